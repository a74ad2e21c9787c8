"""Scene data path on the CPU: the oracle's rays, log conversion, bilinear
projection reads and batch draws pinned by closed forms and independent
formulas; the NPY / NPZ readers against numpy's own writer; the scene
directory parser; argument checks of the C ABI before any device work."""
import ctypes as C
import math
import os

import numpy as np
import pytest
import torch

from scene_dir import tilt_quaternion, write_scene


def _rotmat(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _unit_quats(n, seed):
    q = np.random.default_rng(seed).normal(size=(n, 4))
    return q / np.linalg.norm(q, axis=1, keepdims=True)


def _ulp_close(a, b, ulps=1):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float64)
    tol = ulps * np.spacing(np.maximum(np.abs(b), 1e-30).astype(np.float32)).astype(np.float64)
    return np.all(np.abs(a.astype(np.float64) - b) <= tol + 1e-12)


def test_rotation_matches_matrix_form(oracle):
    for q in _unit_quats(50, 1):
        v = np.random.default_rng(int(abs(q[0]) * 1e6)).normal(size=3)
        np.testing.assert_allclose(oracle.rotate_point(q, v), _rotmat(q) @ v, rtol=0, atol=1e-14)
    # identity, and a quarter turn about +z maps x to y
    v = np.array([0.3, -0.2, 0.7])
    assert np.array_equal(oracle.rotate_point([0, 0, 0, 1.0], v), v)
    s = math.sqrt(0.5)
    np.testing.assert_allclose(oracle.rotate_point([0, 0, s, s], [1.0, 0, 0]), [0, 1, 0], atol=1e-15)
    # q and -q are the same rotation
    q = _unit_quats(1, 2)[0]
    np.testing.assert_allclose(oracle.rotate_point(q, v), oracle.rotate_point(-q, v), atol=1e-15)


def test_orthographic_rays_match_tilt_series(oracle):
    """A parallel-beam tilt series written as a rig (rotation about +y,
    K mapping pixel x to 2 x / (W - 1) - 1) gives the rays of the closed
    form R_y(theta) (p, 0) - sqrt(2) R_y(theta) z."""
    from adaptivetensordf_amd import phantom

    W, H = 64, 48
    angles = phantom.tilt_angles(61)
    tilt, px, py = phantom.sample_pixels(4000, 61, W, H, 3)
    o_ref, d_ref = phantom.rays_for_pixels(tilt, px, py, W, H, angles)
    quats = np.stack([tilt_quaternion(t) for t in angles])
    K = np.array([[(W - 1) / 2, (H - 1) / 2, (W - 1) / 2, (H - 1) / 2, 0.0]])
    uv = np.stack([px / np.float32(W - 1), py / np.float32(H - 1)], 1).astype(np.float32)
    o, d = oracle.generate_rays(1, W, H, K, quats, None, uv, tilt, np.zeros_like(tilt))
    np.testing.assert_allclose(d, d_ref, rtol=0, atol=3e-7)
    np.testing.assert_allclose(o, o_ref, rtol=0, atol=2e-6)


def test_rays_against_matrix_restatement(oracle):
    rng = np.random.default_rng(4)
    W, H, n = 40, 30, 3000
    K = np.array([[35.0, 33.0, 19.25, 14.75, 0.0], [50.0, 52.0, 20.5, 15.5, 0.0]])
    quats = _unit_quats(7, 5)
    trans = rng.uniform(-3, 3, size=(7, 3))
    vol = np.float32([0.1, -0.2, 0.05])
    uv = rng.uniform(-0.1, 1.1, size=(n, 2)).astype(np.float32)
    img = rng.integers(0, 7, n).astype(np.int32)
    cam = rng.integers(0, 2, n).astype(np.int32)
    for align in (True, False):
        if align:
            px = (uv * np.float32([W - 1, H - 1])).astype(np.float32)
        else:
            px = (uv * np.float32([W, H]) - np.float32(0.5)).astype(np.float32)
        u = (px.astype(np.float64) - K[cam, 2:4]) / K[cam, 0:2]
        R = np.stack([_rotmat(q) for q in quats])[img]
        # pinhole
        un = np.concatenate([u, np.ones((n, 1))], 1)
        un /= np.linalg.norm(un, axis=1, keepdims=True)
        o, d = oracle.generate_rays(0, W, H, K, quats, trans, uv, img, cam, vol, align)
        assert _ulp_close(d, np.einsum("nij,nj->ni", R, un), 2)
        assert np.array_equal(o, trans[img].astype(np.float32))
        # orthographic
        p = np.concatenate([u, np.zeros((n, 1))], 1)
        dz = R[:, :, 2]
        want_o = np.einsum("nij,nj->ni", R, p) + vol.astype(np.float64) - math.sqrt(2) * dz
        o, d = oracle.generate_rays(1, W, H, K, quats, None, uv, img, cam, vol, align)
        assert _ulp_close(d, dz, 2)
        err = np.abs(o.astype(np.float64) - want_o) / (1 + np.abs(want_o))
        assert err.max() < 4e-7, err.max()
    # the principal point looks straight down the rotated z axis
    q = quats[:1]
    uvc = np.float32([[K[0, 2] / (W - 1), K[0, 3] / (H - 1)]])
    _, d = oracle.generate_rays(0, W, H, K[:1], q, trans[:1], uvc, [0], [0])
    np.testing.assert_allclose(d[0], _rotmat(q[0])[:, 2], atol=1e-6)


def test_log_conversion_closed_forms(oracle):
    xmin, xmax, pf = 1200.0, 61000.0, 2.5
    raw = np.float32([xmax, xmin, 0.0, -3.0, 1e-7, 1e-5, 30000.0, 5.0])
    out10 = oracle.log_convert(raw, xmin, xmax, pf, log10=True)
    oute = oracle.log_convert(raw, xmin, xmax, pf, log10=False)
    assert out10[0] == 0.0 and oute[0] == 0.0
    np.testing.assert_allclose([out10[1], oute[1]], [pf, pf], rtol=2e-6)
    # clamping at 1e-5
    assert out10[2] == out10[3] == out10[4] == out10[5]
    # the base cancels: -log(x) / -log(m) is base independent
    np.testing.assert_allclose(out10, oute, rtol=3e-6)
    want = -np.log10(np.maximum(raw.astype(np.float64), 1e-5) / xmax) * pf / -np.log10(xmin / xmax)
    np.testing.assert_allclose(out10, want, rtol=3e-6)
    # log-space input: one f32 division
    ls = oracle.log_convert(raw, xmin, xmax, pf, log_space_input=True)
    assert np.array_equal(ls, (raw / np.float32(xmax)).astype(np.float32))
    # the mask multiplies once
    m = np.float32([1, 0, 0.5, 1, 1, 0.25, 0.75, 1])
    assert np.array_equal(oracle.log_convert(raw, xmin, xmax, pf, mask=m), (out10 * m).astype(np.float32))


def test_bilinear_read_closed_forms(oracle):
    w, h = 17, 9  # w - 1 and h - 1 powers of two: pixel uv map exactly to pixels
    rng = np.random.default_rng(6)
    imgs = rng.normal(size=(3, h, w)).astype(np.float32)
    ys, xs = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    uv = np.stack([xs.ravel() / np.float32(16), ys.ravel() / np.float32(8)], 1).astype(np.float32)
    for k in range(3):
        t, tm = oracle.sample_projection(imgs, uv, np.full(uv.shape[0], k))
        assert np.array_equal(t, imgs[k].ravel()) and np.all(tm == 1.0)
    # a bilinear function is reproduced between pixels (align corners and not)
    a, b, c, d = 0.3, -0.7, 0.45, 0.11
    f = lambda x, y: a + b * x + c * y + d * x * y  # noqa: E731
    lin = f(xs, ys).astype(np.float32)[None]
    uvr = rng.uniform(0, 1, size=(2000, 2)).astype(np.float32)
    t, _ = oracle.sample_projection(lin, uvr, np.zeros(2000))
    np.testing.assert_allclose(t, f(uvr[:, 0] * 16.0, uvr[:, 1] * 8.0), atol=2e-5)
    t, _ = oracle.sample_projection(lin, uvr, np.zeros(2000), align_corners=False)
    xq = np.clip(uvr[:, 0] * 17.0 - 0.5, 0, 16)
    yq = np.clip(uvr[:, 1] * 9.0 - 0.5, 0, 8)
    np.testing.assert_allclose(t, f(xq, yq), atol=2e-5)
    # border padding: outside [0, 1] reads the edge
    out = np.float32([[-0.4, 0.5], [1.6, 0.5], [0.5, -2.0], [0.5, 3.0], [-1, -1], [2, 2]])
    edge = np.float32([[0.0, 0.5], [1.0, 0.5], [0.5, 0.0], [0.5, 1.0], [0, 0], [1, 1]])
    t1, _ = oracle.sample_projection(imgs, out, np.ones(6))
    t2, _ = oracle.sample_projection(imgs, edge, np.ones(6))
    assert np.array_equal(t1, t2)
    assert t2[4] == imgs[1, 0, 0] and t2[5] == imgs[1, h - 1, w - 1]
    # a mask is read the same way
    m = rng.uniform(0, 1, size=(3, h, w)).astype(np.float32)
    t, tm = oracle.sample_projection(imgs, uv, np.full(uv.shape[0], 2), mask=m)
    assert np.array_equal(tm, m[2].ravel())


def test_batch_draws(oracle):
    idx = np.int32([4, 9, 2])
    cams = np.int32([1, 0, 1])
    uv, img, cam = oracle.random_pixels(idx, cams, 12, 7, 20000, 11, interpolate=False)
    x = uv[:, 0] * np.float32(11)
    y = uv[:, 1] * np.float32(6)
    assert np.all(np.abs(x - np.round(x)) < 1e-5) and np.all(np.abs(y - np.round(y)) < 1e-5)
    assert set(np.round(x).astype(int)) == set(range(12)) and set(np.round(y).astype(int)) == set(range(7))
    assert set(img) == {4, 9, 2}
    assert np.all(cam == np.where(img == 9, 0, 1))
    counts = np.array([(img == i).sum() for i in idx])
    assert np.all(np.abs(counts - 20000 / 3) < 5 * math.sqrt(20000 / 3))
    uv2, img2, _ = oracle.random_pixels(idx, cams, 12, 7, 20000, 11, interpolate=False)
    assert np.array_equal(uv, uv2) and np.array_equal(img, img2)
    uvi, _, _ = oracle.random_pixels(idx, None, 12, 7, 20000, 12, interpolate=True)
    assert uvi.min() >= 0 and uvi.max() < 1 and abs(uvi.mean() - 0.5) < 0.01
    uvc, _, _ = oracle.random_pixels(idx, None, 12, 7, 100, 11, interpolate=False, align_corners=False)
    x = uvc[:, 0] * np.float32(12) - np.float32(0.5)
    assert np.all(np.abs(x - np.round(x)) < 1e-5)


# ---------------------------------------------------------------------------
# NPY / NPZ readers against numpy's writer
# ---------------------------------------------------------------------------
_DTYPES = ["<f4", "<f8", ">f4", ">f8", "|u1", "<u2", ">u2", "<i2", "<i4", ">i4", "<i8", "|i1", "|b1", "<u4"]
_TORCH = {"f4": torch.float32, "f8": torch.float64, "u1": torch.uint8, "u2": torch.uint16, "i2": torch.int16,
          "i4": torch.int32, "i8": torch.int64, "i1": torch.int8, "b1": torch.bool, "u4": torch.uint32}


@pytest.mark.parametrize("dt", _DTYPES)
def test_npy_reader_dtypes(native, tmp_path, dt):
    rng = np.random.default_rng(7)
    for shape in ((5, 7), (3,), (2, 3, 4), (0,), ()):
        a = (rng.normal(size=shape) * 100).astype(dt) if dt[1] != "b" else rng.uniform(size=shape) > 0.5
        a = np.asarray(a, dtype=dt)
        p = str(tmp_path / "a.npy")
        np.save(p, a)
        t = native.LoadNumpy(p)
        assert t.dtype == _TORCH[dt[1:]] and tuple(t.shape) == shape
        assert np.array_equal(t.to(torch.float64).numpy(), a.astype(np.float64))


def test_npz_reader(native, tmp_path):
    rng = np.random.default_rng(8)
    proj = rng.normal(size=(33, 47)).astype(np.float32)
    other = rng.normal(size=(5,))
    for saver in (np.savez, np.savez_compressed):
        p = str(tmp_path / f"{saver.__name__}.npz")
        saver(p, other=other, proj=proj, proj2=proj.astype(np.float64))
        t = native.LoadNumpyNpz(p, "proj")
        assert t.dtype == torch.float32 and np.array_equal(t.numpy(), proj)
        assert np.array_equal(native.LoadNumpyNpz(p, "proj2").numpy(), proj.astype(np.float64))
        assert np.array_equal(native.LoadNumpyNpz(p, "other").numpy(), other)
        with pytest.raises(RuntimeError, match="no array 'missing'"):
            native.LoadNumpyNpz(p, "missing")
    big = rng.integers(0, 65535, size=(512, 700)).astype(np.uint16)
    p = str(tmp_path / "big.npz")
    np.savez_compressed(p, proj=big)
    assert np.array_equal(native.LoadNumpyNpz(p, "proj").numpy(), big)


def test_npy_reader_rejects(native, tmp_path):
    p = str(tmp_path / "f.npy")
    np.save(p, np.asfortranarray(np.arange(6, dtype=np.float32).reshape(2, 3)))
    with pytest.raises(RuntimeError, match="Fortran"):
        native.LoadNumpy(p)
    np.save(p, np.arange(6, dtype=np.complex64))
    with pytest.raises(RuntimeError, match="unsupported dtype"):
        native.LoadNumpy(p)
    np.save(p, np.arange(100, dtype=np.float32))
    data = open(p, "rb").read()
    open(p, "wb").write(data[:-8])
    with pytest.raises(RuntimeError, match="shorter"):
        native.LoadNumpy(p)
    open(p, "wb").write(b"not numpy at all")
    with pytest.raises(RuntimeError, match="not an NPY"):
        native.LoadNumpy(p)


def _npy_with_shape(shape_text):
    """A little-endian f4 NPY whose header states `shape_text` and holds
    16 bytes of data."""
    hdr = "{'descr': '<f4', 'fortran_order': False, 'shape': %s, }" % shape_text
    hdr += " " * ((64 - (10 + len(hdr) + 1) % 64) % 64) + "\n"
    return b"\x93NUMPY\x01\x00" + len(hdr).to_bytes(2, "little") + hdr.encode() + bytes(16)


def test_npy_reader_rejects_bad_shapes(native, tmp_path):
    p = str(tmp_path / "s.npy")
    open(p, "wb").write(_npy_with_shape("(2, 2)"))
    assert tuple(native.LoadNumpy(p).shape) == (2, 2)
    for bad, msg in (("(-1, 4)", "negative dimension"), ("(3037000500, 3037000500)", "shorter"),
                     ("(4611686018427387904, 4)", "shorter"), ("(2, 3)", "shorter")):
        open(p, "wb").write(_npy_with_shape(bad))
        with pytest.raises(RuntimeError, match=msg):
            native.LoadNumpy(p)


def test_npz_reader_rejects_corrupt_directory(native, tmp_path):
    p = str(tmp_path / "c.npz")
    np.savez(p, proj=np.arange(12, dtype=np.float32))
    good = bytearray(open(p, "rb").read())
    cd = good.rfind(b"PK\x01\x02")
    assert cd > 0
    # name length running past the end of the archive
    z = bytearray(good)
    z[cd + 28:cd + 30] = (0xFFFF).to_bytes(2, "little")
    open(p, "wb").write(bytes(z))
    with pytest.raises(RuntimeError):
        native.LoadNumpyNpz(p, "proj")
    # extra field whose length runs past the entry's extra area
    z = bytearray(good)
    name_len = int.from_bytes(z[cd + 28:cd + 30], "little")
    extra_len = int.from_bytes(z[cd + 30:cd + 32], "little")
    if extra_len >= 4:
        x = cd + 46 + name_len
        z[x + 2:x + 4] = (extra_len + 100).to_bytes(2, "little")
        open(p, "wb").write(bytes(z))
        with pytest.raises(RuntimeError, match="extra field"):
            native.LoadNumpyNpz(p, "proj")
    # local header offset outside the archive
    z = bytearray(good)
    z[cd + 42:cd + 46] = (len(z) + 10).to_bytes(4, "little")
    open(p, "wb").write(bytes(z))
    with pytest.raises(RuntimeError, match="local header"):
        native.LoadNumpyNpz(p, "proj")
    # truncated archive
    open(p, "wb").write(bytes(good[: len(good) // 2]))
    with pytest.raises(RuntimeError):
        native.LoadNumpyNpz(p, "proj")


# ---------------------------------------------------------------------------
# scene directory
# ---------------------------------------------------------------------------
def test_scene_directory_parse(native, tmp_path):
    root, raw, poses = write_scene(str(tmp_path / "scene_a"), num_images=7, scale_poses=3.0, scene_scale=0.5,
                                   projection_factor=1.75)
    s = native.SceneBase(root)
    dp = s.dataset_params
    assert s.scene_name == "scene_a" and s.NumImages() == 7
    assert dp.camera_model == "orthographic" and s.CameraModelId() == 1
    assert dp.projection_factor == 1.75 and dp.scene_scale == 0.5 and dp.xray_min == 1000.0
    assert dp.xray_max == 60000.0 and dp.use_log10_conversion and not dp.log_space_input
    assert dp.image_dir == "projections" and dp.mask_dir == ""
    assert list(dp.roi_min) == [-1, -1, -0.5] and list(dp.roi_max) == [1, 1, 0.5]
    cam = s.cameras[0]
    assert (cam.w, cam.h) == (24, 16) and list(cam.K) == [11.5, 7.5, 11.5, 7.5, 0.0]
    for i, f in enumerate(s.frames):
        q = poses[i, :4]
        np.testing.assert_allclose(f.rotation, q / np.linalg.norm(q), rtol=0, atol=1e-16)
        assert abs(np.linalg.norm(f.rotation) - 1) < 1e-15
        assert np.array_equal(f.translation, poses[i, 4:] * 0.5)
        assert f.image_file == f"proj_{i:03d}." + ("npy" if i % 2 == 0 else "npz")
        assert f.camera_id == 0 and not f.loaded
    s.LoadSplit("exp_uniform_50")
    assert s.train_indices == [0, 2, 3, 5, 6] and s.test_indices == [1, 4]
    assert native.SceneBase.ReadIndexFile(os.path.join(root, "exp_uniform_50", "train.txt")) == [6, 5, 3, 2, 0]
    s.Finalize("cpu", [0.25, -0.5, 0.125])
    assert np.array_equal(s.translation.numpy(), poses[:, 4:] * 0.5 - np.array([0.25, -0.5, 0.125]))
    assert np.array_equal(s.rotation.numpy(), np.array([f.rotation for f in s.frames]))
    assert np.array_equal(s.intrinsics.numpy(), [[11.5, 7.5, 11.5, 7.5, 0.0]])
    with pytest.raises(RuntimeError, match="Finalize"):
        s.LoadImagesCT([0])


def test_scene_directory_errors(native, tmp_path):
    with pytest.raises(RuntimeError, match="does not exist"):
        native.SceneBase(str(tmp_path / "nowhere"))
    root, _, _ = write_scene(str(tmp_path / "s"), num_images=3)
    txt = open(os.path.join(root, "dataset.ini")).read()
    open(os.path.join(root, "dataset.ini"), "w").write(txt.replace("orthographic", "fisheye", 1))
    with pytest.raises(RuntimeError, match="camera_model"):
        native.SceneBase(root)
    open(os.path.join(root, "dataset.ini"), "w").write(txt)
    open(os.path.join(root, "camera_indices.txt"), "w").write("0\n0\n")
    with pytest.raises(RuntimeError, match="camera_indices"):
        native.SceneBase(root)
    open(os.path.join(root, "camera_indices.txt"), "w").write("0\n1\n0\n")
    with pytest.raises(RuntimeError, match="camera 1"):
        native.SceneBase(root)
    os.remove(os.path.join(root, "camera_indices.txt"))
    s = native.SceneBase(root)
    with pytest.raises(RuntimeError, match="does not exist"):
        s.LoadSplit("no_split")


def test_scene_boundary_without_device_work():
    from adaptivetensordf_amd import capi

    lib = capi.lib()
    one = C.c_void_p(1)
    rig = capi.CameraRig(2, 8, 8, 1, 1, 1, one, one, one, (C.c_float * 3)(0, 0, 0))
    st = lib.atdf_generate_rays(C.byref(rig), one, one, one, 4, one, one, None)
    assert st == capi.ERR_INVALID_ARGUMENT and b"camera_model" in lib.atdf_last_error()
    rig = capi.CameraRig(0, 8, 8, 1, 1, 1, one, one, None, (C.c_float * 3)(0, 0, 0))
    st = lib.atdf_generate_rays(C.byref(rig), one, one, one, 4, one, one, None)
    assert st == capi.ERR_INVALID_ARGUMENT and b"null rig" in lib.atdf_last_error()
    rig = capi.CameraRig(1, 8, 8, 1, 1, 1, one, one, None, (C.c_float * 3)(0, 0, 0))
    assert lib.atdf_generate_rays(C.byref(rig), None, None, None, 0, None, None, None) == capi.OK
    st = lib.atdf_log_convert(one, None, 4, 0.0, 100.0, 1.0, 1, 0, one, None)
    assert st == capi.ERR_INVALID_ARGUMENT and b"xray_min" in lib.atdf_last_error()
    assert lib.atdf_log_convert(one, None, 0, 0.0, 100.0, 1.0, 1, 1, one, None) == capi.OK
    ps = capi.ProjectionSet(None, None, 8, 8, 1, 1)
    st = lib.atdf_sample_projection(C.byref(ps), one, one, 4, one, None, None)
    assert st == capi.ERR_INVALID_ARGUMENT and b"null projection" in lib.atdf_last_error()
    st = lib.atdf_random_ray_batch(None, None, one, 0, None, 8, 8, 1, 1, 0, 4, one, one, one, None, None, None,
                                   None, None)
    assert st == capi.ERR_INVALID_ARGUMENT and b"bad sizes" in lib.atdf_last_error()
    st = lib.atdf_random_ray_batch(None, None, one, 2, None, 8, 8, 1, 1, 0, 4, one, one, one, one, None, None,
                                   None, None)
    assert st == capi.ERR_INVALID_ARGUMENT and b"projection set" in lib.atdf_last_error()


def test_point_in_any_image_closed_forms(oracle):
    """Visibility against the projection written with the rotation matrix:
    R^T (p - t) mapped through K and rounded half away from zero."""
    rng = np.random.default_rng(12)
    w, h = 30, 20
    roi_min, roi_max = np.float32([-1, -1, -1]), np.float32([1, 1, 1])
    q = _unit_quats(1, 13)
    R = _rotmat(q[0])
    # pinhole: camera 4 units back along its own z axis
    t = -4.0 * R[:, 2]
    K = np.array([[60.0, 58.0, 14.5, 9.5, 0.7]])  # skew ignored
    # pixel targets away from rounding ties, some inside, some outside
    px = np.concatenate([rng.uniform(-3, w + 3, 400), [-0.6, -0.4, w - 0.6, w - 0.4]])
    py = np.concatenate([rng.uniform(-3, h + 3, 400), [5.2, 5.2, 5.2, 5.2]])
    z = rng.uniform(3.5, 4.5, px.size)
    view = np.stack([(px - K[0, 2]) / K[0, 0] * z, (py - K[0, 3]) / K[0, 1] * z, z], 1)
    pts = (view @ R.T + t).astype(np.float32)
    inside_roi = np.all(np.abs(pts) <= 1, 1)
    vis = oracle.point_in_any_image(0, w, h, K, q, t[None], roi_min, roi_max, pts)
    # recompute the pixel from the float32 points, then round half away from zero
    v = (pts.astype(np.float64) - t) @ R
    ix, iy = K[0, 0] * v[:, 0] / v[:, 2] + K[0, 2], K[0, 1] * v[:, 1] / v[:, 2] + K[0, 3]
    rx, ry = np.sign(ix) * np.floor(np.abs(ix) + 0.5), np.sign(iy) * np.floor(np.abs(iy) + 0.5)
    want = inside_roi & (rx >= 0) & (rx < w) & (ry >= 0) & (ry < h)
    assert np.array_equal(vis, want.astype(np.float32))
    assert 0.15 < want.mean() < 0.85
    assert list(vis[-4:]) == [0.0, 1.0, 1.0, 0.0] or not inside_roi[-4:].all()
    # orthographic: the same with (fx x + cx, fy y + cy); "any" over images
    Ko = np.array([[8.0, 8.0, 14.5, 9.5, 0.0]])
    q2 = np.stack([tilt_quaternion(0.0), tilt_quaternion(math.pi / 2)])
    t2 = np.zeros((2, 3))
    p = np.float32([[1.5, 0.0, 0.0], [0.0, 0.0, 0.5], [0.9, 0.0, 0.9], [0.1, -0.2, 0.3], [0.2, 1.3, 0.0]])
    # image 0 sees -1.875 < x < 1.875 (pixels 0..29 after rounding), image 1
    # (90 degrees about y) sees z in the same range
    vis = oracle.point_in_any_image(1, w, h, Ko, q2, t2, np.float32([-2] * 3), np.float32([2] * 3), p)
    assert list(vis) == [1.0, 1.0, 1.0, 1.0, 0.0]
    vis = oracle.point_in_any_image(1, w, h, Ko, q2[:1], t2[:1], np.float32([-2] * 3), np.float32([2] * 3),
                                    np.float32([[1.9, 0, 0], [-1.9, 0, 0], [1.8, 0, 0]]))
    assert list(vis) == [0.0, 0.0, 1.0]
    # the ROI box clips first
    vis = oracle.point_in_any_image(1, w, h, Ko, q2, t2, np.float32([-0.5] * 3), np.float32([0.5] * 3), p)
    assert list(vis) == [0.0, 1.0, 0.0, 1.0, 0.0]


def test_node_lattice_points_closed_form(oracle):
    t = oracle.tree_build(2)
    ids = np.int64([0, 1, 12])
    pos = oracle.node_lattice_points(t["node_min"], t["node_max"], ids, 3)
    for a, node in enumerate(ids):
        mn, mx = t["node_min"][node].astype(np.float64), t["node_max"][node].astype(np.float64)
        for p in range(27):
            idx = np.array([p % 3, (p // 3) % 3, p // 9])
            assert np.array_equal(pos[a, p], (mn + idx / 2.0 * (mx - mn)).astype(np.float32))
