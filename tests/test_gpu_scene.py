"""Scene data path on the GPU, through the C ABI: rays bit-exact with the
oracle for both camera models and uv conventions, the log conversion within
a few ulp (the transcendental is the device library's), bilinear projection
reads bit-exact, the fused random batch equal to the composition of its
parts, and a scene directory loaded and sampled end to end."""
import ctypes as C
import math

import numpy as np
import pytest
import torch

from scene_dir import tilt_quaternion, write_scene

pytestmark = pytest.mark.gpu


def _rig(capi, model, w, h, K, rot, trans, vol=(0, 0, 0), align=True):
    return capi.CameraRig(model, w, h, int(align), K.shape[0], rot.shape[0], capi.ptr(K), capi.ptr(rot),
                          capi.ptr(trans), (C.c_float * 3)(*vol))


def _rig_case(cuda, seed, n=5000, cams=2, images=9, w=96, h=64):
    rng = np.random.default_rng(seed)
    K = np.stack([[80.0 + 7 * c, 75.0 - 3 * c, 47.3 + c, 31.6 - c, 0.0] for c in range(cams)])
    q = rng.normal(size=(images, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    t = rng.uniform(-3, 3, size=(images, 3))
    uv = rng.uniform(-0.05, 1.05, size=(n, 2)).astype(np.float32)
    img = rng.integers(0, images, n).astype(np.int32)
    cam = rng.integers(0, cams, n).astype(np.int32)
    dev = {k: torch.from_numpy(np.ascontiguousarray(v)).to(cuda) for k, v in
           dict(K=K, q=q, t=t, uv=uv, img=img, cam=cam).items()}
    return (K, q, t, uv, img, cam), dev


@pytest.mark.parametrize("model", [0, 1])
@pytest.mark.parametrize("align", [True, False])
def test_generate_rays_bit_exact(oracle, cuda, model, align):
    from adaptivetensordf_amd import capi

    (K, q, t, uv, img, cam), d = _rig_case(cuda, 10 + model + 2 * align)
    vol = (0.125, -0.3, 0.07)
    rig = _rig(capi, model, 96, 64, d["K"], d["q"], d["t"], vol, align)
    n = uv.shape[0]
    o = torch.empty(n, 3, device=cuda)
    dr = torch.empty(n, 3, device=cuda)
    print(f"generate_rays model {model} align {align} n {n}", flush=True)
    capi.call("atdf_generate_rays", C.byref(rig), capi.ptr(d["uv"]), capi.ptr(d["img"]), capi.ptr(d["cam"]), n,
              capi.ptr(o), capi.ptr(dr), capi.stream_ptr())
    torch.cuda.synchronize()
    oo, od = oracle.generate_rays(model, 96, 64, K, q, t if model == 0 else None, uv, img, cam, vol, align)
    assert np.array_equal(o.cpu().numpy(), oo) and np.array_equal(dr.cpu().numpy(), od)


def test_generate_rays_bad_index_gives_nan(cuda):
    from adaptivetensordf_amd import capi

    _, d = _rig_case(cuda, 3, n=64)
    d["img"][5] = 9
    d["cam"][7] = -1
    rig = _rig(capi, 0, 96, 64, d["K"], d["q"], d["t"])
    o = torch.empty(64, 3, device=cuda)
    dr = torch.empty(64, 3, device=cuda)
    capi.call("atdf_generate_rays", C.byref(rig), capi.ptr(d["uv"]), capi.ptr(d["img"]), capi.ptr(d["cam"]), 64,
              capi.ptr(o), capi.ptr(dr), capi.stream_ptr())
    bad = torch.isnan(dr).any(1).cpu().numpy()
    assert bad[5] and bad[7] and bad.sum() == 2


def test_orthographic_tilt_series_full_batch(cuda):
    """2^18 detector rays of a 61-tilt parallel-beam series against the
    closed-form tilt geometry (size-independent: every ray within f32
    rounding of R_y(theta) (p, 0) - sqrt(2) R_y(theta) z)."""
    from adaptivetensordf_amd import capi, phantom

    W = H = 512
    angles = phantom.tilt_angles(61)
    tilt, px, py = phantom.sample_pixels(1 << 18, 61, W, H, 5)
    o_ref, d_ref = phantom.rays_for_pixels(tilt, px, py, W, H, angles)
    q = torch.from_numpy(np.stack([tilt_quaternion(a) for a in angles])).to(cuda)
    K = torch.tensor([[(W - 1) / 2, (H - 1) / 2, (W - 1) / 2, (H - 1) / 2, 0.0]], dtype=torch.float64, device=cuda)
    uv = torch.from_numpy(np.stack([px / np.float32(W - 1), py / np.float32(H - 1)], 1).astype(np.float32)).to(cuda)
    img = torch.from_numpy(tilt.astype(np.int32)).to(cuda)
    cam = torch.zeros_like(img)
    rig = _rig(capi, 1, W, H, K, q, None)
    n = uv.shape[0]
    o = torch.empty(n, 3, device=cuda)
    d = torch.empty(n, 3, device=cuda)
    capi.call("atdf_generate_rays", C.byref(rig), capi.ptr(uv), capi.ptr(img), capi.ptr(cam), n, capi.ptr(o),
              capi.ptr(d), capi.stream_ptr())
    np.testing.assert_allclose(d.cpu().numpy(), d_ref, rtol=0, atol=3e-7)
    np.testing.assert_allclose(o.cpu().numpy(), o_ref, rtol=0, atol=2e-6)


@pytest.mark.parametrize("mode", ["log10", "ln", "log_space"])
def test_log_convert_against_oracle(oracle, cuda, mode):
    from adaptivetensordf_amd import capi

    rng = np.random.default_rng(21)
    raw = rng.uniform(-10, 70000, size=200003).astype(np.float32)
    raw[:50] = np.float32([0, 1e-7, 1e-5, 60000, 1300]) .repeat(10)
    mask = rng.uniform(0, 1, size=raw.size).astype(np.float32)
    log10, ls = mode == "log10", mode == "log_space"
    for m in (None, mask):
        r = torch.from_numpy(raw).to(cuda)
        md = None if m is None else torch.from_numpy(m).to(cuda)
        out = torch.empty_like(r)
        capi.call("atdf_log_convert", capi.ptr(r), capi.ptr(md), raw.size, 1300.0, 60000.0, 1.5, int(log10), int(ls),
                  capi.ptr(out), capi.stream_ptr())
        want = oracle.log_convert(raw, 1300.0, 60000.0, 1.5, log10=log10, log_space_input=ls, mask=m)
        got = out.cpu().numpy()
        if ls:
            assert np.array_equal(got, want)
        else:
            np.testing.assert_allclose(got, want, rtol=4e-7, atol=1e-7)


def test_sample_projection_bit_exact(oracle, cuda):
    from adaptivetensordf_amd import capi

    rng = np.random.default_rng(22)
    n_img, h, w, n = 5, 37, 53, 40000
    proj = rng.normal(size=(n_img, h, w)).astype(np.float32)
    mask = rng.uniform(size=(n_img, h, w)).astype(np.float32)
    uv = rng.uniform(-0.1, 1.1, size=(n, 2)).astype(np.float32)
    uv[:100] = np.float32([[0, 0], [1, 1], [1, 0], [0, 1], [0.5, 1]]).repeat(20, 0)
    img = rng.integers(0, n_img, n).astype(np.int32)
    P, M, U, I = (torch.from_numpy(a).to(cuda) for a in (proj, mask, uv, img))
    for align in (True, False):
        for use_mask in (False, True):
            ps = capi.ProjectionSet(P.data_ptr(), M.data_ptr() if use_mask else None, w, h, n_img, int(align))
            t = torch.empty(n, device=cuda)
            tm = torch.empty(n, device=cuda)
            capi.call("atdf_sample_projection", C.byref(ps), capi.ptr(U), capi.ptr(I), n, capi.ptr(t), capi.ptr(tm),
                      capi.stream_ptr())
            wt, wm = oracle.sample_projection(proj, uv, img, mask=mask if use_mask else None, align_corners=align)
            assert np.array_equal(t.cpu().numpy(), wt) and np.array_equal(tm.cpu().numpy(), wm)


@pytest.mark.parametrize("interpolate", [True, False])
def test_random_batch_is_its_composition(oracle, cuda, interpolate):
    from adaptivetensordf_amd import capi

    (K, q, t, _, _, _), d = _rig_case(cuda, 30, n=8, cams=2, images=9, w=96, h=64)
    rng = np.random.default_rng(31)
    proj = rng.normal(size=(9, 64, 96)).astype(np.float32)
    P = torch.from_numpy(proj).to(cuda)
    idx = np.int32([8, 1, 4, 6])
    icam = np.int32([1, 0, 1, 1])
    Idx, Icam = torch.from_numpy(idx).to(cuda), torch.from_numpy(icam).to(cuda)
    n = 100000
    rig = _rig(capi, 0, 96, 64, d["K"], d["q"], d["t"])
    ps = capi.ProjectionSet(P.data_ptr(), None, 96, 64, 9, 1)
    uv = torch.empty(n, 2, device=cuda)
    im = torch.empty(n, dtype=torch.int32, device=cuda)
    cm = torch.empty(n, dtype=torch.int32, device=cuda)
    tg = torch.empty(n, device=cuda)
    tm = torch.empty(n, device=cuda)
    o = torch.empty(n, 3, device=cuda)
    dr = torch.empty(n, 3, device=cuda)
    capi.call("atdf_random_ray_batch", C.byref(rig), C.byref(ps), capi.ptr(Idx), 4, capi.ptr(Icam), 96, 64, 1,
              int(interpolate), 99, n, capi.ptr(uv), capi.ptr(im), capi.ptr(cm), capi.ptr(tg), capi.ptr(tm),
              capi.ptr(o), capi.ptr(dr), capi.stream_ptr())
    wuv, wim, wcm = oracle.random_pixels(idx, icam, 96, 64, n, 99, interpolate=interpolate)
    assert np.array_equal(uv.cpu().numpy(), wuv)
    assert np.array_equal(im.cpu().numpy(), wim) and np.array_equal(cm.cpu().numpy(), wcm)
    wt, _ = oracle.sample_projection(proj, wuv, wim)
    assert np.array_equal(tg.cpu().numpy(), wt) and torch.all(tm == 1)
    wo, wd = oracle.generate_rays(0, 96, 64, K, q, t, wuv, wim, wcm)
    assert np.array_equal(o.cpu().numpy(), wo) and np.array_equal(dr.cpu().numpy(), wd)


def test_scene_directory_end_to_end(oracle, native, cuda, tmp_path):
    w, h, n_img = 40, 24, 6
    rng = np.random.default_rng(40)
    masks = (rng.uniform(size=(n_img, h, w)) > 0.2).astype(np.uint8) * 255
    raw = rng.integers(0, 65535, size=(n_img, h, w)).astype(np.uint16)
    root, _, poses = write_scene(str(tmp_path / "scene"), camera_model="pinhole", w=w, h=h, num_images=n_img,
                                 raw=raw, masks=masks, scene_scale=1.5, xray_min=900.0, xray_max=64000.0,
                                 log10=False, projection_factor=3.0)
    s = native.SceneBase(root)
    s.LoadSplit("exp_uniform_50")
    vol = [0.05, -0.1, 0.2]
    s.Finalize("cuda:0", vol)
    s.LoadImagesCT(s.train_indices)
    print("scene loaded", flush=True)
    for i in range(n_img):
        assert s.frames[i].loaded == (i in s.train_indices)
    P = s.projections.cpu().numpy()
    for i in s.train_indices:
        want = oracle.log_convert(raw[i].astype(np.float32), 900.0, 64000.0, 3.0, log10=False,
                                  mask=masks[i].astype(np.float32) / np.float32(255))
        np.testing.assert_allclose(P[i], want, rtol=4e-7, atol=1e-7)
    for i in s.test_indices:
        assert np.all(P[i] == 0)
    # rays of the scene's rig
    uv = torch.rand(3000, 2, device=cuda)
    img = torch.tensor(np.resize(s.train_indices, 3000), dtype=torch.int32, device=cuda)
    cam = torch.zeros_like(img)
    rays = s.GetRays(uv, img, cam)
    trans = poses[:, 4:] * 1.5 - np.array(vol, dtype=np.float32).astype(np.float64)
    rot = np.array([f.rotation for f in s.frames])
    K = np.array([s.cameras[0].K])
    wo, wd = oracle.generate_rays(0, w, h, K, rot, trans, uv.cpu().numpy(), img.cpu().numpy(), cam.cpu().numpy())
    assert np.array_equal(rays.origin.cpu().numpy(), wo) and np.array_equal(rays.direction.cpu().numpy(), wd)
    t, m = s.SampleProjection(uv, img)
    wt, wm = oracle.sample_projection(P, uv.cpu().numpy(), img.cpu().numpy(), mask=s.masks.cpu().numpy())
    assert np.array_equal(t.cpu().numpy(), wt) and np.array_equal(m.cpu().numpy(), wm)
    with pytest.raises(RuntimeError, match="out of range"):
        s.GetRays(uv[:2], torch.tensor([0, 6], dtype=torch.int32, device=cuda), cam[:2])
    # a random batch over the training images
    b = s.RandomRayBatch(torch.tensor(s.train_indices, dtype=torch.int32), 50000, 5, False)
    assert set(b.image_id.cpu().numpy().tolist()) == set(s.train_indices)
    r2 = s.GetRays(b.uv, b.image_id, b.camera_id)
    assert torch.equal(r2.origin, b.rays.origin) and torch.equal(r2.direction, b.rays.direction)
    t2, m2 = s.SampleProjection(b.uv, b.image_id)
    assert torch.equal(t2, b.target) and torch.equal(m2, b.target_mask)
    assert math.isclose(float(b.uv[:, 0].mul(w - 1).round().max()), w - 1)


def _visibility_scene(tmp_path, model):
    # a narrow detector (visible |x| <~ 0.5 in view space) and the ROI's
    # z limit of +-0.5 leave part of the cube unseen
    w, h = 24, 40
    K = (w - 1.0, (h - 1) / 2.0, (w - 1) / 2.0, (h - 1) / 2.0, 0.3) if model == "orthographic" else \
        (20.0, 20.0, (w - 1) / 2.0, (h - 1) / 2.0, 0.3)
    root, _, _ = write_scene(str(tmp_path / f"vis_{model}"), camera_model=model, w=w, h=h, num_images=3, K=K)
    return root


@pytest.mark.parametrize("model", ["orthographic", "pinhole"])
def test_point_in_any_image_bit_exact(oracle, native, cuda, tmp_path, model):
    from scene_util import adaptive_active, device_tree

    s = native.SceneBase(_visibility_scene(tmp_path, model))
    s.Finalize("cuda:0", [0.1, 0.0, -0.05])
    t = oracle.tree_build(3)
    act = adaptive_active(t, budget=200, seed=3)
    tree = device_tree(3, act)
    ids = oracle.active_ids(act)
    pos = tree.UniformGlobalSamples(torch.from_numpy(ids).to(cuda), 7)
    want_pos = oracle.node_lattice_points(t["node_min"], t["node_max"], ids, 7)
    assert tuple(pos.shape) == (ids.size, 7, 7, 7, 3)
    assert np.array_equal(pos.reshape(ids.size, -1, 3).cpu().numpy(), want_pos)
    mask = s.PointInAnyImage(pos)
    assert tuple(mask.shape) == (ids.size, 7, 7, 7)
    dp = s.dataset_params
    want = oracle.point_in_any_image(s.CameraModelId(), 24, 40, s.intrinsics.cpu().numpy(), s.rotation.cpu().numpy(),
                                     s.translation.cpu().numpy(), dp.roi_min, dp.roi_max, want_pos.reshape(-1, 3))
    got = mask.reshape(-1).cpu().numpy()
    print(f"{model}: visible {got.mean():.3f}", flush=True)
    assert np.array_equal(got, want)
    assert 0.05 < got.mean() < 0.95


def test_compute_max_density_with_visibility(oracle, native, cuda, tmp_path):
    from test_gpu_project import make_case

    cs = make_case(oracle, C_=16, seed=41, budget=200)
    s = native.SceneBase(_visibility_scene(tmp_path, "orthographic"))
    s.Finalize("cuda:0", [0.0, 0.0, 0.0])
    grid, mlp, tree = cs["grid"].cuda(), cs["mlp"].cuda(), cs["tree"]
    n = 9
    native.ComputeMaxDensity(tree, grid, mlp, cs["fp"], n, s)
    ids = oracle.active_ids(cs["act"])
    gmax = tree.node_max_density.cpu().numpy()[ids]
    fm = oracle.FusedModel(cs["t"]["node_min"], cs["t"]["node_max"], cs["act"], cs["grid"].numpy(),
                           cs["mlp"].numpy(), 16)
    _, _, _, dens = oracle.node_density_stats(fm, 0, n, want_densities=True)
    pts = oracle.node_lattice_points(cs["t"]["node_min"], cs["t"]["node_max"], ids, n).reshape(-1, 3)
    dp = s.dataset_params
    vis = oracle.point_in_any_image(1, 24, 40, s.intrinsics.cpu().numpy(), s.rotation.cpu().numpy(),
                                    s.translation.cpu().numpy(), dp.roi_min, dp.roi_max, pts)
    omax = (dens.reshape(ids.size, -1) * vis.reshape(ids.size, -1)).max(1)
    unseen = vis.reshape(ids.size, -1).max(1) == 0
    print(f"nodes {ids.size}, unseen {unseen.sum()}, partly seen "
          f"{((vis.reshape(ids.size, -1).min(1) == 0) & ~unseen).sum()}", flush=True)
    assert unseen.any() and not unseen.all()
    assert np.all(gmax[unseen] == 0)
    assert np.abs(gmax - omax).max() <= 1e-5 * np.abs(omax).max()
    # without a scene the same call takes every lattice point
    native.ComputeMaxDensity(tree, grid, mlp, cs["fp"], n)
    full = tree.node_max_density.cpu().numpy()[ids]
    assert np.all(full >= gmax) and np.any(full > gmax)
