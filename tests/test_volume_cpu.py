"""Volume export on the CPU: the lattice-point oracle against closed forms,
and the Analyze 7.5 writer / reader (SaveHDRImageTensor, LoadHDRImageTensor)
checked field by field against the format's published header layout."""
import os
import struct

import numpy as np
import torch


def test_volume_points_closed_forms(oracle):
    # mode 0 with size - 1 a power of two: i / 8 * 2 - 1 is exact
    pos = oracle.volume_points(0, (9, 5, 3))
    ii = np.arange(9 * 5 * 3)
    c = np.stack([ii % 9, (ii // 9) % 5, ii // 45], 1)
    want = c / (np.array([9, 5, 3]) - 1) * 2 - 1
    assert np.array_equal(pos.astype(np.float64), want)
    # swap_xy: axis 2 fastest
    pos = oracle.volume_points(0, (3, 5, 9), swap_xy=True)
    c = np.stack([ii // 45, (ii // 9) % 5, ii % 9], 1)
    assert np.array_equal(pos.astype(np.float64), c / (np.array([3, 5, 9]) - 1) * 2 - 1)
    # mode 1: roi_min + step * i, one float32 rounding per operation
    roi, step = np.float32([-0.7, -0.3, 0.1]), np.float32([0.013, 0.021, 0.007])
    pos = oracle.volume_points(1, (7, 4, 6), roi, step)
    ii = np.arange(7 * 4 * 6)
    c = np.stack([ii % 7, (ii // 7) % 4, ii // 28], 1).astype(np.float32)
    assert np.array_equal(pos, (roi + (step * c).astype(np.float32)).astype(np.float32))
    # mode 0 at a general size: rounding of the quotient then of 2q - 1
    pos = oracle.volume_points(0, (11, 2, 2))
    q = (np.arange(11, dtype=np.float32) / np.float32(10)).astype(np.float32)
    assert np.array_equal(pos[:11, 0], (q * np.float32(2) - np.float32(1)).astype(np.float32))


def test_analyze_round_trip_and_header(native, tmp_path):
    rng = np.random.default_rng(0)
    vol = torch.from_numpy(rng.normal(size=(1, 5, 7, 9)).astype(np.float32))
    path = str(tmp_path / "vol.hdr")
    native.SaveHDRImageTensor(vol, path)
    hdr = open(str(tmp_path / "vol.hdr"), "rb").read()
    img = open(str(tmp_path / "vol.img"), "rb").read()
    assert len(hdr) == 348 and len(img) == 4 * vol.numel()
    # Analyze 7.5: sizeof_hdr @0, extents @32, regular @38, dim[8] @40,
    # datatype @70, bitpix @72, pixdim[8] @76, vox_offset @108
    assert struct.unpack_from("<i", hdr, 0)[0] == 348
    assert hdr[38:39] == b"r"
    dim = struct.unpack_from("<8h", hdr, 40)
    assert dim[:5] == (4, 9, 7, 5, 1)
    assert struct.unpack_from("<h", hdr, 70)[0] == 16 and struct.unpack_from("<h", hdr, 72)[0] == 32
    assert struct.unpack_from("<8f", hdr, 76)[1:4] == (1.0, 1.0, 1.0)
    assert struct.unpack_from("<f", hdr, 108)[0] == 0.0
    assert np.array_equal(np.frombuffer(img, "<f4"), vol.numpy().ravel())  # width fastest
    back = native.LoadHDRImageTensor(path)
    assert back.shape == (1, 5, 7, 9) and torch.equal(back, vol)
    # a big-endian int16 volume written by hand reads back the same
    d = rng.integers(-300, 300, size=(3, 4, 6)).astype(">i2")
    h = bytearray(348)
    struct.pack_into(">i", h, 0, 348)
    struct.pack_into(">8h", h, 40, 3, 6, 4, 3, 1, 1, 1, 1)
    struct.pack_into(">hh", h, 70, 4, 16)
    stem = str(tmp_path / "be")
    open(stem + ".hdr", "wb").write(bytes(h))
    open(stem + ".img", "wb").write(d.tobytes())
    got = native.LoadHDRImageTensor(stem + ".hdr").numpy()
    assert got.shape == (1, 3, 4, 6) and np.array_equal(got[0], d.astype(np.float32))
    assert os.path.exists(stem + ".img")
