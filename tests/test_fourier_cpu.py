"""Decoder on feature rows and the Fourier regulariser on the CPU: the
oracle's FP64 decoder and its adjoint against libtorch autograd, the ring
means against closed forms and against the reference's own sequence of
libtorch calls (arange, meshgrid, grid_sample with border padding and
align_corners), the FFT band values against direct DFT sums, invariances of
the whole term, and the C ABI's argument checks."""
import ctypes as C
import math

import numpy as np
import pytest
import torch
import torch.nn.functional as F


def _mlp_params(rng, Cc, H, scale=0.4):
    n = H * Cc + H + H * H + H + H + 1
    return (rng.normal(size=n) * scale).astype(np.float32)


def _torch_decode(params, Cc, H, beta, x):
    p = torch.from_numpy(params).double().requires_grad_()
    xt = torch.from_numpy(x).double().requires_grad_()
    o = 0
    W1 = p[o:o + H * Cc].view(H, Cc); o += H * Cc
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].view(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + H].view(1, H); o += H
    b3 = p[o:o + 1]
    h = torch.relu(xt @ W1.T + b1)
    h = torch.relu(h @ W2.T + b2)
    y = F.softplus(h @ W3.T + b3, beta=beta, threshold=20.0)[:, 0]
    return y, p, xt


@pytest.mark.parametrize("Cc", [16, 8])
def test_decode_oracle_against_torch_autograd(oracle, Cc):
    rng = np.random.default_rng(Cc)
    H, beta = 64, 2.0
    params = _mlp_params(rng, Cc, H)
    x = rng.normal(size=(257, Cc)).astype(np.float32)
    dd = rng.normal(size=257)
    dens, gx, gm = oracle.decode_features(params, Cc, H, beta, x, dd)
    y, p, xt = _torch_decode(params, Cc, H, beta, x)
    (y * torch.from_numpy(dd)).sum().backward()
    assert np.abs(dens - y.detach().numpy()).max() <= 1e-12 * np.abs(dens).max()
    assert np.abs(gx - xt.grad.numpy()).max() <= 1e-12 * np.abs(gx).max()
    assert np.abs(gm - p.grad.numpy()).max() <= 1e-12 * np.abs(gm).max()
    # the FP32 oracle decoder agrees to float rounding
    _, d32 = oracle.mlp_fwd(params, Cc, H, beta, x)
    assert np.abs(dens - d32).max() <= 1e-5 * np.abs(dens).max()


def test_decode_oracle_bf16_rounding(oracle):
    # BF16 emulation: inputs and the first two layers' weights rounded
    rng = np.random.default_rng(7)
    Cc, H, beta = 16, 64, 2.0
    params = _mlp_params(rng, Cc, H)
    x = rng.normal(size=(64, Cc)).astype(np.float32)
    dens_bf, gx_bf, gm_bf = oracle.decode_features(params, Cc, H, beta, x, np.ones(64), bf=True)
    bf = lambda a: torch.from_numpy(a).to(torch.bfloat16).double().numpy()
    o = H * Cc
    pr = params.astype(np.float64).copy()
    pr[:o] = bf(params[:o])
    pr[o + H:o + H + H * H] = bf(params[o + H:o + H + H * H])
    # layer 1 and 2 inputs rounded too: h1 = relu(W1 bf(x) + b1); h2 uses bf(h1)
    W1 = pr[:o].reshape(H, Cc); b1 = pr[o:o + H]
    W2 = pr[o + H:o + H + H * H].reshape(H, H); b2 = pr[o + H + H * H:o + 2 * H + H * H]
    W3 = pr[o + 2 * H + H * H:o + 3 * H + H * H]; b3 = pr[-1]
    h1 = np.maximum(bf(x.astype(np.float32)) @ W1.T + b1, 0)
    h2 = np.maximum(bf(h1.astype(np.float32)) @ W2.T + b2, 0)
    yv = h2 @ W3 + b3
    want = np.where(yv * beta > 20, yv, np.log1p(np.exp(yv * beta)) / beta)
    assert np.abs(dens_bf - want).max() <= 1e-6 * np.abs(want).max()
    # the adjoint stays near the FP64 one (BF16 rounding flips a few ReLU
    # masks, so entries can differ; the norms agree)
    _, gx, gm = oracle.decode_features(params, Cc, H, beta, x, np.ones(64))
    assert np.linalg.norm(gx_bf - gx) <= 0.1 * np.linalg.norm(gx)
    assert np.linalg.norm(gm_bf - gm) <= 0.1 * np.linalg.norm(gm)


def _angles(i):
    n = int(2 * math.pi * i)
    step = float(np.float32(360.0 / n))
    return n, step, len(torch.arange(0, 360, step))


def test_ring_angle_count_follows_torch_arange():
    # arange(0, 360, (float)(360 / n)) holds n or n + 1 angles: one more
    # whenever the float step rounds below 360 / n
    counts = {i: _angles(i) for i in range(1, 40)}
    assert counts[1][2] == 6 and counts[2][2] == 12 and counts[3][2] == 18
    assert counts[4][0] == 25 and counts[4][2] == 26
    extra = [i for i, (n, s, m) in counts.items() if m == n + 1]
    assert 5 < len(extra) < 35


def test_ring_means_closed_forms(oracle):
    X = Z = 41
    out = oracle.fourier_ring_means(np.full((2, X, Z), 2.5, np.float32))
    assert np.all(np.isinf(out[0])) and np.allclose(out[1:], 2.5, rtol=1e-6)
    # a linear image a r + b c + g: the angle sums vanish over a whole turn,
    # so ring i holds a (X/2 + i e_i) + b Z/2 + g, e_i = 1/m for the rings
    # whose extra angle sits at ~360 degrees (cos 1, sin 0), else 0
    a, b, g = 0.75, -1.5, 3.0
    r, c = np.meshgrid(np.arange(X), np.arange(Z), indexing="ij")
    img = (a * r + b * c + g).astype(np.float32)[None]
    out = oracle.fourier_ring_means(img)
    for i in range(1, X // 2):
        n, _, m = _angles(i)
        e = (1.0 / m) if m == n + 1 else 0.0
        want = a * (X // 2 + i * e) + b * (Z // 2) + g
        assert abs(out[i, 0] - want) <= 2e-4 * abs(want), i


@pytest.mark.parametrize("X,Z", [(17, 17), (21, 13)])
def test_ring_means_against_torch_grid_sample(oracle, X, Z):
    rng = np.random.default_rng(X + Z)
    Y = 3
    img = rng.uniform(0, 1, (Y, X, Z)).astype(np.float32)
    out = oracle.fourier_ring_means(img)
    fv = torch.from_numpy(img)[None]
    ox, oy = X // 2, Z // 2
    for i in range(1, X // 2):
        n = int(2 * math.pi * i)
        step = float(np.float32(360.0 / n))
        th = torch.arange(0, 360, step)
        xr = i * torch.cos(th / 180.0 * math.pi) + ox
        yr = i * torch.sin(th / 180.0 * math.pi) + oy
        gx, gy = torch.meshgrid(xr, yr, indexing="ij")
        grid = torch.stack([gy, gx], -1) / (X - 1) * 2 - 1
        v = F.grid_sample(fv, grid[None], mode="bilinear", padding_mode="border", align_corners=True)
        want = v.mean((2, 3))[0].numpy()
        assert np.abs(out[i] - want).max() <= 2e-6 * np.abs(want).max(), i


def test_fourier_bands_against_direct_dft(oracle):
    rng = np.random.default_rng(3)
    X, Y, Z = 15, 4, 15
    d = rng.uniform(0, 1, (X, Y, Z))
    for rate_w, rows in ((0.01, [0]), (0.3, [-1, 0, 1])):
        _, parts = oracle.fourier_terms(d, rate=0.2, rate_w=rate_w)
        cx = X // 2
        xs, zs = np.arange(X), np.arange(Z)
        for y in range(Y):
            sl = d[:, y, :]
            # centre column of fx: mean over the band's x frequencies of |DFT| at z frequency 0
            mags = [abs((sl * np.exp(-2j * np.pi * k * xs[:, None] / X)).sum()) / (X * Z) for k in rows]
            assert parts["fx"][y, Z // 2] == pytest.approx(np.mean(mags), rel=1e-12)
        if rate_w == 0.01:
            for y in range(Y):
                sl = d[:, y, :]
                for k in (1, 2, -3):
                    want = abs((sl * np.exp(-2j * np.pi * k * xs[:, None] / X)).sum()) / (X * Z)
                    assert parts["fy"][y, cx + k] == pytest.approx(want, rel=1e-12)
                kz = 2
                want = abs((sl * np.exp(-2j * np.pi * kz * zs[None, :] / Z)).sum()) / (X * Z)
                assert parts["fx"][y, Z // 2 + kz] == pytest.approx(want, rel=1e-12)


def test_fourier_term_invariances(oracle):
    rng = np.random.default_rng(11)
    d = rng.uniform(0, 1, (21, 5, 21))
    base, parts = oracle.fourier_terms(d)
    assert np.isfinite(base) and base > 0
    assert parts["mask_x"].any() and parts["mask_y"].any()
    # |FFT| is linear in a positive scale and blind to circular shifts in x and z
    assert oracle.fourier_terms(2.5 * d)[0] == pytest.approx(2.5 * base, rel=1e-6)
    assert oracle.fourier_terms(np.roll(d, 4, axis=0))[0] == pytest.approx(base, rel=1e-6)
    assert oracle.fourier_terms(np.roll(d, -3, axis=2))[0] == pytest.approx(base, rel=1e-6)
    # the centre of the ring profile is 0, its neighbours the undefined row
    total = 21 // 2
    assert np.all(parts["circle"][:, total] == 0) and np.all(np.isinf(parts["circle"][:, total + 1]))
    assert np.all(parts["factor"][:, total] == 0)


def test_decode_and_ring_boundary_without_device_work():
    from adaptivetensordf_amd import capi

    lib = capi.lib()
    one = C.c_void_p(1)
    f = capi.Field()
    f.C, f.H = 12, 64
    st = lib.atdf_decode_features(one, 4, C.byref(f), one, None)
    assert st == capi.ERR_INVALID_ARGUMENT and b"not built" in lib.atdf_last_error()
    f.C, f.mlp = 16, None
    st = lib.atdf_decode_features(one, 4, C.byref(f), one, None)
    assert st == capi.ERR_INVALID_ARGUMENT and b"null decoder" in lib.atdf_last_error()
    f.mlp = 1
    assert lib.atdf_decode_features(one, 0, C.byref(f), one, None) == 0
    assert lib.atdf_decode_features_bwd_workspace_bytes(0, C.byref(f)) == 0
    np_ = 64 * 16 + 64 + 64 * 64 + 64 + 64 + 1
    assert lib.atdf_decode_features_bwd_workspace_bytes(1000, C.byref(f)) == 4 * 64 * np_  # 63 tiles -> 16 workgroups of 4 waves
    st = lib.atdf_decode_features_bwd(one, 1000, C.byref(f), one, one, one, one, 16, None)
    assert st == capi.ERR_INVALID_ARGUMENT and b"workspace" in lib.atdf_last_error()
    st = lib.atdf_fourier_ring_means(one, 2, 1, 5, one, one, 1 << 20, None)
    assert st == capi.ERR_INVALID_ARGUMENT and b"too small" in lib.atdf_last_error()
    st = lib.atdf_fourier_ring_means(one, 2, 3001, 5, one, one, 1 << 30, None)
    assert st == capi.ERR_INVALID_ARGUMENT and b"angles" in lib.atdf_last_error()
    st = lib.atdf_fourier_ring_means(one, 2, 17, 9, one, one, 8, None)
    assert st == capi.ERR_INVALID_ARGUMENT and b"workspace" in lib.atdf_last_error()
    assert lib.atdf_fourier_ring_workspace_bytes(17, 9) == 8 * 8 * (17 + 9) + 4 * 4 * 8
