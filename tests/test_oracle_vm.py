"""Pins of the 2-D / 1-D gather, GradientCorrection and ROI-test oracles
against torch CPU library routines and closed forms."""
import numpy as np
import torch
import torch.nn.functional as F


def _torch_2d(grid, index, uv, dtype=torch.float64):
    """Per-sample torch grid_sample(bilinear, border, align_corners=True)."""
    g = torch.from_numpy(grid).to(dtype)
    u = torch.from_numpy(uv).to(dtype)
    outs = []
    for s in range(uv.shape[0]):
        o = F.grid_sample(g[index[s]][None], u[s].view(1, 1, 1, 2), mode="bilinear", padding_mode="border",
                          align_corners=True)
        outs.append(o.view(-1))
    return torch.stack(outs)


def test_grid_sample2d_matches_torch(oracle):
    rng = np.random.default_rng(0)
    grid = rng.normal(size=(5, 6, 7, 9)).astype(np.float32)
    S = 300
    index = rng.integers(0, 5, S)
    uv = rng.uniform(-1.2, 1.2, (S, 2)).astype(np.float32)
    uv[:4] = [[-1, -1], [1, 1], [1, -1], [0, 0]]
    out = oracle.grid_sample2d_fwd(grid, index, uv)
    ref = _torch_2d(grid, index, uv).numpy()
    assert np.abs(out - ref).max() <= 2e-6 * np.abs(ref).max()
    # backward: grid and uv gradients against torch autograd in FP64
    go = rng.normal(size=(S, 6)).astype(np.float32)
    gg, gu = oracle.grid_sample2d_bwd(grid, index, uv, go)
    g_t = torch.from_numpy(grid).double().requires_grad_()
    u_t = torch.from_numpy(uv).double().requires_grad_()
    outs = [F.grid_sample(g_t[index[s]][None], u_t[s].view(1, 1, 1, 2), mode="bilinear", padding_mode="border",
                          align_corners=True).view(-1) for s in range(S)]
    torch.stack(outs).backward(torch.from_numpy(go).double())
    assert np.allclose(gg, g_t.grad.numpy(), rtol=1e-9, atol=1e-12)
    inside = np.all(np.abs(uv) < 0.999, axis=1)
    assert np.allclose(gu[inside], u_t.grad.numpy()[inside], rtol=1e-6, atol=2e-6)  # torch CPU rounds at ~4e-7 here


def test_grid_sample1d_matches_torch(oracle):
    rng = np.random.default_rng(1)
    grid = rng.normal(size=(2, 5, 11, 1)).astype(np.float32)
    S = 200
    index = rng.integers(0, 2, S)
    u = rng.uniform(-1.1, 1.1, S).astype(np.float32)
    out = oracle.grid_sample1d_fwd(grid, index, u)
    uv2 = np.stack([np.zeros_like(u), u], 1).astype(np.float32)  # x spans W = 1, y spans the line
    ref = _torch_2d(grid, index, uv2).numpy()
    assert np.abs(out - ref).max() <= 2e-6 * np.abs(ref).max()
    go = rng.normal(size=(S, 5)).astype(np.float32)
    gg, gu = oracle.grid_sample1d_bwd(grid, index, u, go)
    g_t = torch.from_numpy(grid).double().requires_grad_()
    u_t = torch.from_numpy(uv2).double().requires_grad_()
    outs = [F.grid_sample(g_t[index[s]][None], u_t[s].view(1, 1, 1, 2), mode="bilinear", padding_mode="border",
                          align_corners=True).view(-1) for s in range(S)]
    torch.stack(outs).backward(torch.from_numpy(go).double())
    assert np.allclose(gg, g_t.grad.numpy(), rtol=1e-9, atol=1e-12)
    inside = np.abs(u) < 0.999
    assert np.allclose(gu[inside], u_t.grad.numpy()[inside, 1], rtol=1e-6, atol=2e-6)  # torch CPU rounds at ~4e-7 here


def test_gradient_correction_closed_form(oracle):
    rng = np.random.default_rng(2)
    d = rng.normal(size=(40, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d.astype(np.float32)
    ray = rng.integers(0, 40, 500)
    g = rng.normal(size=(500, 3)).astype(np.float32)
    out = oracle.gradient_correction(d, ray, g)
    dd = d[ray].astype(np.float64)
    ref = g - (g * dd).sum(1, keepdims=True) * dd
    assert np.abs(out - ref).max() <= 1e-6
    assert np.abs((out.astype(np.float64) * dd).sum(1)).max() <= 1e-6  # orthogonal to the ray
    # a gradient along the ray is removed entirely
    assert np.abs(oracle.gradient_correction(d, np.arange(40), d * 3.0)).max() <= 1e-6


def test_compute_bound_weight(oracle):
    rng = np.random.default_rng(3)
    node = rng.integers(0, 50, 1000)
    roi = np.array([3, 7, 11, 49])
    out = oracle.compute_bound_weight(node, roi)
    assert np.array_equal(out, np.where(np.isin(node, roi), np.arange(1000), -1))
