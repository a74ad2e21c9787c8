"""The adjoint's two-pass form (decoder pass + LDS scatter pass over a staged
per-sample buffer) against its fused form and the FP64 oracle.

Both forms add the scatter in a run-dependent order, so they agree to FP32
rounding: gradients within 1e-5 of each other (max|a - b| / max|b|), and the
FP32 path within 1e-5 of the FP64 oracle.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from capi_util import Fused
from conftest import random_rays
from scene_util import adaptive_active, device_tree, rays_to_device, tilt_rays

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = a.detach().double().cpu()
    b = b.detach().double().cpu()
    return float((a - b).abs().max() / b.abs().max().clamp_min(1e-30))


def make(oracle, C_=16, bf16_grid=False, bf16_dec=False, depth=3, K=6, nrays=768, seed=0, budget=300):
    from adaptivetensordf_amd import phantom

    t = oracle.tree_build(depth)
    act = adaptive_active(t, budget=budget, seed=seed)
    tree = device_tree(depth, act)
    na = int(act.sum())
    grid32 = phantom.init_grid(na, 11, C_, std=0.3, seed=seed)
    mlp = phantom.init_mlp(C_, 64, seed=seed).cuda()
    o1, d1 = random_rays(nrays // 2, 50 + seed)
    o2, d2 = tilt_rays(nrays - nrays // 2, 60 + seed)
    o, d = np.concatenate([o1, o2]), np.concatenate([d1, d2])
    o_t, d_t = rays_to_device(o, d)
    grid = grid32.cuda().to(torch.bfloat16) if bf16_grid else grid32.cuda()
    fu = Fused(tree, o_t, d_t, K, grid, mlp, C_, jitter_mode=2, jitter_seed=7 + seed, bf16_decoder=bf16_dec)
    return dict(t=t, act=act, tree=tree, grid32=grid32, mlp=mlp, o=o, d=d, fu=fu)


@pytest.mark.parametrize("C_,bf16_grid,bf16_dec", [(16, False, False), (16, True, True), (16, True, False),
                                                   (16, False, True), (8, False, False), (8, True, True)])
def test_two_pass_matches_fused(oracle, cuda, C_, bf16_grid, bf16_dec):
    cs = make(oracle, C_, bf16_grid, bf16_dec, seed=C_ + 2 * bf16_grid + bf16_dec)
    fu = cs["fu"]
    dP = torch.from_numpy(np.random.default_rng(3).normal(size=fu.R).astype(np.float32)).cuda()
    st = fu.staging_bytes()
    print(f"C={C_} bf16 grid={bf16_grid} dec={bf16_dec}: {fu.iv.num_samples} samples, "
          f"{fu.iv.num_intervals} intervals, staging {st} bytes", flush=True)
    assert st >= fu.iv.num_samples * (16 + 4 * C_)
    gg_f, gm_f = fu.backward(dP)
    gg_t, gm_t = fu.backward(dP, staging=st)
    eg, em = rel(gg_t, gg_f), rel(gm_t, gm_f)
    print(f"two-pass vs fused: grid {eg:.2e}, decoder {em:.2e}", flush=True)
    assert gg_f.abs().max() > 0 and gm_f.abs().max() > 0
    assert eg <= 1e-5 and em <= 1e-5


@pytest.mark.parametrize("C_", [16, 8])
def test_two_pass_matches_oracle(oracle, cuda, C_):
    cs = make(oracle, C_, seed=20 + C_)
    fu = cs["fu"]
    fm = oracle.FusedModel(cs["t"]["node_min"], cs["t"]["node_max"], cs["act"], cs["grid32"].numpy(),
                           cs["mlp"].cpu().numpy(), C_, K=6, jitter_mode=2, seed=7 + 20 + C_)
    dP = np.random.default_rng(C_).normal(size=fu.R)
    gg, gm = fu.backward(torch.from_numpy(dP).float().cuda(), staging=fu.staging_bytes())
    ref_g, ref_m = fm.project_bwd(cs["o"], cs["d"], dP)
    eg = np.abs(gg.cpu().double().numpy() - ref_g).max() / np.abs(ref_g).max()
    em = np.abs(gm.cpu().double().numpy() - ref_m).max() / np.abs(ref_m).max()
    print(f"C={C_}: two-pass vs FP64 oracle: grid {eg:.2e}, decoder {em:.2e}", flush=True)
    assert eg <= 1e-5 and em <= 1e-5


def test_two_pass_small_staging_runs_fused(oracle, cuda):
    """A staging region too small for the list's samples leaves the fused
    kernel in charge (decided on the device); results as before."""
    cs = make(oracle, 16, seed=31)
    fu = cs["fu"]
    dP = torch.from_numpy(np.random.default_rng(5).normal(size=fu.R).astype(np.float32)).cuda()
    st = fu.staging_bytes()
    per_sample = 16 + 4 * 16
    gg_f, gm_f = fu.backward(dP)
    gg_s, gm_s = fu.backward(dP, staging=st - per_sample * (fu.iv.num_samples // 2))
    assert rel(gg_s, gg_f) <= 1e-5 and rel(gm_s, gm_f) <= 1e-5


def test_two_pass_prepared_adam(oracle, cuda):
    """atdf_project_prepare + atdf_project_bwd_prepared with the fused Adam,
    with and without the staging region: the same parameters and moments
    (to FP32 rounding; a parameter may move by its last ulp where the first
    Adam step is sign-like), the same decoder gradient, grad_grid left at 0."""
    from adaptivetensordf_amd import capi

    rng = np.random.default_rng(9)
    t = oracle.tree_build(3)
    act = adaptive_active(t, budget=200, seed=4)
    tree = device_tree(3, act)
    na = int(act.sum())
    n = 12000
    o, d = random_rays(n, 77)
    o_t, d_t = rays_to_device(o, d)
    master = (0.3 * torch.randn(na, 11, 11, 11, 16, generator=torch.Generator().manual_seed(2))).cuda()
    from adaptivetensordf_amd import phantom

    mlp = phantom.init_mlp(16, 64, seed=5).cuda()
    m0 = (0.01 * torch.randn_like(master)).abs()
    v0 = 1e-4 * torch.rand_like(master)
    dP = torch.from_numpy(rng.normal(size=n).astype(np.float32) * 1e-2).cuda()
    out = []
    for two_pass in (False, True):
        p, m, v = master.clone(), m0.clone(), v0.clone()
        fu = Fused(tree, o_t, d_t, 8, p, mlp, 16, jitter_mode=2, jitter_seed=3)
        ws = fu.workspace(fu.staging_bytes() if two_pass else 0)
        gg = torch.zeros_like(p)
        gm = torch.zeros_like(mlp)
        adam = capi.GridAdam(p.data_ptr(), None, m.data_ptr(), v.data_ptr(), 1e-2, 0.9, 0.999, 1e-8, 3)
        s = capi.stream_ptr()
        capi.call("atdf_project_prepare", *fu.args(), capi.ptr(ws), ws.numel(), s)
        capi.call("atdf_project_bwd_prepared", *fu.args(), capi.ptr(dP), capi.ptr(gg), capi.ptr(gm), C.byref(adam),
                  capi.ptr(ws), ws.numel(), s)
        torch.cuda.synchronize()
        out.append((p, m, v, gg, gm))
    (p0, m0_, v0_, g0, gm0), (p1, m1, v1, g1, gm1) = out
    moved = (p0 - master).abs().max().item()
    dp = (p1 - p0).abs().max().item()
    print(f"max update {moved:.3e}, |dp| {dp:.2e}, m {rel(m1, m0_):.2e}, v {rel(v1, v0_):.2e}, "
          f"decoder {rel(gm1, gm0):.2e}", flush=True)
    assert moved > 1e-3
    assert g0.abs().max() == 0 and g1.abs().max() == 0
    assert dp <= 1e-5 * moved
    assert rel(m1, m0_) <= 1e-5 and rel(v1, v0_) <= 1e-5 and rel(gm1, gm0) <= 1e-5


def test_two_pass_rays_that_miss(oracle, cuda):
    """A list in which some rays miss the tree and some nodes get no sample."""
    cs = make(oracle, 16, seed=41, nrays=256)
    fu0 = cs["fu"]
    o = np.concatenate([cs["o"], np.tile(np.array([[5.0, 5.0, 5.0]], np.float32), (64, 1))])
    d = np.concatenate([cs["d"], np.tile(np.array([[1.0, 0.0, 0.0]], np.float32), (64, 1))])
    o_t, d_t = rays_to_device(o, d)
    fu = Fused(cs["tree"], o_t, d_t, 6, fu0.grid, fu0.mlp, 16, jitter_mode=2, jitter_seed=48)
    dP = torch.from_numpy(np.random.default_rng(6).normal(size=fu.R).astype(np.float32)).cuda()
    gg_f, gm_f = fu.backward(dP)
    gg_t, gm_t = fu.backward(dP, staging=fu.staging_bytes())
    assert rel(gg_t, gg_f) <= 1e-5 and rel(gm_t, gm_f) <= 1e-5
