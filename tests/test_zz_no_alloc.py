"""atdf_project_bwd_noalloc (the adjoint for callers that cannot allocate)
against atdf_project_bwd with a caller workspace, through the C ABI: two
calls in a row, the second with a smaller input (the library's scratch
array keeps the first call's contents), F32 and BF16 fields; and the
rejection of an input whose workspace exceeds the scratch array."""
import ctypes as C

import numpy as np
import pytest
import torch

from capi_util import Fused
from conftest import random_rays
from scene_util import adaptive_active, device_tree, rays_to_device, tilt_rays

pytestmark = pytest.mark.gpu


def _case(oracle, nrays, seed, bf16):
    from adaptivetensordf_amd import phantom

    t = oracle.tree_build(4)
    act = adaptive_active(t, budget=900, seed=seed)
    tree = device_tree(4, act)
    o1, d1 = random_rays(nrays // 2, 70 + seed)
    o2, d2 = tilt_rays(nrays - nrays // 2, 80 + seed)
    o_t, d_t = rays_to_device(np.concatenate([o1, o2]), np.concatenate([d1, d2]))
    grid = phantom.init_grid(int(act.sum()), 11, 16, std=0.3, seed=seed).cuda()
    if bf16:
        grid = grid.to(torch.bfloat16)
    mlp = phantom.init_mlp(16, 64, seed=seed).cuda()
    fu = Fused(tree, o_t, d_t, 8, grid, mlp, 16, jitter_seed=seed, bf16_decoder=bf16)
    dP = torch.from_numpy(np.random.default_rng(seed).normal(size=nrays).astype(np.float32)).cuda()
    return fu, dP


def test_noalloc_adjoint_matches_workspace_adjoint(oracle, cuda):
    from adaptivetensordf_amd import capi

    lib = capi.lib()
    cap = lib.atdf_project_bwd_noalloc_max_bytes()
    assert cap == 256 << 20
    for nrays, seed, bf16 in ((20000, 1, False), (3000, 2, True)):
        fu, dP = _case(oracle, nrays, seed, bf16)
        need = lib.atdf_project_bwd_workspace_bytes(fu.R, fu.iv.num_intervals, C.byref(fu.field))
        print(f"{nrays} rays, {fu.iv.num_intervals} intervals, bf16={bf16}: workspace {need} of {cap} bytes",
              flush=True)
        assert need <= cap
        gg_ref, gm_ref = fu.backward(dP)
        gg = torch.full(tuple(fu.grid.shape), 0.0, dtype=torch.float32, device="cuda")
        gm = torch.zeros_like(fu.mlp)
        capi.call("atdf_project_bwd_noalloc", *fu.args(), capi.ptr(dP), capi.ptr(gg), capi.ptr(gm), capi.stream_ptr())
        torch.cuda.synchronize()
        eg = ((gg - gg_ref).abs().max() / gg_ref.abs().max()).item()
        em = ((gm - gm_ref).abs().max() / gm_ref.abs().max()).item()
        print(f"  noalloc vs workspace adjoint: grid {eg:.2e}, decoder {em:.2e}", flush=True)
        # the same kernels; only the LDS accumulation order of shared cells differs
        assert eg <= 1e-5 and em <= 1e-5
        assert gg_ref.abs().max().item() > 0
    # an input whose workspace exceeds the scratch array is rejected before any device work
    fu, dP = _case(oracle, 512, 3, False)
    civ = capi.Intervals(fu.iv.node.data_ptr(), fu.iv.tmin.data_ptr(), fu.iv.tmax.data_ptr(), fu.iv.count.data_ptr(),
                         fu.iv.first.data_ptr(), fu.iv.ray.data_ptr(), 40_000_000)
    a = list(fu.args())
    a[7] = C.byref(civ)
    gg = torch.zeros(tuple(fu.grid.shape), dtype=torch.float32, device="cuda")
    gm = torch.zeros_like(fu.mlp)
    st = lib.atdf_project_bwd_noalloc(*a, capi.ptr(dP), capi.ptr(gg), capi.ptr(gm), capi.stream_ptr())
    assert st == capi.ERR_INVALID_ARGUMENT and b"scratch" in lib.atdf_last_error()
    torch.cuda.synchronize()
    assert gg.abs().max().item() == 0.0
