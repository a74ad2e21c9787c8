"""Per-sample operators of the reference API on the GPU vs the oracle /
torch library routines: indirect 3-D gather (bit-exact forward), its
backward, line integral, FP64 per-node loss, MSE, Adam, layout."""
import ctypes as C

import numpy as np
import pytest
import torch
import torch.nn.functional as F

pytestmark = pytest.mark.gpu


def test_indirect_grid_sample3d(oracle, cuda):
    from adaptivetensordf_amd import native

    rng = np.random.default_rng(0)
    grid = rng.normal(size=(40, 16, 11, 11, 11)).astype(np.float32)
    S = 20000
    index = rng.integers(0, 40, S)
    uv = rng.uniform(-1, 1, (S, 3)).astype(np.float32)
    uv[:6] = [[-1, -1, -1], [1, 1, 1], [-1.0000001, 0, 0], [1.0000001, 0.5, -0.5], [0, 0, 0], [0.1, -1, 1]]
    g_t = torch.from_numpy(grid).cuda().requires_grad_()
    uv_t = torch.from_numpy(uv).cuda().requires_grad_()
    out = native().IndirectGridSample3D(g_t, torch.from_numpy(index).cuda(), uv_t)
    ref = oracle.grid_sample3d_fwd(grid, index, uv)
    assert np.array_equal(out.detach().cpu().numpy().view(np.uint32), ref.view(np.uint32))
    go = rng.normal(size=(S, 16)).astype(np.float32)
    gg, gu = torch.autograd.grad(out, [g_t, uv_t], torch.from_numpy(go).cuda())
    rg, ru = oracle.grid_sample3d_bwd(grid, index, uv, go)
    assert np.abs(gg.cpu().double().numpy() - rg).max() <= 1e-5 * np.abs(rg).max()
    assert np.abs(gu.cpu().double().numpy() - ru).max() <= 1e-5 * np.abs(ru).max()


def test_integrate_xray_and_loss_per_node(oracle, cuda):
    from adaptivetensordf_amd import native

    rng = np.random.default_rng(1)
    S, R, N = 50000, 3000, 500
    v = rng.normal(size=(S, 1)).astype(np.float32)
    w = rng.uniform(0, 0.1, (S, 1)).astype(np.float32)
    ray = np.sort(rng.integers(0, R, S))
    P = native().IntegrateSamplesXRay(torch.from_numpy(v).cuda(), torch.from_numpy(w).cuda(),
                                      torch.from_numpy(ray).cuda(), 1, R).cpu().numpy()
    ref = oracle.integrate(v, w[:, 0], ray, R)
    assert np.abs(P - ref).max() <= 1e-5 * np.abs(ref).max()
    sl = native().SampleList()
    sl.ray_index = torch.from_numpy(ray).cuda()
    sl.weight = torch.from_numpy(w[:, 0].copy()).cuda()
    node = rng.integers(0, N, S)
    sl.node_id = torch.from_numpy(node).cuda()
    prl = rng.uniform(0, 1, (1, R)).astype(np.float32)
    ls, ws = native().AccumulateSampleLossPerNode(sl, torch.from_numpy(prl).cuda(), N)
    rl, rw = oracle.loss_per_node(prl, ray, w[:, 0], node, N)
    assert ls.dtype == torch.float64
    assert np.allclose(ls.cpu().numpy(), rl, rtol=1e-12, atol=1e-15)
    assert np.allclose(ws.cpu().numpy(), rw, rtol=1e-12, atol=1e-15)


def test_mse_adam_layout(cuda):
    from adaptivetensordf_amd import capi

    s = capi.stream_ptr()
    R = 100003
    P = torch.rand(R, device="cuda")
    T = torch.rand(R, device="cuda")
    dP = torch.empty(R, device="cuda")
    loss = torch.empty(1, dtype=torch.float64, device="cuda")
    ws = torch.empty(capi.lib().atdf_mse_workspace_bytes(R), dtype=torch.uint8, device="cuda")
    capi.call("atdf_mse_loss", capi.ptr(P), capi.ptr(T), R, C.c_float(1.0), capi.ptr(dP), capi.ptr(loss),
              capi.ptr(ws), ws.numel(), s)
    ref = ((P.double() - T.double()) ** 2).mean()
    assert abs(loss.item() - ref.item()) <= 1e-12 * ref.item()
    assert torch.allclose(dP, 2 * (P - T) / R, rtol=1e-6)
    # Adam against torch.optim.Adam, 5 steps
    n = 10001
    p0 = torch.randn(n, device="cuda")
    p_ref = p0.clone().requires_grad_()
    opt = torch.optim.Adam([p_ref], lr=1e-2)
    p = p0.clone()
    g = torch.zeros(n, device="cuda")
    m = torch.zeros(n, device="cuda")
    v = torch.zeros(n, device="cuda")
    for step in range(1, 6):
        grad = torch.randn(n, device="cuda", generator=torch.Generator("cuda").manual_seed(step))
        g.copy_(grad)
        capi.call("atdf_adam_step", capi.ptr(p), capi.ptr(g), capi.ptr(m), capi.ptr(v), n, C.c_float(1e-2),
                  C.c_float(0.9), C.c_float(0.999), C.c_float(1e-8), step, 1, s)
        p_ref.grad = grad.clone()
        opt.step()
    assert torch.allclose(p, p_ref.detach(), rtol=1e-5, atol=1e-6)
    assert torch.all(g == 0)
    # layout round trip and bf16 rounding
    x = torch.randn(7, 16, 5, 5, 5, device="cuda")
    y = torch.empty(7, 5, 5, 5, 16, device="cuda")
    z = torch.empty_like(x)
    capi.call("atdf_grid_layout", capi.ptr(x), capi.ptr(y), 7, 16, 5, 0, s)
    capi.call("atdf_grid_layout", capi.ptr(y), capi.ptr(z), 7, 16, 5, 1, s)
    assert torch.equal(y, x.permute(0, 2, 3, 4, 1)) and torch.equal(z, x)
    b = torch.empty(x.numel(), dtype=torch.int16, device="cuda")
    capi.call("atdf_f32_to_bf16", capi.ptr(x), capi.ptr(b), x.numel(), s)
    assert torch.equal(b.view(torch.bfloat16), x.reshape(-1).to(torch.bfloat16))


def test_invalid_arguments_are_reported(cuda):
    from adaptivetensordf_amd import capi

    with pytest.raises(capi.AtdfError) as e:
        capi.call("atdf_grid_sample3d_fwd", None, 0, 16, 11, 11, 11, None, None, 10, None, capi.stream_ptr())
    assert e.value.status == capi.ERR_INVALID_ARGUMENT


def test_device_limits_for_node_major_adjoint(cuda):
    """The node-major adjoint keeps a node's 11^3 x 16 f32 gradient in LDS."""
    props = torch.cuda.get_device_properties(0)
    lds = getattr(props, "shared_memory_per_block", None)
    print(f"{props.name} {props.gcnArchName}: max LDS per workgroup {lds} bytes, {props.multi_processor_count} CUs",
          flush=True)
    assert lds is None or lds >= 150000
