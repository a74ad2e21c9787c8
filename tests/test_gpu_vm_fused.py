"""The tensorial (VM) geometry through the fused projection kernels
(atdf_project_fwd / atdf_project_bwd with an atdf_vm field) against the
oracle's composition of the same steps: traversal samples, local
coordinates, the 2-D / 1-D / 3-D gathers of GeometryTensorQO
(geometry_tensor_qo.cpp:257-380, tensor_op "add"), the FP64 decoder and the
FP64 line integral; gradients of every plane, line, explicit grid and the
decoder against the oracle's backward of each part.  Also: the ctypes
layout of atdf_vm, the fused projection against the composed
ComputeImageVM on identical samples, and TrainStepVM's gradients and Adam
step."""
import ctypes as C

import numpy as np
import pytest
import torch

from conftest import random_rays
from scene_util import adaptive_active, device_tree

pytestmark = pytest.mark.gpu

NAMES = ["plane_yz", "plane_xz", "plane_xy", "line_x", "line_y", "line_z", "grid_out"]
PAIRS = (("plane_yz", "line_x"), ("plane_xz", "line_y"), ("plane_xy", "line_z"))


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)


def _setup(native, oracle, C_, Gp, Gl, Go, lo, hi, seed=0, depth=3, budget=150, n_rays=300):
    rng = np.random.default_rng(seed)
    t = oracle.tree_build(depth)
    act = adaptive_active(t, budget=budget, seed=seed)
    tree = device_tree(depth, act)
    roi = tree.RoiSplit(lo, hi, 0.0)
    n_in, n_out = int(roi.in_roi_node_index.numel()), int(roi.out_roi_node_index.numel())
    arr = {}
    f = native.VMFields()
    shapes = {"plane_yz": (n_in, C_, Gp, Gp), "plane_xz": (n_in, C_, Gp, Gp), "plane_xy": (n_in, C_, Gp, Gp),
              "line_x": (1, C_, Gl, 1), "line_y": (1, C_, Gl, 1), "line_z": (1, C_, Gl, 1),
              "grid_out": (n_out, C_, Go, Go, Go)}
    for k in NAMES:
        if k == "grid_out" and n_out == 0:
            continue
        arr[k] = rng.uniform(-0.6, 0.6, shapes[k]).astype(np.float32)
        setattr(f, k, torch.from_numpy(arr[k]).cuda().requires_grad_())
    f.tensor_op = "add"
    f.roi_min, f.roi_max = lo, hi
    mlp = (rng.normal(size=64 * C_ + 64 + 64 * 64 + 64 + 64 + 1) * 0.2).astype(np.float32)
    o, d = random_rays(n_rays, 40 + seed, spread=0.9)
    return t, act, tree, roi, f, arr, mlp, o, d


def _oracle(oracle, t, act, roi, arr, mlp, o, d, K, jm, seed, C_, lohi, dLdP=None, beta=2.0, bf=False):
    """Projection [R] (FP64) and, with dLdP, the gradients of every field
    and of the decoder."""
    tr = oracle.traverse(t["node_min"], t["node_max"], oracle.active_ids(act), o, d, K, jitter_mode=jm, seed=seed)
    g, nid, w, ray = tr["global_coordinate"], tr["node_id"], tr["weight"], tr["ray_index"]
    pin, pout = roi.prefix_inroi.cpu().numpy(), roi.prefix_outroi.cpu().numpy()
    flag = pin[nid] >= 0
    loc = oracle.local_samples(t["node_min"], t["node_max"], g, nid)
    li, gi = loc[flag], g[flag]
    slot = pin[nid[flag]]
    lo, hi = np.float32(lohi[0][0]), np.float32(lohi[1][0])
    pl = {"plane_yz": li[:, [1, 2]], "plane_xz": li[:, [0, 2]], "plane_xy": li[:, [0, 1]]}
    # the ROI-normalised x of the line, with a true division
    ln = {"line_x": (np.float32(2.0) * (gi[:, 0] - lo)) / (hi - lo) - np.float32(1.0), "line_y": li[:, 1],
          "line_z": gi[:, 2]}
    pl = {k: np.ascontiguousarray(v, np.float32) for k, v in pl.items()}
    ln = {k: np.ascontiguousarray(v, np.float32).reshape(-1, 1) for k, v in ln.items()}
    zero = np.zeros(li.shape[0], np.int64)
    P = {k: oracle.grid_sample2d_fwd(arr[k], slot, pl[k]) for k in pl}
    L = {k: oracle.grid_sample1d_fwd(arr[k], zero, ln[k]) for k in ln}
    fin = P["plane_yz"] * L["line_x"]
    fin = fin + P["plane_xz"] * L["line_y"]
    fin = fin + P["plane_xy"] * L["line_z"]
    feats = np.zeros((g.shape[0], C_), np.float32)
    feats[flag] = fin
    out_slot = pout[nid[~flag]]
    if (~flag).any():
        feats[~flag] = oracle.grid_sample3d_fwd(arr["grid_out"], out_slot, loc[~flag])
    R = o.shape[0]
    dd = None if dLdP is None else dLdP[ray].astype(np.float64) * w.astype(np.float64)
    dens, gx, gm = oracle.decode_features(mlp, C_, 64, beta, feats, dd, bf=bf)
    proj = np.zeros(R)
    np.add.at(proj, ray, dens * w.astype(np.float64))
    info = dict(n_in_samples=int(flag.sum()), n_out_samples=int((~flag).sum()))
    if dLdP is None:
        return proj, None, info
    grads = {"mlp": gm}
    gin = gx[flag]
    for pk, lk in PAIRS:
        dP = (gin * L[lk]).astype(np.float32)
        dL = (gin * P[pk]).astype(np.float32)
        grads[pk], _ = oracle.grid_sample2d_bwd(arr[pk], slot, pl[pk], dP)
        grads[lk], _ = oracle.grid_sample1d_bwd(arr[lk], zero, ln[lk], dL)
    if (~flag).any():
        grads["grid_out"], _ = oracle.grid_sample3d_bwd(arr["grid_out"], out_slot, loc[~flag],
                                                        gx[~flag].astype(np.float32))
    return proj, grads, info


def _gpu(native, tree, roi, f, mlp, o, d, K, jm, seed, C_, Go, bf=False, two_pass=True):
    sp = native.SamplingParams()
    sp.max_samples_per_node = K
    sp.jitter_mode = jm
    sp.jitter_seed = seed
    fp = native.FieldParams()
    fp.C, fp.G, fp.H = C_, Go, 64
    fp.bf16_decoder = bf
    fp.two_pass_adjoint = two_pass
    m = torch.from_numpy(mlp).cuda().requires_grad_()
    rays = native.RayList(torch.from_numpy(o).cuda(), torch.from_numpy(d).cuda())
    return native.ProjectRaysVM(tree, roi, f, m, sp, fp, rays), m, sp, fp, rays


@pytest.mark.parametrize("two_pass", [True, False])
@pytest.mark.parametrize("C_,Gp,Gl,Go,jm", [(16, 9, 17, 5, 0), (16, 11, 33, 11, 2), (8, 9, 17, 5, 2)])
def test_fused_vm_projection_and_adjoint(oracle, native, cuda, C_, Gp, Gl, Go, jm, two_pass):
    """Both forms of the adjoint: the two-pass kernels (decoder pass, staged
    plane / line gradients, LDS scatter passes) and the fused kernel."""
    lo, hi = [-0.55, -0.45, -0.6], [0.5, 0.6, 0.4]
    t, act, tree, roi, f, arr, mlp, o, d = _setup(native, oracle, C_, Gp, Gl, Go, lo, hi)
    K, seed = 4, 11
    P, m, *_ = _gpu(native, tree, roi, f, mlp, o, d, K, jm, seed, C_, Go, two_pass=two_pass)
    rng = np.random.default_rng(3)
    dLdP = rng.normal(size=o.shape[0]).astype(np.float32)
    ref, rgrads, info = _oracle(oracle, t, act, roi, arr, mlp, o, d, K, jm, seed, C_, (lo, hi), dLdP)
    print(f"C={C_} Gp={Gp} Gl={Gl} Go={Go} jitter={jm} two-pass={two_pass}: {info}", flush=True)
    assert info["n_in_samples"] > 0 and info["n_out_samples"] > 0
    Pn = P.detach().cpu().double().numpy()
    err = np.abs(Pn - ref).max() / np.abs(ref).max()
    print(f"  projection rel err {err:.2e}", flush=True)
    assert err <= 1e-5
    params = [getattr(f, k) for k in NAMES] + [m]
    grads = torch.autograd.grad(P, params, torch.from_numpy(dLdP).cuda())
    for k, g in zip(NAMES + ["mlp"], grads):
        e = _rel(g.cpu().double().numpy(), rgrads[k])
        print(f"  grad {k}: rel err {e:.2e}", flush=True)
        assert e <= 1e-5, k


@pytest.mark.parametrize("two_pass", [True, False])
def test_fused_vm_all_nodes_in_roi(oracle, native, cuda, two_pass):
    """An ROI holding every node: no explicit grids at all (num_grids 0)."""
    lo, hi = [-1.0, -1.0, -1.0], [1.0, 1.0, 1.0]
    C_, Gp, Gl, Go = 16, 7, 13, 5
    t, act, tree, roi, f, arr, mlp, o, d = _setup(native, oracle, C_, Gp, Gl, Go, lo, hi, seed=2)
    assert int(roi.out_roi_node_index.numel()) == 0
    P, m, *_ = _gpu(native, tree, roi, f, mlp, o, d, 3, 2, 5, C_, Go, two_pass=two_pass)
    dLdP = np.random.default_rng(4).normal(size=o.shape[0]).astype(np.float32)
    ref, rgrads, info = _oracle(oracle, t, act, roi, arr, mlp, o, d, 3, 2, 5, C_, (lo, hi), dLdP)
    assert info["n_out_samples"] == 0
    assert np.abs(P.detach().cpu().double().numpy() - ref).max() <= 1e-5 * np.abs(ref).max()
    names = NAMES[:6]
    grads = torch.autograd.grad(P, [getattr(f, k) for k in names] + [m], torch.from_numpy(dLdP).cuda())
    for k, g in zip(names + ["mlp"], grads):
        assert _rel(g.cpu().double().numpy(), rgrads[k]) <= 1e-5, k


def test_fused_vm_bf16_decoder(oracle, native, cuda):
    """BF16-MFMA decoder over the VM features: against the oracle's
    BF16-emulating decoder (same rounding points)."""
    lo, hi = [-0.55, -0.45, -0.6], [0.5, 0.6, 0.4]
    C_, Gp, Gl, Go = 16, 9, 17, 5
    t, act, tree, roi, f, arr, mlp, o, d = _setup(native, oracle, C_, Gp, Gl, Go, lo, hi, seed=1)
    P, *_ = _gpu(native, tree, roi, f, mlp, o, d, 4, 2, 7, C_, Go, bf=True)
    ref, _, _ = _oracle(oracle, t, act, roi, arr, mlp, o, d, 4, 2, 7, C_, (lo, hi), bf=True)
    err = np.abs(P.detach().cpu().double().numpy() - ref).max() / np.abs(ref).max()
    print(f"bf16 decoder: projection rel err {err:.2e} against the emulating oracle", flush=True)
    assert err <= 1e-5


def test_fused_vm_through_ctypes(oracle, native, cuda):
    """atdf_project_fwd with an atdf_vm built in ctypes equals ProjectRaysVM
    bit for bit (the struct layout of include/atdf.h)."""
    from adaptivetensordf_amd import capi

    lo, hi = [-0.55, -0.45, -0.6], [0.5, 0.6, 0.4]
    C_, Gp, Gl, Go = 16, 9, 17, 5
    t, act, tree, roi, f, arr, mlp, o, d = _setup(native, oracle, C_, Gp, Gl, Go, lo, hi, seed=3)
    P, m, sp, fp, rays = _gpu(native, tree, roi, f, mlp, o, d, 4, 2, 9, C_, Go)
    planes = torch.stack([f.plane_yz, f.plane_xz, f.plane_xy], 1).permute(0, 1, 3, 4, 2).contiguous().detach()
    lines = torch.stack([f.line_x[0, :, :, 0], f.line_y[0, :, :, 0], f.line_z[0, :, :, 0]], 0).permute(0, 2, 1)
    lines = lines.contiguous().detach()
    grid = f.grid_out.permute(0, 2, 3, 4, 1).contiguous().detach()
    n2p = roi.prefix_inroi.to(torch.int32).contiguous()
    n2g = roi.prefix_outroi.to(torch.int32).contiguous()
    vm = capi.VM(planes.data_ptr(), lines.data_ptr(), planes.shape[0], n2p.data_ptr(), Gp, Gl, lo[0], hi[0], None,
                 None)
    fld = capi.Field(C_, 64, Go, capi.GRID_F32, 2.0, 0, grid.data_ptr(), grid.shape[0], n2g.data_ptr(), m.data_ptr(),
                     capi.DECODER_F32, 0, C.pointer(vm))
    iv = tree.CreateIntervalsForRays(rays, sp)
    civ = capi.Intervals(iv.node.data_ptr(), iv.tmin.data_ptr(), iv.tmax.data_ptr(), iv.count.data_ptr(),
                         iv.first.data_ptr(), iv.ray.data_ptr(), iv.num_intervals)
    csp = capi.Sampling(4, 256, 2, 0, 9, None, None)
    out = torch.empty(o.shape[0], device="cuda")
    packed = tree.PackedTree()
    capi.call("atdf_project_fwd", capi.ptr(packed), tree.NumNodes(), capi.ptr(rays.origin), capi.ptr(rays.direction),
              o.shape[0], C.byref(csp), capi.ptr(iv.ray_offset), C.byref(civ), C.byref(fld), capi.ptr(out),
              capi.stream_ptr())
    torch.cuda.synchronize()
    assert torch.equal(out, P.detach())
    # the training step's prepared forward: node-major over the counting sort
    # when the workspace holds the two-pass staging region (per-interval sums
    # added per ray, so within a few ulp), the ray-major forward without it
    base = capi.lib().atdf_project_bwd_workspace_bytes(o.shape[0], iv.num_intervals, C.byref(fld))
    stage = capi.lib().atdf_project_bwd_staging_bytes(iv.num_intervals, iv.num_samples, C.byref(fld))
    assert stage > 0
    for nbytes, exact in ((base + stage, False), (base, True)):
        ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        out2 = torch.full_like(out, float("nan"))
        args = (capi.ptr(packed), tree.NumNodes(), capi.ptr(rays.origin), capi.ptr(rays.direction), o.shape[0],
                C.byref(csp), capi.ptr(iv.ray_offset), C.byref(civ), C.byref(fld))
        capi.call("atdf_project_prepare", *args, capi.ptr(ws), ws.numel(), capi.stream_ptr())
        capi.call("atdf_project_fwd_prepared", *args, capi.ptr(out2), capi.ptr(ws), ws.numel(), capi.stream_ptr())
        torch.cuda.synchronize()
        if exact:
            assert torch.equal(out2, out)
        else:
            err = float(((out2 - out).abs() / out.abs().clamp_min(1e-3 * float(out.abs().max()))).max())
            print(f"node-major tensorial forward vs ray-major: {err:.2e}", flush=True)
            assert err <= 1e-5


def test_fused_vm_matches_composed(oracle, native, cuda):
    """ProjectRaysVM against ComputeImageVM on the same midpoint samples;
    they differ only in the line x coordinate (the composed path's torch
    division by a scalar multiplies by its reciprocal)."""
    lo, hi = [-0.55, -0.45, -0.6], [0.5, 0.6, 0.4]
    C_, Gp, Gl, Go = 16, 9, 17, 5
    t, act, tree, roi, f, arr, mlp, o, d = _setup(native, oracle, C_, Gp, Gl, Go, lo, hi, seed=4)
    P, m, sp, fp, rays = _gpu(native, tree, roi, f, mlp, o, d, 4, 0, 0, C_, Go)
    samples = tree.CreateSamplesForRays(rays, 4, False)
    Pc = native.ComputeImageVM(tree, roi, f, m, fp, samples, o.shape[0])[0]
    assert torch.allclose(P, Pc, rtol=0, atol=1e-5 * float(Pc.detach().abs().max()))
    rng = np.random.default_rng(8)
    go = torch.from_numpy(rng.normal(size=o.shape[0]).astype(np.float32)).cuda()
    params = [getattr(f, k) for k in NAMES] + [m]
    g1 = torch.autograd.grad(P, params, go)
    g2 = torch.autograd.grad(Pc, params, go)
    for k, a, b in zip(NAMES + ["mlp"], g1, g2):
        e = _rel(a.cpu().double().numpy(), b.cpu().double().numpy())
        print(f"fused vs composed grad {k}: {e:.2e}", flush=True)
        assert e <= 1e-4, k


def test_train_step_vm(oracle, native, cuda):
    """TrainStepVM: the loss of the fused forward, and the first Adam step
    -lr g / (|g| + eps) on every parameter entry, g the gradient of the same
    MSE through ProjectRaysVM."""
    lo, hi = [-0.55, -0.45, -0.6], [0.5, 0.6, 0.4]
    C_, Gp, Gl, Go = 16, 9, 17, 5
    t, act, tree, roi, f, arr, mlp, o, d = _setup(native, oracle, C_, Gp, Gl, Go, lo, hi, seed=5)
    P, m, sp, fp, rays = _gpu(native, tree, roi, f, mlp, o, d, 4, 2, 13, C_, Go)
    target = (P.detach() + 0.05 * torch.randn_like(P)).contiguous()
    lr = 1e-3
    st = native.MakeVMTrainState(tree, roi, f, m, lr, fp)
    loss = native.TrainStepVM(tree, roi, st, rays, target, sp, fp)
    ref_loss = ((P.detach().double() - target.double()) ** 2).mean()
    assert abs(float(loss.item()) - float(ref_loss)) <= 1e-6 * float(ref_loss)
    params = [getattr(f, k) for k in NAMES] + [m]
    gref = torch.autograd.grad(((P - target) ** 2).mean(), params)
    after = native.VMTrainStateFields(st, f)
    new = [getattr(after, k) for k in NAMES] + [st.mlp]
    for k, p0, p1, g in zip(NAMES + ["mlp"], params, new, gref):
        sel = (g.abs() > 1e-3 * g.abs().max()) & (g.abs() > 1e-7)
        assert int(sel.sum()) > 0, k
        step = (p1 - p0.detach())[sel]
        want = -lr * g[sel] / (g[sel].abs() + 1e-8)
        print(f"TrainStepVM {k}: {int(sel.sum())} entries, max |step - want| {float((step - want).abs().max()):.2e}",
              flush=True)
        assert torch.allclose(step, want, rtol=0, atol=2e-3 * lr), k
    assert all(float(gr.abs().max()) == 0.0 for gr in st.grads)  # cleared by the Adam passes


def _adjoint(native, tree, roi, f, mlp, o, d, K, seed, C_, Go, bf, two_pass, dLdP):
    P, m, *_ = _gpu(native, tree, roi, f, mlp, o, d, K, 2, seed, C_, Go, bf=bf, two_pass=two_pass)
    names = NAMES if int(roi.out_roi_node_index.numel()) > 0 else NAMES[:6]
    grads = torch.autograd.grad(P, [getattr(f, k) for k in names] + [m], torch.from_numpy(dLdP).cuda())
    return dict(zip(names + ["mlp"], [g.detach().cpu().double().numpy() for g in grads]))


@pytest.mark.parametrize("C_,bf,depth,budget,n_rays,K,lohi", [
    (16, False, 4, 400, 3000, 6, ([-0.55, -0.45, -0.6], [0.5, 0.6, 0.4])),
    (16, True, 4, 400, 3000, 6, ([-0.55, -0.45, -0.6], [0.5, 0.6, 0.4])),
    (8, False, 4, 400, 3000, 6, ([-0.55, -0.45, -0.6], [0.5, 0.6, 0.4])),
    (8, True, 4, 400, 2000, 6, ([-0.55, -0.45, -0.6], [0.5, 0.6, 0.4])),
    # few large nodes, many samples each: nodes split over several chunks
    (16, False, 2, 20, 6000, 8, ([-0.6, -0.6, -0.6], [0.6, 0.6, 0.6])),
])
def test_vm_two_pass_matches_fused(oracle, native, cuda, C_, bf, depth, budget, n_rays, K, lohi):
    """The two-pass adjoint of a tensorial field against its fused kernel on
    larger lists (FP32 and BF16-MFMA decoder, C = 16 and 8, nodes split over
    chunks): every plane, line, explicit-grid and decoder gradient within
    1e-5 (both sum in run-dependent orders)."""
    lo, hi = lohi
    Gp, Gl, Go = 11, 65, 11
    t, act, tree, roi, f, arr, mlp, o, d = _setup(native, oracle, C_, Gp, Gl, Go, lo, hi, seed=6, depth=depth,
                                                  budget=budget, n_rays=n_rays)
    dLdP = np.random.default_rng(12).normal(size=o.shape[0]).astype(np.float32)
    ga = _adjoint(native, tree, roi, f, mlp, o, d, K, 17, C_, Go, bf, True, dLdP)
    gb = _adjoint(native, tree, roi, f, mlp, o, d, K, 17, C_, Go, bf, False, dLdP)
    print(f"C={C_} bf16 decoder={bf} depth={depth}: {int(roi.in_roi_node_index.numel())} in-ROI / "
          f"{int(roi.out_roi_node_index.numel())} explicit nodes, {o.shape[0]} rays", flush=True)
    for k in ga:
        e = _rel(ga[k], gb[k])
        print(f"  two-pass vs fused grad {k}: {e:.2e}", flush=True)
        assert np.abs(gb[k]).max() > 0, k
        assert e <= 1e-5, k
