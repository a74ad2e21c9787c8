"""Tensorial (VM) feature path on the GPU against the oracle: 2-D / 1-D
indirect gathers (forward bit-exact, backward within 1e-5 of FP64),
GradientCorrection, ComputeBoundWeight, and the ROI-routed feature
composition SampleVolumeIndirectFeature ("add" and "cat") with its
autograd gradients."""
import numpy as np
import pytest
import torch

from conftest import random_rays
from scene_util import adaptive_active, device_tree

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)


def test_grid_sample2d_1d(oracle, native, cuda):
    rng = np.random.default_rng(0)
    for C_, H, W in ((16, 11, 9), (8, 5, 13)):
        grid = rng.normal(size=(30, C_, H, W)).astype(np.float32)
        S = 20000
        index = rng.integers(0, 30, S)
        uv = rng.uniform(-1.05, 1.05, (S, 2)).astype(np.float32)
        uv[:5] = [[-1, -1], [1, 1], [1, -1], [0, 0], [1.0000001, -1.0000001]]
        g_t = torch.from_numpy(grid).cuda().requires_grad_()
        u_t = torch.from_numpy(uv).cuda().requires_grad_()
        out = native.IndirectGridSample2D(g_t, torch.from_numpy(index).cuda(), u_t)
        ref = oracle.grid_sample2d_fwd(grid, index, uv)
        assert np.array_equal(out.detach().cpu().numpy().view(np.uint32), ref.view(np.uint32))
        go = rng.normal(size=(S, C_)).astype(np.float32)
        gg, gu = torch.autograd.grad(out, [g_t, u_t], torch.from_numpy(go).cuda())
        rg, ru = oracle.grid_sample2d_bwd(grid, index, uv, go)
        print(f"2-D C={C_} {H}x{W}: grad_grid {_rel(gg.cpu().numpy(), rg):.2e}, grad_uv {_rel(gu.cpu().numpy(), ru):.2e}",
              flush=True)
        assert _rel(gg.cpu().double().numpy(), rg) <= 1e-5 and _rel(gu.cpu().double().numpy(), ru) <= 1e-5
    line = rng.normal(size=(2, 16, 23, 1)).astype(np.float32)
    S = 30000
    index = rng.integers(0, 2, S)
    u = rng.uniform(-1.1, 1.1, (S, 1)).astype(np.float32)
    l_t = torch.from_numpy(line).cuda().requires_grad_()
    u_t = torch.from_numpy(u).cuda().requires_grad_()
    out = native.IndirectGridSample1D(l_t, torch.from_numpy(index).cuda(), u_t)
    ref = oracle.grid_sample1d_fwd(line, index, u)
    assert np.array_equal(out.detach().cpu().numpy().view(np.uint32), ref.view(np.uint32))
    go = rng.normal(size=(S, 16)).astype(np.float32)
    gg, gu = torch.autograd.grad(out, [l_t, u_t], torch.from_numpy(go).cuda())
    rg, ru = oracle.grid_sample1d_bwd(line, index, u, go)
    print(f"1-D: grad_grid {_rel(gg.cpu().numpy(), rg):.2e}, grad_uv {_rel(gu.cpu().numpy().ravel(), ru):.2e}",
          flush=True)
    assert _rel(gg.cpu().double().numpy(), rg) <= 1e-5 and _rel(gu.cpu().double().numpy().ravel(), ru) <= 1e-5


def test_gradient_correction_and_bound_weight(oracle, native, cuda):
    rng = np.random.default_rng(1)
    d = rng.normal(size=(100, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    ray = rng.integers(0, 100, 5000)
    pos = torch.from_numpy(rng.normal(size=(5000, 3)).astype(np.float32)).cuda().requires_grad_()
    out = native.GradientCorrection(pos, torch.from_numpy(d).cuda(), torch.from_numpy(ray).cuda())
    assert torch.equal(out, pos)
    go = rng.normal(size=(5000, 3)).astype(np.float32)
    (g,) = torch.autograd.grad(out, pos, torch.from_numpy(go).cuda())
    assert np.array_equal(g.cpu().numpy(), oracle.gradient_correction(d, ray, go))
    # bound weight against the reference's scan
    t = oracle.tree_build(2)
    tree = native.HyperTreeBase(3, 2)
    tree.settree()
    tree.SetActive(2)
    tree.to("cuda")
    roi = tree.RoiSplit([-0.5, -0.5, -0.5], [0.5, 0.9, 0.2], 0.0)
    nodes = np.nonzero(t["depth"] == 2)[0]
    node_id = rng.choice(nodes, 7000)
    bw = native.ComputeBoundWeight(torch.from_numpy(node_id).cuda(), roi).cpu().numpy()
    assert np.array_equal(bw, oracle.compute_bound_weight(node_id, roi.in_roi_node_index.cpu().numpy()))


def _vm_setup(native, oracle, op, seed=0):
    rng = np.random.default_rng(seed)
    depth = 3
    t = oracle.tree_build(depth)
    act = adaptive_active(t, budget=150, seed=seed)
    tree = device_tree(depth, act)
    lo, hi = [-0.55, -0.45, -0.6], [0.5, 0.6, 0.4]
    roi = tree.RoiSplit(lo, hi, 0.0)
    n_in = int(roi.in_roi_node_index.numel())
    n_out = int(roi.out_roi_node_index.numel())
    C_, G, Gl = 8, 9, 17
    Cg = C_ * (3 if op == "cat" else 1)
    f = native.VMFields()
    arr = {}
    for name, shape in (("plane_yz", (n_in, C_, G, G)), ("plane_xz", (n_in, C_, G, G)),
                        ("plane_xy", (n_in, C_, G, G)), ("line_x", (1, C_, Gl, 1)), ("line_y", (1, C_, Gl, 1)),
                        ("line_z", (1, C_, Gl, 1)), ("grid_out", (n_out, Cg, 5, 5, 5))):
        arr[name] = rng.uniform(-0.5, 0.5, shape).astype(np.float32)
        setattr(f, name, torch.from_numpy(arr[name]).cuda().requires_grad_())
    f.tensor_op = op
    f.roi_min, f.roi_max = lo, hi
    o, d = random_rays(300, 40 + seed, spread=0.9)
    tr = oracle.traverse(t["node_min"], t["node_max"], oracle.active_ids(act), o, d, 4)
    return t, act, tree, roi, f, arr, tr, (lo, hi)


def _vm_oracle(oracle, t, act, roi, arr, tr, lohi, op):
    """The same composition from the oracle's gathers, in the kernels' order."""
    g = tr["global_coordinate"]
    nid = tr["node_id"]
    flag = roi.in_roi_flag.cpu().numpy()[nid].astype(bool)
    pin, pout = roi.prefix_inroi.cpu().numpy(), roi.prefix_outroi.cpu().numpy()
    loc = oracle.local_samples(t["node_min"], t["node_max"], g, nid)
    lo, hi = np.float32(lohi[0][0]), np.float32(lohi[1][0])
    li = loc[flag]
    gi = g[flag]
    slot = pin[nid[flag]]
    pl = {"plane_yz": li[:, [1, 2]], "plane_xz": li[:, [0, 2]], "plane_xy": li[:, [0, 1]]}
    ln = {"line_x": np.float32(2.0) * (gi[:, 0] - lo) / (hi - lo) - np.float32(1.0), "line_y": li[:, 1],
          "line_z": gi[:, 2]}
    zero = np.zeros(li.shape[0], np.int64)
    P = {k: oracle.grid_sample2d_fwd(arr[k], slot, v) for k, v in pl.items()}
    L = {k: oracle.grid_sample1d_fwd(arr[k], zero, v.astype(np.float32)) for k, v in ln.items()}
    pairs = (("plane_yz", "line_x"), ("plane_xz", "line_y"), ("plane_xy", "line_z"))
    if op == "add":
        fin = P["plane_yz"] * L["line_x"]
        fin = fin + P["plane_xz"] * L["line_y"]
        fin = fin + P["plane_xy"] * L["line_z"]
    else:
        fin = np.concatenate([P[a] for a, _ in pairs], 1) * np.concatenate([L[b] for _, b in pairs], 1)
    out_slot = pout[nid[~flag]]
    grid_ncz = arr["grid_out"]
    fout = oracle.grid_sample3d_fwd(grid_ncz, out_slot, loc[~flag])
    feats = np.zeros((g.shape[0], fin.shape[1]), np.float32)
    feats[flag] = fin
    feats[~flag] = fout
    return feats, dict(flag=flag, slot=slot, out_slot=out_slot, loc=loc, pl=pl, ln=ln, P=P, L=L, pairs=pairs)


@pytest.mark.parametrize("op", ["add", "cat"])
def test_sample_volume_indirect_feature(oracle, native, cuda, op):
    t, act, tree, roi, f, arr, tr, lohi = _vm_setup(native, oracle, op)
    g = torch.from_numpy(tr["global_coordinate"]).cuda()
    nid = torch.from_numpy(tr["node_id"]).cuda()
    feats = native.SampleVolumeIndirectFeature(tree, roi, f, g, nid)
    ref, aux = _vm_oracle(oracle, t, act, roi, arr, tr, lohi, op)
    flag = aux["flag"]
    print(f"{op}: {int(flag.sum())} in-ROI / {int((~flag).sum())} out-of-ROI samples", flush=True)
    assert 0 < flag.sum() < flag.size
    assert np.abs(feats.detach().cpu().numpy() - ref).max() <= 1e-6 * np.abs(ref).max()
    # gradients of every field against the FP64 oracle backward of each part
    rng = np.random.default_rng(5)
    go = rng.normal(size=ref.shape).astype(np.float32)
    names = ["plane_yz", "plane_xz", "plane_xy", "line_x", "line_y", "line_z", "grid_out"]
    grads = torch.autograd.grad(feats, [getattr(f, k) for k in names], torch.from_numpy(go).cuda())
    got = dict(zip(names, [x.cpu().double().numpy() for x in grads]))
    gin = go[flag]
    C_ = arr["line_x"].shape[1]
    zero = np.zeros(gin.shape[0], np.int64)
    for i, (pk, lk) in enumerate(aux["pairs"]):
        gi = gin if op == "add" else gin[:, i * C_:(i + 1) * C_]
        dP = (gi * aux["L"][lk]).astype(np.float32)
        dL = (gi * aux["P"][pk]).astype(np.float32)
        rp, _ = oracle.grid_sample2d_bwd(arr[pk], aux["slot"], aux["pl"][pk], dP)
        rl, _ = oracle.grid_sample1d_bwd(arr[lk], zero, aux["ln"][lk].astype(np.float32), dL)
        assert _rel(got[pk], rp) <= 1e-5, pk
        assert _rel(got[lk], rl) <= 1e-5, lk
    rgo, _ = oracle.grid_sample3d_bwd(arr["grid_out"], aux["out_slot"], aux["loc"][~flag], go[~flag])
    assert _rel(got["grid_out"], rgo) <= 1e-5


@pytest.mark.parametrize("shape,cl", [((3, 8, 33), False), ((40, 8, 9, 9), False), ((30, 8, 5, 5, 5), False),
                                      ((500, 11, 11, 11, 16), True)])
def test_tv_loss_matches_oracle(oracle, cuda, shape, cl):
    import ctypes as C

    from adaptivetensordf_amd import capi

    rng = np.random.default_rng(sum(shape))
    x = rng.normal(size=shape).astype(np.float32)
    x[..., 0] = x[..., 1]  # ties
    X = torch.from_numpy(x).to(cuda)
    nd = len(shape) - 2
    if cl:
        n, st = shape[1:-1], [s // 4 for s in x.strides]
        sh = capi.TVShape(shape[0], shape[-1], nd, (C.c_int32 * 3)(*(list(n) + [1] * (3 - nd))), st[0], 1,
                          (C.c_int64 * 3)(*(st[1:-1] + [0] * (3 - nd))))
    else:
        n, st = shape[2:], [s // 4 for s in x.strides]
        sh = capi.TVShape(shape[0], shape[1], nd, (C.c_int32 * 3)(*(list(n) + [1] * (3 - nd))), st[0], st[1],
                          (C.c_int64 * 3)(*(st[2:] + [0] * (3 - nd))))
    for weight in (None, rng.uniform(0.5, 2.0, size=shape[0]).astype(np.float32)):
        W = None if weight is None else torch.from_numpy(weight).to(cuda)
        g0 = rng.normal(size=shape).astype(np.float32)
        G = torch.from_numpy(g0).to(cuda)
        loss = torch.zeros(1, dtype=torch.float64, device=cuda)
        ws = torch.empty(capi.lib().atdf_tv_workspace_bytes(x.size), dtype=torch.uint8, device=cuda)
        print(f"tv {shape} cl={cl} weight={weight is not None}", flush=True)
        capi.call("atdf_tv_loss", capi.ptr(X), C.byref(sh), capi.ptr(W), 0.25, capi.ptr(loss), capi.ptr(G),
                  capi.ptr(ws), ws.numel(), capi.stream_ptr())
        want, wg = oracle.tv_loss(x, channels_last=cl, weight=weight, scale=0.25, grad=g0)
        assert abs(loss.item() - want) <= 1e-12 * abs(want)
        assert np.array_equal(G.cpu().numpy(), wg)


def test_tv_loss_linear_field_full_size(native, cuda):
    """Explicit grids at the bench's node count (8191 x 11^3 x 16,
    channels-last): a linear field's loss is C * sum_d |a_d| whatever the
    size, and the gradient vanishes inside every node."""
    N, G, C_ = 8191, 11, 16
    a = (0.5, -0.25, 0.125)
    s = torch.arange(G, dtype=torch.float32, device=cuda)
    field = a[0] * s.view(G, 1, 1) + a[1] * s.view(1, G, 1) + a[2] * s.view(1, 1, G)
    x = (field.view(1, G, G, G, 1) + torch.arange(C_, device=cuda).view(1, 1, 1, 1, C_)).expand(N, -1, -1, -1, -1)
    x = x.contiguous().requires_grad_()
    loss = native.TVLoss(x, channels_last=True)
    assert abs(loss.item() - C_ * sum(abs(v) for v in a)) <= 1e-6
    loss.backward()
    assert torch.all(x.grad[:, 1:-1, 1:-1, 1:-1, :] == 0) and torch.any(x.grad != 0)


def _torch_tv(t):
    nc, D = t.shape[1], t.dim() - 2
    total = 0
    for i in range(D):
        d = i + 2
        n = t.shape[d]
        total = total + (t.narrow(d, 0, n - 1) - t.narrow(d, 1, n - 1)).abs().mean()
    return total * nc


@pytest.mark.parametrize("tv3d", [True, False])
def test_volume_regularizer(oracle, native, cuda, tv3d):
    t, act, tree, roi, f, arr, tr, lohi = _vm_setup(native, oracle, "add", seed=3)
    p = native.RegularizerParams()
    p.loss_tv, p.use_loss_tv_3d, p.tv_size = 0.01, tv3d, [7, 13, 7]
    p.plane_vec_only, p.test_para, p.loss_l2 = False, 1.0, 0.02
    p.test_line_x, p.test_line_y, p.test_line_z = 1.0, 0.5, 2.0
    names = ["plane_yz", "plane_xz", "plane_xy", "line_x", "line_y", "line_z", "grid_out"]
    fields = [getattr(f, k) for k in names]
    loss = native.VolumeRegularizer(tree, roi, f, p)
    got = torch.autograd.grad(loss, fields, allow_unused=True)
    # the same terms written with libtorch ops
    if tv3d:
        ids = roi.in_roi_node_index
        N = ids.numel()
        lin = [torch.linspace(-1, 1, k, device=cuda) for k in p.tv_size]
        mesh = torch.meshgrid(*lin, indexing="ij")
        cell = torch.stack(mesh, -1)[None]
        mn = tree.node_position_min.index_select(0, ids).view(N, 1, 1, 1, 3)
        sz = tree.node_position_max.index_select(0, ids).view(N, 1, 1, 1, 3) - mn
        pos = ((cell + 1) / 2 * sz + mn).reshape(-1, 3)
        nid = ids.view(N, 1, 1, 1).expand(N, *p.tv_size).reshape(-1)
        feat = native.SampleVolumeIndirectFeature(tree, roi, f, pos, nid).reshape(N, *p.tv_size, -1)
        ref = _torch_tv(feat.movedim(-1, 1)) * p.loss_tv
    else:
        ref = sum(_torch_tv(getattr(f, k)) * p.loss_tv for k in ("plane_xz", "plane_yz", "plane_xy"))
        for k, w in (("line_x", 1.0), ("line_y", 0.5), ("line_z", 2.0)):
            ref = ref + _torch_tv(getattr(f, k).squeeze(-1)) * (p.loss_tv * w)
        ref = ref + _torch_tv(f.grid_out) * p.loss_tv
    for k in ("plane_xz", "plane_yz", "plane_xy", "line_x", "line_y", "line_z"):
        ref = ref + getattr(f, k).square().mean() * p.loss_l2
    ref = ref + f.grid_out.square().mean()
    want = torch.autograd.grad(ref, fields, allow_unused=True)
    print(f"tv3d={tv3d}: loss {loss.item():.6e} torch {ref.item():.6e}", flush=True)
    assert abs(loss.item() - ref.item()) <= 1e-5 * abs(ref.item())
    for k, a, b in zip(names, got, want):
        if b is None:
            assert a is None or torch.all(a == 0), k
            continue
        assert _rel(a.cpu().double().numpy(), b.cpu().double().numpy()) <= 1e-4, k


@pytest.mark.parametrize("use_roi", [False, True])
@pytest.mark.parametrize("quad", [False, True])
def test_node_neighbor_samples_bit_exact(oracle, native, cuda, use_roi, quad):
    for depth, budget in ((3, 150), (4, 500)):
        t = oracle.tree_build(depth)
        act = adaptive_active(t, budget=budget, seed=depth)
        tree = device_tree(depth, act)
        # the quad variant's four-child descent only reaches the z-lowest
        # nodes of an octree, so its ROI reaches down to z = -1
        roi = ((-0.6, -0.7, -1.2 if quad else -0.5), (0.55, 0.65, 0.6)) if use_roi else None
        lo, hi = (list(roi[0]), list(roi[1])) if roi else ([], [])
        if quad:
            s = tree.NodeNeighborSamplesPlane2D([4, 3, 5], 1e-3, 0, lo, hi, use_roi, False)
        else:
            s = tree.NodeNeighborSamples([4, 3, 5], 1e-3, lo, hi, use_roi)
        pos, node = oracle.neighbor_pairs(t["node_min"], t["node_max"], t["children"], act, (4, 3, 5), eps=1e-3,
                                          roi=roi, quad=quad)
        print(f"depth {depth} roi={use_roi} quad={quad}: {pos.shape[0] // 2} pairs", flush=True)
        assert pos.shape[0] > 0
        assert np.array_equal(s.global_coordinate.cpu().numpy(), pos)
        assert np.array_equal(s.node_id.cpu().numpy(), node)
        assert np.array_equal(s.local_index_in_ray.cpu().numpy(), np.arange(pos.shape[0]) % 2)
        if quad:
            # use_quad: NodeNeighborSamples2D, which the reference runs on
            # whatever tree it is given (here an octree, four-child descent)
            s2 = tree.NodeNeighborSamplesPlane2D([4, 3, 5], 1e-3, 1, lo, hi, use_roi, True)
            pos2, node2 = oracle.neighbor_pairs(t["node_min"], t["node_max"], t["children"], act, (4, 3, 5),
                                                eps=1e-3, roi=roi, plane=1)
            assert np.array_equal(s2.global_coordinate.cpu().numpy(), pos2)
            assert np.array_equal(s2.node_id.cpu().numpy(), node2)


def test_edge_loss_and_regularizer(oracle, native, cuda):
    t, act, tree, roi, f, arr, tr, lohi = _vm_setup(native, oracle, "add", seed=5)
    pairs = tree.NodeNeighborSamples([5, 5, 5], 1e-3)
    p = native.RegularizerParams()
    p.loss_tv, p.loss_l2, p.loss_edge = 0.0, -1.0, 0.05
    names = ["plane_yz", "plane_xz", "plane_xy", "line_x", "line_y", "line_z", "grid_out"]
    fields = [getattr(f, k) for k in names]
    loss = native.VolumeRegularizer(tree, roi, f, p, pairs)
    got = torch.autograd.grad(loss, fields, allow_unused=True)
    # the same with libtorch ops: features of both samples, per-group mean of
    # sqrt(mean_c (f1 - f2)^2) * loss_edge * C
    feat = native.SampleVolumeIndirectFeature(tree, roi, f, pairs.global_coordinate, pairs.node_id)
    e = (feat[0::2] - feat[1::2]).pow(2).mean(1).sqrt()
    inr = native.ComputeBoundWeight(pairs.node_id[0::2].contiguous(), roi) >= 0
    print(f"{int(inr.sum())} in-ROI pairs, {int((~inr).sum())} outside", flush=True)
    assert 0 < int(inr.sum()) < inr.numel()
    ref = e[inr].mean() * (0.05 * feat.shape[1]) + e[~inr].mean() * (0.05 * feat.shape[1])
    want = torch.autograd.grad(ref, fields, allow_unused=True)
    assert abs(loss.item() - ref.item()) <= 1e-6 * abs(ref.item())
    for k, a, b in zip(names, got, want):
        if b is None:
            assert a is None or torch.all(a == 0), k
            continue
        assert _rel(a.cpu().double().numpy(), b.cpu().double().numpy()) <= 1e-5, k
    # without pairs (or with the term off) there is no edge term
    assert native.VolumeRegularizer(tree, roi, f, p) is None


@pytest.mark.parametrize("D,N,shape,S", [(1, 1, (193,), 100_000), (2, 9000, (11, 11), 60_000),
                                         (3, 40, (11, 11, 11), 50_000), (3, 9000, (5, 5, 5), 30_000)])
def test_grid_sample_bwd_slots(oracle, cuda, D, N, shape, S):
    """The slot-major backward through the raw ABI: one line taking every
    sample (25 chunks of one slot), more slots than an LDS histogram holds
    (global-atomic sort), multi-chunk 3-D slots of 85 KB; against the FP64
    oracle backward."""
    import ctypes as C
    from adaptivetensordf_amd import capi

    lib = capi.lib()
    rng = np.random.default_rng(D * 100 + N)
    Cc = 16
    grid = rng.normal(size=(N, Cc) + shape).astype(np.float32)
    index = np.sort(rng.integers(0, N, S)) if D == 3 and N == 40 else rng.integers(0, N, S)
    uv = rng.uniform(-1.05, 1.05, (S, D)).astype(np.float32)
    # keep coordinates off the lattice points, where the derivative of the
    # interpolation jumps and float32 and FP64 can take different sides
    while True:
        bad = np.zeros(S, bool)
        for d in range(D):
            n = np.float32(shape[D - 1 - d] - 1)
            ix = ((uv[:, d] + np.float32(1)) / np.float32(2)) * n
            bad |= np.abs(ix - np.round(ix)) < 1e-3
        if not bad.any():
            break
        uv[bad] = rng.uniform(-1.05, 1.05, (int(bad.sum()), D)).astype(np.float32)
    go = rng.normal(size=(S, Cc)).astype(np.float32)
    print(f"slots D={D} N={N} shape={shape} S={S}", flush=True)
    t = {k: torch.from_numpy(v).to(cuda) for k, v in (("grid", grid), ("index", index), ("uv", uv), ("go", go))}
    gg = torch.zeros_like(t["grid"])
    gu = torch.zeros_like(t["uv"])
    ws = torch.empty(lib.atdf_grid_sample_bwd_workspace_bytes(N, S), dtype=torch.uint8, device=cuda)
    size = (C.c_int32 * 3)(*(list(shape) + [1] * (3 - D)))
    assert lib.atdf_grid_sample_bwd_max_slot_floats() >= Cc * int(np.prod(shape))
    st = lib.atdf_grid_sample_bwd_slots(D, t["grid"].data_ptr(), N, Cc, size, t["index"].data_ptr(), t["uv"].data_ptr(),
                                        S, t["go"].data_ptr(), gg.data_ptr(), gu.data_ptr(), ws.data_ptr(), ws.numel(),
                                        None)
    assert st == 0, lib.atdf_last_error()
    torch.cuda.synchronize()
    if D == 3:
        rg, ru = oracle.grid_sample3d_bwd(grid, index, uv, go)
    elif D == 2:
        rg, ru = oracle.grid_sample2d_bwd(grid, index, uv, go)
    else:
        rg, ru = oracle.grid_sample1d_bwd(grid[..., None], index, uv, go)
        rg = rg.reshape(grid.shape)
    ru = np.asarray(ru).reshape(uv.shape)
    print(f"  grad_grid {_rel(gg.cpu().double().numpy(), rg):.2e} grad_uv {_rel(gu.cpu().double().numpy(), ru):.2e}",
          flush=True)
    assert _rel(gg.cpu().double().numpy(), rg) <= 1e-5
    assert _rel(gu.cpu().double().numpy(), ru) <= 1e-5
