"""Parity at the benchmark configurations' own shapes (the 512^3 workload:
depth-5 adaptive octree with 8191 active nodes, 61 tilts on a 512^2
detector, K = 12 for ~128 stratified samples per ray, C = 16, 11^3 grids).

  c1  plain FP32: on a 4096-ray subset, intervals bit-exact, projections
      within 1e-5 of the FP32 and FP64 oracles, grid and decoder gradients
      within 1e-5 of the FP64 oracle adjoint; on a full 2^18-ray batch,
      size-independent properties (interval weights telescope to the chord,
      samples inside their nodes, determinism, node-major == ray-major
      forward, a constant field's closed form, linearity and ray-additivity
      of the adjoint).
  c2  the hand-written adjoint against central finite differences
      (h = 1e-3, FP64 oracle forward) on 1000 random touched grid entries.
  c3  BF16 grids + BF16-MFMA decoder after GPU training: the field on the
      full 256^3 crop of the 512^3 lattice and projections of rays through
      that crop, against the FP64 oracle (f32 master grid: the stated
      precision cost) and the BF16-emulating FP64 oracle (parity).
  c4  the learned proposal trained on the GPU at the 512^2 and 768^2
      detectors: placement bit-exact, the mixed interval list bit-exact,
      projections over it within 1e-5 (FP32) of the FP64 oracle, and on a
      full batch the fallback rays' intervals equal the stratified ones.

Gradient checks leave out rays whose decoder pre-activations come within
3e-6 of a ReLU kink, where FP32 and FP64 can take different masks
(|preact| above that bound exceeds the FP32 rounding of the dot products).
"""
import numpy as np
import pytest
import torch

import workload as W
from capi_util import Fused
from scene_util import device_tree, rays_to_device

pytestmark = pytest.mark.gpu
KINK = 3e-6
JS = 7  # jitter seed (hash jitter, mode 2)


def _rel(P, ref):
    scale = np.maximum(np.abs(ref), 1e-3 * np.abs(ref).max())
    return float(np.max(np.abs(P - ref) / scale))


def _maxrel(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).max() / np.abs(b).max())


@pytest.fixture(scope="module")
def c1(oracle, cuda):
    from adaptivetensordf_amd import phantom

    t, act, ell = W.tree_arrays(oracle)
    tree = device_tree(W.DEPTH, act)
    na = int(act.sum())
    grid = phantom.init_grid(na, W.G, W.C, std=0.3, seed=0)
    mlp = phantom.init_mlp(W.C, W.H, seed=0)
    o, d = W.rays(4096, 5)
    fo, fd = W.rays(1 << 18, 100)
    fm = oracle.FusedModel(t["node_min"], t["node_max"], act, grid.numpy(), mlp.numpy(), W.C, K=W.K, jitter_mode=2,
                           seed=JS)
    print(f"workload tree: {na} active nodes of {t['depth'].size}", flush=True)
    return dict(t=t, act=act, ell=ell, tree=tree, grid=grid, mlp=mlp, o=o, d=d, fo=fo, fd=fd, fm=fm)


def _fused(c, o, d, grid=None, mlp=None, bf16_decoder=False, intervals=None, jitter_mode=2):
    o_t, d_t = rays_to_device(o, d)
    g = (c["grid"] if grid is None else grid).cuda()
    m = (c["mlp"] if mlp is None else mlp).cuda()
    return Fused(c["tree"], o_t, d_t, W.K, g, m, W.C, jitter_mode=jitter_mode, jitter_seed=JS,
                 bf16_decoder=bf16_decoder, intervals=intervals)


def test_c1_intervals_subset_bit_exact(oracle, c1):
    fu = _fused(c1, c1["o"], c1["d"])
    iv = fu.iv
    ref = oracle.traverse(c1["t"]["node_min"], c1["t"]["node_max"], oracle.active_ids(c1["act"]), c1["o"], c1["d"],
                          W.K, jitter_mode=2, seed=JS)
    spr = iv.num_samples / fu.R
    print(f"c1 subset: {iv.num_intervals} intervals, {spr:.1f} samples/ray", flush=True)
    assert 120 < spr < 136
    assert np.array_equal(iv.ray_offset.cpu().numpy(), ref["ray_interval_offset"])
    assert np.array_equal(iv.node.cpu().numpy(), ref["iv_node"])
    assert np.array_equal(iv.count.cpu().numpy(), ref["iv_count"])
    assert np.array_equal(iv.first.cpu().numpy(), ref["iv_first"])
    assert np.array_equal(iv.tmin.cpu().numpy().view(np.uint32), ref["iv_tmin"].view(np.uint32))
    assert np.array_equal(iv.tmax.cpu().numpy().view(np.uint32), ref["iv_tmax"].view(np.uint32))


def test_c1_samples_full_batch_invariants(oracle, native, c1):
    """2^18 rays, ~33 M samples: per ray the midpoint weights telescope to
    the chord through the root box (the active nodes tile it), every sample
    lies in its node's box, and the traversal is deterministic."""
    o_t, d_t = rays_to_device(c1["fo"], c1["fd"])
    rays = native.RayList(o_t, d_t)
    s1 = c1["tree"].CreateSamplesForRays(rays, W.K, False)
    s2 = c1["tree"].CreateSamplesForRays(rays, W.K, False)
    S = s1.size()
    print(f"c1 full batch: {S} samples ({S / o_t.shape[0]:.1f} per ray)", flush=True)
    assert torch.equal(s1.node_id, s2.node_id) and torch.equal(s1.ray_t, s2.ray_t)
    R = o_t.shape[0]
    wsum = torch.zeros(R, dtype=torch.float64, device=o_t.device).index_add_(0, s1.ray_index, s1.weight.double())
    chord = W.chord_through_root(c1["fo"], c1["fd"])
    wsum = wsum.cpu().numpy()
    err = np.abs(wsum - chord).max() / chord.max()
    print(f"  sum of weights vs FP64 chord: max err {err:.2e} (relative to the longest chord)", flush=True)
    # float rounding of ~60 interval ends, and intervals shorter than 1e-5
    # that the traversal drops (HyperTree.cu:334)
    assert err <= 2e-5
    nmin = c1["tree"].node_position_min[s1.node_id]
    nmax = c1["tree"].node_position_max[s1.node_id]
    pos = s1.global_coordinate
    eps = 2e-6
    assert bool(((pos >= nmin - eps) & (pos <= nmax + eps)).all())


def test_c1_projection_subset(oracle, c1):
    fu = _fused(c1, c1["o"], c1["d"])
    P = fu.forward().cpu().double().numpy()
    P32 = c1["fm"].project(c1["o"], c1["d"], precision=0)
    P64 = c1["fm"].project(c1["o"], c1["d"], precision=1)
    e32, e64 = _rel(P, P32), _rel(P, P64)
    print(f"c1 projection (4096 rays): max rel err vs FP32 oracle {e32:.2e}, vs FP64 oracle {e64:.2e}", flush=True)
    assert e32 <= 1e-5 and e64 <= 1e-5


def test_c1_projection_full_batch_invariants(c1):
    """Full 2^18-ray batch: the forward is deterministic, the node-major
    forward (counting sort shared with the adjoint) equals the ray-major one
    to the rounding of the per-interval sums, and a constant density field
    (decoder weights 0, output bias b) projects to softplus_2(b) x chord."""
    import ctypes as C

    from adaptivetensordf_amd import capi

    fu = _fused(c1, c1["fo"], c1["fd"])
    P1 = fu.forward()
    P2 = fu.forward()
    assert torch.equal(P1, P2)
    ws = fu.workspace()
    Pn = torch.full_like(P1, float("nan"))
    capi.call("atdf_project_prepare", *fu.args(), capi.ptr(ws), ws.numel(), capi.stream_ptr())
    capi.call("atdf_project_fwd_prepared", *fu.args(), capi.ptr(Pn), capi.ptr(ws), ws.numel(), capi.stream_ptr())
    e_nm = ((Pn - P1).abs().max() / P1.abs().max()).item()
    mlp_c = torch.zeros_like(c1["mlp"])
    b = 0.7
    mlp_c[-1] = b
    fc = _fused(c1, c1["fo"], c1["fd"], mlp=mlp_c, intervals=fu.iv)
    Pc = fc.forward().cpu().double().numpy()
    want = np.log1p(np.exp(2 * b)) / 2 * W.chord_through_root(c1["fo"], c1["fd"])
    e_c = np.abs(Pc - want).max() / want.max()
    print(f"c1 full batch: node-major vs ray-major {e_nm:.2e}; constant field vs closed form {e_c:.2e}", flush=True)
    assert e_nm <= 2e-6 and e_c <= 2e-5


def _clear_weights(c, o, d, seed, fm=None):
    clear = (c["fm"] if fm is None else fm).min_preact(o, d) >= KINK
    w = np.random.default_rng(seed).normal(size=o.shape[0]) * clear
    return w, clear


def test_c1_adjoint_subset(oracle, c1):
    fu = _fused(c1, c1["o"], c1["d"])
    w, clear = _clear_weights(c1, c1["o"], c1["d"], 3)
    gg, gm = fu.backward(torch.from_numpy(w).float().cuda())
    ref_g, ref_m = c1["fm"].project_bwd(c1["o"], c1["d"], w)
    eg = _maxrel(gg.cpu().numpy(), ref_g)
    em = _maxrel(gm.cpu().numpy(), ref_m)
    touched = np.count_nonzero(np.abs(ref_g).reshape(ref_g.shape[0], -1).max(1))
    print(f"c1 adjoint ({int(clear.sum())} of {clear.size} rays clear of kinks, {touched} nodes touched): "
          f"grid {eg:.2e}, decoder {em:.2e}", flush=True)
    assert clear.mean() > 0.8
    assert eg <= 1e-5 and em <= 1e-5


def test_c1_adjoint_full_batch_linearity(c1):
    """Full batch: the adjoint is linear in dL/dP and additive over a split
    of the rays (to FP32 accumulation-order rounding)."""
    fu = _fused(c1, c1["fo"], c1["fd"])
    R = fu.R
    gen = torch.Generator(device="cuda").manual_seed(0)
    u = torch.randn(R, device="cuda", generator=gen) * 1e-3
    v = torch.randn(R, device="cuda", generator=gen) * 1e-3
    gu, mu = fu.backward(u)
    gv, mv = fu.backward(v)
    guv, muv = fu.backward(2 * u - 3 * v)
    lin_g = ((guv - (2 * gu - 3 * gv)).abs().max() / guv.abs().max()).item()
    lin_m = ((muv - (2 * mu - 3 * mv)).abs().max() / muv.abs().max()).item()
    half = torch.arange(R, device="cuda") < R // 2
    ga, ma = fu.backward(u * half)
    gb, mb = fu.backward(u * ~half)
    add_g = ((gu - (ga + gb)).abs().max() / gu.abs().max()).item()
    add_m = ((mu - (ma + mb)).abs().max() / mu.abs().max()).item()
    print(f"c1 adjoint full batch: linearity grid {lin_g:.2e} decoder {lin_m:.2e}; additivity grid {add_g:.2e} "
          f"decoder {add_m:.2e}", flush=True)
    assert max(lin_g, add_g) <= 1e-5 and max(lin_m, add_m) <= 1e-5


def test_c2_finite_differences(oracle, c1):
    """Central differences (h = 1e-3) of the FP64 oracle forward on 1000
    random touched entries of the 8191 x 11^3 x 16 grid; each difference is
    taken over the rays that cross the entry's node (the others cancel).
    Midpoint samples: hash jitter depends on a ray's index in its batch,
    and each difference projects a sub-batch."""
    o, d = c1["o"], c1["d"]
    t = c1["t"]
    grid = c1["grid"].numpy().copy()
    fm = oracle.FusedModel(t["node_min"], t["node_max"], c1["act"], grid, c1["mlp"].numpy(), W.C, K=W.K)
    fu = _fused(c1, o, d, jitter_mode=0)
    w, _ = _clear_weights(c1, o, d, 4, fm)
    gg, _ = fu.backward(torch.from_numpy(w).float().cuda())
    gg = gg.cpu().double().numpy()
    ref_g, _ = fm.project_bwd(o, d, w)
    assert _maxrel(gg, ref_g) <= 1e-5
    # rays crossing each node
    tr = oracle.traverse(t["node_min"], t["node_max"], oracle.active_ids(c1["act"]), o, d, W.K)
    ray_of_iv = np.repeat(np.arange(o.shape[0]), np.diff(tr["ray_interval_offset"]))
    slot_of_iv = oracle.prefix_sum(c1["act"])[tr["iv_node"]]
    flat_g = gg.reshape(-1)
    touched = np.nonzero(flat_g)[0]
    rng = np.random.default_rng(1)
    idx = rng.choice(touched, 1000, replace=False)
    per_slot = W.G ** 3 * W.C
    flat = fm.grid.reshape(-1)
    h = np.float32(1e-3)
    errs = []
    for i in idx:
        slot = i // per_slot
        rs = np.unique(ray_of_iv[slot_of_iv == slot])
        ro, rd, rw = o[rs], d[rs], w[rs]
        keep = flat[i]
        flat[i] = keep + h
        gp = float(flat[i])
        pp = np.dot(rw, fm.project(ro, rd, 1))
        flat[i] = keep - h
        gn = float(flat[i])
        pn = np.dot(rw, fm.project(ro, rd, 1))
        flat[i] = keep
        fdv = (pp - pn) / (gp - gn)
        errs.append(abs(fdv - flat_g[i]) / max(abs(fdv), 1e-3 * np.abs(flat_g).max()))
    errs = np.array(errs)
    print(f"c2 FD on 1000 entries: median rel err {np.median(errs):.2e}, 90% {np.quantile(errs, 0.9):.2e}, "
          f"max {errs.max():.2e}", flush=True)
    assert np.median(errs) < 1e-5 and np.quantile(errs, 0.9) < 1e-4


@pytest.fixture(scope="module")
def trained_bf16(oracle, native, c1):
    """A BF16 field (FP32 master) trained 60 steps on the GPU on 2^18-ray
    batches of the analytic phantom projections."""
    from adaptivetensordf_amd import phantom

    dev = torch.device("cuda:0")
    na = int(c1["act"].sum())
    st = native.MakeTrainState(phantom.init_grid(na, W.G, W.C, std=0.01, seed=0, device=dev),
                               phantom.init_mlp(W.C, W.H, seed=0).to(dev), True, 1e-3, 1e-3)
    fp = native.FieldParams()
    fp.C, fp.G, fp.H = W.C, W.G, W.H
    fp.bf16_decoder = True
    sp = native.SamplingParams()
    sp.max_samples_per_node = W.K
    sp.jitter_mode = 2
    batches = []
    for b in range(2):
        o, d = W.rays(1 << 18, 200 + b)
        o_t, d_t = rays_to_device(o, d)
        batches.append((native.RayList(o_t, d_t), phantom.analytic_projection_torch(o_t, d_t, c1["ell"]).float()))
    losses = []
    for i in range(60):
        sp.jitter_seed = i
        losses.append(native.TrainStep(c1["tree"], st, batches[i % 2][0], batches[i % 2][1], sp, fp).item())
    print(f"BF16 field trained 60 steps: loss {losses[0]:.4e} -> {losses[-1]:.4e}", flush=True)
    assert losses[-1] < 0.5 * losses[0]
    return dict(st=st, fp=fp, sp=sp, batches=batches)


def test_c3_bf16_crop_volume_and_projection(oracle, native, c1, trained_bf16):
    st, fp = trained_bf16["st"], trained_bf16["fp"]
    t, act = c1["t"], c1["act"]
    n, lo = 256, 128
    step = 2.0 / 511.0
    roi = [float(np.float32(-1.0 + lo * step))] * 3
    vol = native.SampleVolumeBySlice(c1["tree"], st.grid, st.mlp, fp, [n, n, n], roi, [step] * 3)
    vol = vol.cpu().numpy().reshape(-1).astype(np.float64)
    pos = oracle.volume_points(1, (n, n, n), roi, [step] * 3)
    node, mask = oracle.node_id_for_position(t["node_min"], t["node_max"], t["children"], act, pos)
    assert mask.min() == 1.0
    master = st.grid_master.cpu().numpy()
    gb = st.grid.float().cpu().numpy()
    mlp = st.mlp.cpu().numpy()
    f64 = oracle.FusedModel(t["node_min"], t["node_max"], act, master, mlp, W.C)
    femu = oracle.FusedModel(t["node_min"], t["node_max"], act, gb, mlp, W.C, decoder_bf16=True)
    d64 = oracle.field_density(f64, pos, node, precision=1)
    # the emulating oracle (parity) on every 8th lattice point
    demu = oracle.field_density(femu, pos[::8], node[::8], precision=1)
    scale = np.abs(d64).max()
    e64 = np.abs(vol - d64) / scale
    eemu = np.abs(vol[::8] - demu) / scale
    rms = np.sqrt(np.mean((vol - d64) ** 2) / np.mean(d64 ** 2))
    print(f"c3 256^3 crop ({vol.size} points): BF16 field vs FP64(f32 master) max {e64.max():.2e} rms {rms:.2e}; "
          f"vs BF16-emulating FP64 ({demu.size} points) max {eemu.max():.2e} q99.9 {np.quantile(eemu, 0.999):.2e}",
          flush=True)
    # a single bf16 rounding flip moves a point by up to ~1e-3 of the field's
    # range; a misplaced rounding would move most points
    assert eemu.max() <= 4e-3 and np.quantile(eemu, 0.999) <= 1e-5
    assert e64.max() <= 3e-2 and rms <= 1e-2
    # projections of rays through the crop (the centre 256^2 of the detector)
    o, d = W.window_rays(4096, 11)
    fb = _fused(c1, o, d, grid=st.grid, mlp=st.mlp, bf16_decoder=True)
    P = fb.forward().cpu().double().numpy()
    pm64 = oracle.FusedModel(t["node_min"], t["node_max"], act, master, mlp, W.C, K=W.K, jitter_mode=2, seed=JS)
    pmemu = oracle.FusedModel(t["node_min"], t["node_max"], act, gb, mlp, W.C, K=W.K, jitter_mode=2, seed=JS,
                              decoder_bf16=True)
    ep64 = _rel(P, pm64.project(o, d, precision=1))
    epemu = _rel(P, pmemu.project(o, d, precision=1))
    print(f"c3 projections through the crop: BF16 vs FP64(f32 master) {ep64:.2e}, vs BF16-emulating {epemu:.2e}",
          flush=True)
    assert epemu <= 2e-4 and ep64 <= 2e-2


@pytest.mark.parametrize("det", [512, 768])
def test_c4_learned_at_workload_shape(oracle, native, c1, trained_bf16, det):
    from adaptivetensordf_amd import phantom
    from test_gpu_proposal import _oracle_mixed_list

    st, fp, sp = trained_bf16["st"], trained_bf16["fp"], trained_bf16["sp"]
    t, act, tree = c1["t"], c1["act"], c1["tree"]
    pp = native.ProposalParams()
    pp.mix = 0.3
    pp.kl_threshold = 0.3
    ps = native.MakeProposalState(phantom.init_proposal(0).cuda(), 3e-3, pp)
    o_tr, d_tr = W.rays(1 << 15, 300, det)
    train_rays = native.RayList(*rays_to_device(o_tr, d_tr))
    sp.jitter_seed = 5
    loss = [native.ProposalTrain(tree, ps, train_rays, sp, st.grid, st.mlp, fp, 1)]
    loss.append(native.ProposalTrain(tree, ps, train_rays, sp, st.grid, st.mlp, fp, 100))
    assert loss[1] < loss[0]
    table = oracle.leaf_table(t["node_min"], t["node_max"], t["children"], act, W.DEPTH)
    prop = oracle.Proposal(t["node_min"], t["node_max"], table, W.DEPTH, ps.coarse.cpu().numpy(),
                           ps.params.cpu().numpy(), pp.mix, pp.kl_threshold, 2, JS)
    o, d = W.rays(4096, 21, det)
    rays = native.RayList(*rays_to_device(o, d))
    spl = native.SamplingParams()
    spl.max_samples_per_node = W.K
    spl.jitter_mode = 2
    spl.jitter_seed = JS
    ref = prop.sample(o, d)
    fb, nl, node, tlo, thi, kl = [x.cpu().numpy() for x in native.ProposalSample(tree, ps, rays, spl)]
    assert np.array_equal(fb, ref["fallback"]) and np.array_equal(nl, ref["n_learned"])
    assert np.array_equal(kl.view(np.uint32), ref["kl"].view(np.uint32))
    slot = np.arange(32)[:, None] < nl[None, :]
    assert np.array_equal(node[slot], ref["node"][slot])
    assert np.array_equal(tlo[slot].view(np.uint32), ref["tmin"][slot].view(np.uint32))
    assert np.array_equal(thi[slot].view(np.uint32), ref["tmax"][slot].view(np.uint32))
    iv, fbt = native.CreateProposalIntervals(tree, ps, rays, spl)
    off, inode, ilo, ihi, cnt, first, ray, fb_ref = _oracle_mixed_list(oracle, t, act, prop, o, d, W.K, 2, JS)
    assert np.array_equal(fbt.cpu().numpy(), fb_ref)
    assert np.array_equal(iv.ray_offset.cpu().numpy(), off) and np.array_equal(iv.node.cpu().numpy(), inode)
    assert np.array_equal(iv.count.cpu().numpy(), cnt) and np.array_equal(iv.first.cpu().numpy(), first)
    assert np.array_equal(iv.tmin.cpu().numpy().view(np.uint32), ilo.view(np.uint32))
    assert np.array_equal(iv.tmax.cpu().numpy().view(np.uint32), ihi.view(np.uint32))
    spr = iv.num_samples / o.shape[0]
    # FP32 projection over the mixed list (f32 master grid) against FP64
    master = st.grid_master
    fu = _fused(c1, o, d, grid=master, mlp=st.mlp, intervals=iv)
    P = fu.forward().cpu().double().numpy()
    fm = oracle.FusedModel(t["node_min"], t["node_max"], act, master.cpu().numpy(), st.mlp.cpu().numpy(), W.C, K=W.K,
                           jitter_mode=2, seed=JS)
    fm.set_intervals(off, inode, ilo, ihi, cnt)
    e32 = _rel(P, fm.project(o, d, precision=1))
    # the learned leg's own arithmetic (BF16 grid, BF16 decoder) against the emulating oracle
    fbf = _fused(c1, o, d, grid=st.grid, mlp=st.mlp, bf16_decoder=True, intervals=iv)
    Pb = fbf.forward().cpu().double().numpy()
    fe = oracle.FusedModel(t["node_min"], t["node_max"], act, st.grid.float().cpu().numpy(), st.mlp.cpu().numpy(),
                           W.C, K=W.K, jitter_mode=2, seed=JS, decoder_bf16=True)
    fe.set_intervals(off, inode, ilo, ihi, cnt)
    eb = _rel(Pb, fe.project(o, d, precision=1))
    print(f"c4 det {det}: proposal loss {loss[0]:.4f} -> {loss[1]:.4f}; {int(fb.sum())} of {fb.size} rays fall back; "
          f"{spr:.1f} samples/ray; projection FP32 vs FP64 {e32:.2e}, BF16 vs emulating {eb:.2e}", flush=True)
    assert e32 <= 1e-5 and eb <= 2e-4
    # full batch: learned rays carry one-sample cells, fallback rays the stratified intervals
    fo, fdd = W.rays(1 << 18, 400, det)
    frays = native.RayList(*rays_to_device(fo, fdd))
    ivf, fbf_ = native.CreateProposalIntervals(tree, ps, frays, spl)
    ivs = tree.CreateIntervalsForRays(frays, spl)
    fbm = fbf_.bool()
    n_l = (ivf.ray_offset[1:] - ivf.ray_offset[:-1])
    cnt_l = torch.zeros_like(n_l).index_add_(0, ivf.ray.long(), ivf.count.long())
    assert bool((n_l[~fbm] <= 32).all()) and bool((cnt_l[~fbm] == n_l[~fbm]).all())
    sel_f = fbm[ivf.ray.long()]
    sel_s = fbm[ivs.ray.long()]
    for a, b in ((ivf.node, ivs.node), (ivf.count, ivs.count), (ivf.tmin, ivs.tmin), (ivf.tmax, ivs.tmax)):
        assert torch.equal(a[sel_f], b[sel_s])
    print(f"  full batch: {int(fbm.sum())} fallback rays, {ivf.num_samples / frays.size():.1f} samples/ray",
          flush=True)
