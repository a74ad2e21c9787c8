"""Traversal + sample placement on the GPU, through the C ABI and through the
libtorch CreateSamplesForRays, against the oracle: node / interval / sample
indices bit-exact, t, weights and positions bit-exact."""
import ctypes as C

import numpy as np
import pytest
import torch

from conftest import edge_case_rays, random_rays
from scene_util import adaptive_active, device_tree, rays_to_device, tilt_rays, uniform_active

pytestmark = pytest.mark.gpu


def capi_traverse(tree, o_t, d_t, K, max_inter=256, jitter_mode=0, seed=0, jitter=None, cache_slots=0,
                  samples=True, select=None):
    """Count -> scan -> write through the raw C ABI; returns numpy arrays.
    cache_slots > 0: the cached entry points (the write pass replays the
    count pass's hits of every ray with at most cache_slots of them).
    samples=False: the write pass writes the interval list alone."""
    from adaptivetensordf_amd import capi

    dev = o_t.device
    R = o_t.shape[0]
    s = capi.stream_ptr()
    sp = capi.Sampling(K, max_inter, jitter_mode, 0, seed, None if jitter is None else jitter.data_ptr(),
                       None if select is None else select.data_ptr())
    packed = torch.empty(capi.lib().atdf_tree_packed_bytes(tree.NumNodes()), dtype=torch.uint8, device=dev)
    tdesc = capi.Tree(tree.NumNodes(), tree.node_position_min.data_ptr(), tree.node_position_max.data_ptr(),
                      tree.node_children.data_ptr(), tree.node_parent.data_ptr(), tree.node_active.data_ptr())
    capi.call("atdf_tree_pack", C.byref(tdesc), capi.ptr(packed), s)
    niv = torch.empty(R, dtype=torch.int32, device=dev)
    ns = torch.empty(R, dtype=torch.int32, device=dev)
    status = torch.zeros(4, dtype=torch.int32, device=dev)
    cache = None
    if cache_slots > 0:
        # poisoned, so a record the write pass reads without the count pass
        # having written it shows up as a wrong interval
        cache = torch.full((capi.lib().atdf_traverse_cache_bytes(R, cache_slots),), 0x7F, dtype=torch.uint8,
                           device=dev)
        capi.call("atdf_traverse_count_cached", capi.ptr(packed), tree.NumNodes(), capi.ptr(o_t), capi.ptr(d_t), R,
                  C.byref(sp), capi.ptr(niv), capi.ptr(ns), capi.ptr(status), capi.ptr(cache), cache_slots, s)
    else:
        capi.call("atdf_traverse_count", capi.ptr(packed), tree.NumNodes(), capi.ptr(o_t), capi.ptr(d_t), R,
                  C.byref(sp), capi.ptr(niv), capi.ptr(ns), capi.ptr(status), s)
    ws = torch.empty(max(8, capi.lib().atdf_scan_workspace_bytes(R)), dtype=torch.uint8, device=dev)
    ivo = torch.empty(R + 1, dtype=torch.int64, device=dev)
    so = torch.empty(R + 1, dtype=torch.int64, device=dev)
    capi.call("atdf_exclusive_scan_i32", capi.ptr(niv), R, capi.ptr(ivo), capi.ptr(ws), ws.numel(), s)
    capi.call("atdf_exclusive_scan_i32", capi.ptr(ns), R, capi.ptr(so), capi.ptr(ws), ws.numel(), s)
    H, S = int(ivo[-1]), int(so[-1])
    f32 = dict(dtype=torch.float32, device=dev)
    i32 = dict(dtype=torch.int32, device=dev)
    i64 = dict(dtype=torch.int64, device=dev)
    iv = dict(node=torch.empty(H, **i32), tmin=torch.empty(H, **f32), tmax=torch.empty(H, **f32),
              count=torch.empty(H, **i32), first=torch.empty(H, **i32), ray=torch.empty(H, **i32))
    sm = dict(global_coordinate=torch.empty(S, 3, **f32), weight=torch.empty(S, **f32), ray_t=torch.empty(S, **f32),
              ray_index=torch.empty(S, **i64), local_index_in_ray=torch.empty(S, **i64), node_id=torch.empty(S, **i64))
    civ = capi.Intervals(*[iv[k].data_ptr() for k in ("node", "tmin", "tmax", "count", "first", "ray")])
    csm = capi.Samples(*[sm[k].data_ptr() for k in ("global_coordinate", "weight", "ray_t", "ray_index",
                                                    "local_index_in_ray", "node_id")])
    psm = C.byref(csm) if samples else None
    if H > 0 and cache is not None:
        capi.call("atdf_traverse_write_cached", capi.ptr(packed), tree.NumNodes(), capi.ptr(o_t), capi.ptr(d_t), R,
                  C.byref(sp), capi.ptr(ivo), capi.ptr(so), C.byref(civ), psm, capi.ptr(cache), cache_slots, s)
    elif H > 0:  # the write pass needs real buffers; with no hits there is nothing to write
        capi.call("atdf_traverse_write", capi.ptr(packed), tree.NumNodes(), capi.ptr(o_t), capi.ptr(d_t), R,
                  C.byref(sp), capi.ptr(ivo), capi.ptr(so), C.byref(civ), psm, s)
    torch.cuda.synchronize()
    out = {k: v.cpu().numpy() for k, v in {**iv, **sm}.items()}
    out.update(status=status.cpu().numpy(), ray_num_intervals=niv.cpu().numpy(), ray_num_samples=ns.cpu().numpy(),
               ray_interval_offset=ivo.cpu().numpy(), ray_sample_offset=so.cpu().numpy())
    return out


def assert_same_traversal(g, r):
    assert np.array_equal(g["ray_num_intervals"], r["ray_num_intervals"])
    assert np.array_equal(g["ray_num_samples"], r["ray_num_samples"])
    assert np.array_equal(g["ray_interval_offset"], r["ray_interval_offset"])
    assert np.array_equal(g["node"], r["iv_node"])
    assert np.array_equal(g["count"], r["iv_count"])
    assert np.array_equal(g["first"], r["iv_first"])
    assert np.array_equal(g["tmin"].view(np.uint32), r["iv_tmin"].view(np.uint32))
    assert np.array_equal(g["tmax"].view(np.uint32), r["iv_tmax"].view(np.uint32))
    for k in ("ray_index", "local_index_in_ray", "node_id"):
        assert np.array_equal(g[k], r[k]), k
    for k in ("ray_t", "weight", "global_coordinate"):
        assert np.array_equal(g[k].view(np.uint32), r[k].view(np.uint32)), k


@pytest.mark.parametrize("case", ["uniform3", "adaptive4", "uniform1"])
@pytest.mark.parametrize("jitter_mode", [0, 2])
def test_capi_traversal_bit_exact(oracle, cuda, case, jitter_mode):
    depth = {"uniform3": 3, "adaptive4": 4, "uniform1": 1}[case]
    t = oracle.tree_build(depth)
    act = adaptive_active(t, budget=600) if case == "adaptive4" else uniform_active(t, depth)
    tree = device_tree(depth, act)
    o, d = random_rays(700, 21)
    eo, ed = edge_case_rays()
    to_, td = tilt_rays(400, 3)
    o, d = np.concatenate([o, eo, to_]), np.concatenate([d, ed, td])
    K = 7
    print(f"case={case} jitter={jitter_mode} R={o.shape[0]} active={int(act.sum())}", flush=True)
    g = capi_traverse(tree, *rays_to_device(o, d), K, jitter_mode=jitter_mode, seed=99)
    r = oracle.traverse(t["node_min"], t["node_max"], oracle.active_ids(act), o, d, K, jitter_mode=jitter_mode,
                        seed=99)
    assert g["status"][0] == 0 and g["status"][1] == 0
    assert g["status"][2] == r["ray_num_samples"].max()
    assert_same_traversal(g, r)
    assert np.array_equal(g["ray"], np.repeat(np.arange(o.shape[0]), r["ray_num_intervals"]))


@pytest.mark.parametrize("samples", [True, False])
@pytest.mark.parametrize("cache_slots", [64, 3, 1])
@pytest.mark.parametrize("jitter_mode", [0, 2])
def test_capi_traversal_hit_cache(oracle, cuda, cache_slots, jitter_mode, samples):
    """The cached count / write passes: bit-exact against the oracle whether a
    ray's hits come from the cache (at most cache_slots of them) or from a
    second walk (more); with samples (per-ray replay) and for an interval
    list alone (the coalesced wave-cooperative write)."""
    t = oracle.tree_build(4)
    act = adaptive_active(t, budget=600)
    tree = device_tree(4, act)
    o, d = random_rays(700, 23)
    eo, ed = edge_case_rays()
    to_, td = tilt_rays(400, 5)
    o, d = np.concatenate([o, eo, to_]), np.concatenate([d, ed, td])
    print(f"cache_slots={cache_slots} jitter={jitter_mode} R={o.shape[0]}", flush=True)
    g = capi_traverse(tree, *rays_to_device(o, d), 6, jitter_mode=jitter_mode, seed=7, cache_slots=cache_slots,
                      samples=samples)
    r = oracle.traverse(t["node_min"], t["node_max"], oracle.active_ids(act), o, d, 6, jitter_mode=jitter_mode,
                        seed=7)
    n = r["ray_num_intervals"]
    # both paths are exercised: rays within the slots and rays beyond them
    assert (n <= cache_slots).any() and (cache_slots == 64 or (n > cache_slots).any())
    assert g["status"][0] == 0 and g["status"][1] == 0
    if samples:
        assert_same_traversal(g, r)
    else:
        assert np.array_equal(g["ray_interval_offset"], r["ray_interval_offset"])
        assert np.array_equal(g["node"], r["iv_node"])
        assert np.array_equal(g["count"], r["iv_count"])
        assert np.array_equal(g["first"], r["iv_first"])
        assert np.array_equal(g["tmin"].view(np.uint32), r["iv_tmin"].view(np.uint32))
        assert np.array_equal(g["tmax"].view(np.uint32), r["iv_tmax"].view(np.uint32))
    assert np.array_equal(g["ray"], np.repeat(np.arange(o.shape[0]), n))


def test_capi_traversal_hit_cache_with_selection(oracle, cuda):
    """Rays switched off by ray_select (the proposal's fallback mask) get no
    intervals on the cached path either: cached and uncached interval lists
    are identical."""
    t = oracle.tree_build(4)
    act = adaptive_active(t, budget=600)
    tree = device_tree(4, act)
    o, d = random_rays(1500, 29)
    sel = torch.from_numpy((np.random.default_rng(3).random(o.shape[0]) < 0.6).astype(np.int32)).cuda()
    args = (tree, *rays_to_device(o, d), 6)
    a = capi_traverse(*args, jitter_mode=2, seed=1, samples=False, select=sel)
    for slots in (64, 2):
        b = capi_traverse(*args, jitter_mode=2, seed=1, samples=False, select=sel, cache_slots=slots)
        for k in ("ray_interval_offset", "node", "count", "first", "ray"):
            assert np.array_equal(a[k], b[k]), (slots, k)
        for k in ("tmin", "tmax"):
            assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), (slots, k)
    assert (a["ray_num_intervals"][sel.cpu().numpy() == 0] == 0).all()


def test_capi_traversal_explicit_jitter_array(oracle, cuda):
    t = oracle.tree_build(3)
    act = uniform_active(t, 3)
    tree = device_tree(3, act)
    o, d = random_rays(300, 4)
    r0 = oracle.traverse(t["node_min"], t["node_max"], oracle.active_ids(act), o, d, 5)
    S = r0["weight"].shape[0]
    jit = np.random.default_rng(0).uniform(0, 1, S).astype(np.float32)
    g = capi_traverse(tree, *rays_to_device(o, d), 5, jitter_mode=1, jitter=torch.from_numpy(jit).cuda())
    r = oracle.traverse(t["node_min"], t["node_max"], oracle.active_ids(act), o, d, 5, jitter_mode=1, jitter_arr=jit)
    assert_same_traversal(g, r)


def test_create_samples_for_rays_api(oracle, cuda):
    """The libtorch drop-in (HyperTreeBaseImpl::CreateSamplesForRays)."""
    from adaptivetensordf_amd import native

    t = oracle.tree_build(4)
    act = adaptive_active(t, budget=500, seed=2)
    tree = device_tree(4, act)
    o, d = tilt_rays(2000, 8)
    rays = native().RayList(*rays_to_device(o, d))
    sl = tree.CreateSamplesForRays(rays, 6, False)
    r = oracle.traverse(t["node_min"], t["node_max"], oracle.active_ids(act), o, d, 6)
    for k in ("ray_index", "local_index_in_ray", "node_id"):
        assert np.array_equal(getattr(sl, k).cpu().numpy(), r[k])
    for k in ("ray_t", "weight", "global_coordinate"):
        assert np.array_equal(getattr(sl, k).cpu().numpy().view(np.uint32), r[k].view(np.uint32))
    assert sl.max_samples_per_ray == r["ray_num_samples"].max()
    # jittered: same sample slots, every t inside its stratum, weights telescoping
    sj = tree.CreateSamplesForRays(rays, 6, True)
    assert np.array_equal(sj.node_id.cpu().numpy(), r["node_id"])
    w = sj.weight.cpu().numpy()
    seg = np.add.reduceat(w, r["ray_sample_offset"][:-1][r["ray_num_samples"] > 0])
    ref = np.add.reduceat(r["weight"], r["ray_sample_offset"][:-1][r["ray_num_samples"] > 0])
    assert np.allclose(seg, ref, atol=1e-4)
    # ComputeLocalSamples drop-in: bit-exact with the oracle
    loc = tree.ComputeLocalSamples(sl.global_coordinate, sl.node_id).cpu().numpy()
    lo = oracle.local_samples(t["node_min"], t["node_max"], r["global_coordinate"], r["node_id"])
    assert np.array_equal(loc.view(np.uint32), lo.view(np.uint32))


def test_too_many_intersections_is_an_error(oracle, cuda):
    from adaptivetensordf_amd import native

    t = oracle.tree_build(3)
    act = uniform_active(t, 3)
    tree = device_tree(3, act)
    s = 1 / np.sqrt(3)
    o = np.array([[-3 * s + 0.01, -3 * s, -3 * s]], np.float32)
    d = np.array([[s, s, s]], np.float32)
    g = capi_traverse(tree, *rays_to_device(o, d), 2, max_inter=4)
    assert g["status"][0] == 1
    sp = native().SamplingParams()
    sp.max_samples_per_node = 2
    sp.max_intersections = 4
    with pytest.raises(RuntimeError, match="max_intersections"):
        tree.CreateIntervalsForRays(native().RayList(*rays_to_device(o, d)), sp)


def test_empty_and_missing_rays(oracle, cuda):
    t = oracle.tree_build(2)
    act = uniform_active(t, 2)
    tree = device_tree(2, act)
    o = np.array([[5.0, 5.0, -3.0], [0.0, 0.0, 3.0]], np.float32)  # misses / points away
    d = np.array([[0.0, 0.0, 1.0], [0.0, 0.0, 1.0]], np.float32)
    g = capi_traverse(tree, *rays_to_device(o, d), 4)
    assert g["ray_num_samples"].sum() == 0 and g["ray_interval_offset"][-1] == 0
    from adaptivetensordf_amd import native

    sl = tree.CreateSamplesForRays(native().RayList(*rays_to_device(o[:0], d[:0])), 4, False)
    assert sl.size() == 0


def test_traversal_full_size_tree(oracle, cuda):
    """Depth-5 adaptive tree of the 512^3 configuration shape (thousands of
    active nodes) on tilt-series rays of a 512^2 detector: bit-exact on a
    subset, size-independent invariants on the full batch."""
    from adaptivetensordf_amd import native, phantom

    t = oracle.tree_build(5)
    ell = phantom.ellipsoids(200, 1)
    act = phantom.adaptive_active_mask(t["node_min"], t["node_max"], t["children"], t["depth"], ell, 8192, 3, 5)
    tree = device_tree(5, act)
    ang = phantom.tilt_angles(61)
    tl, px, py = phantom.sample_pixels(1 << 16, 61, 512, 512, 5)
    o, d = phantom.rays_for_pixels(tl, px, py, 512, 512, ang)
    print(f"active={int(act.sum())} R={o.shape[0]}", flush=True)
    sp = native().SamplingParams()
    sp.max_samples_per_node = 5
    iv = tree.CreateIntervalsForRays(native().RayList(*rays_to_device(o, d)), sp)
    sub = slice(0, 1500)
    r = oracle.traverse(t["node_min"], t["node_max"], oracle.active_ids(act), o[sub], d[sub], 5)
    H = int(r["ray_interval_offset"][-1])
    assert np.array_equal(iv.node[:H].cpu().numpy(), r["iv_node"])
    assert np.array_equal(iv.count[:H].cpu().numpy(), r["iv_count"])
    assert np.array_equal(iv.tmin[:H].cpu().numpy().view(np.uint32), r["iv_tmin"].view(np.uint32))
    # invariants on every ray: chord through the cube == sum of interval lengths
    off = iv.ray_offset.cpu().numpy()
    lens = (iv.tmax - iv.tmin).double().cpu().numpy()
    per_ray = np.add.reduceat(lens, off[:-1][off[1:] > off[:-1]]) if lens.size else lens
    assert np.all(per_ray > 0) and np.all(per_ray < 2 * np.sqrt(3) + 1e-4)
    assert iv.num_samples == int(iv.ray_num_samples.sum())


def _listed_traverse(tree, o_t, d_t, K, lst, seed=99):
    """Count + scan + interval write with a ray list (the cooperative,
    eight-lanes-per-ray traversal); unlisted rays get no intervals."""
    from adaptivetensordf_amd import capi

    dev = o_t.device
    R = o_t.shape[0]
    s = capi.stream_ptr()
    sel = torch.zeros(R, dtype=torch.int32, device=dev)
    sel[lst[lst >= 0].long()] = 1
    sp = capi.Sampling(K, 256, 2, 0, seed, None, sel.data_ptr(), lst.data_ptr())
    packed = tree.PackedTree()
    niv = torch.zeros(R, dtype=torch.int32, device=dev)
    ns = torch.zeros(R, dtype=torch.int32, device=dev)
    status = torch.zeros(4, dtype=torch.int32, device=dev)
    capi.call("atdf_traverse_count", capi.ptr(packed), tree.NumNodes(), capi.ptr(o_t), capi.ptr(d_t), R,
              C.byref(sp), capi.ptr(niv), capi.ptr(ns), capi.ptr(status), s)
    ws = torch.empty(max(8, capi.lib().atdf_scan_workspace_bytes(R)), dtype=torch.uint8, device=dev)
    ivo = torch.empty(R + 1, dtype=torch.int64, device=dev)
    capi.call("atdf_exclusive_scan_i32", capi.ptr(niv), R, capi.ptr(ivo), capi.ptr(ws), ws.numel(), s)
    H = int(ivo[-1])
    i32 = dict(dtype=torch.int32, device=dev)
    f32 = dict(dtype=torch.float32, device=dev)
    iv = dict(node=torch.empty(H, **i32), tmin=torch.empty(H, **f32), tmax=torch.empty(H, **f32),
              count=torch.empty(H, **i32), first=torch.empty(H, **i32), ray=torch.empty(H, **i32))
    civ = capi.Intervals(*[iv[k].data_ptr() for k in ("node", "tmin", "tmax", "count", "first", "ray")], H)
    if H > 0:
        capi.call("atdf_traverse_write", capi.ptr(packed), tree.NumNodes(), capi.ptr(o_t), capi.ptr(d_t), R,
                  C.byref(sp), capi.ptr(ivo), None, C.byref(civ), None, s)
    torch.cuda.synchronize()
    out = {k: v.cpu().numpy() for k, v in iv.items()}
    out.update(status=status.cpu().numpy(), niv=niv.cpu().numpy(), ns=ns.cpu().numpy(), ivo=ivo.cpu().numpy())
    return out


@pytest.mark.parametrize("case", ["uniform3", "adaptive4", "adaptive5"])
def test_listed_traversal_bit_exact(oracle, cuda, case):
    """Traversal of a ray list (eight lanes per ray, children tested in
    parallel) against the oracle's DFS: counts, node order, tmin / tmax and
    sample counts bit-exact for every listed ray, nothing for the others."""
    from adaptivetensordf_amd import phantom

    depth = {"uniform3": 3, "adaptive4": 4, "adaptive5": 5}[case]
    t = oracle.tree_build(depth)
    if case == "adaptive5":
        ell = phantom.ellipsoids(200, 1)
        act = phantom.adaptive_active_mask(t["node_min"], t["node_max"], t["children"], t["depth"], ell, 2000, 3, 5)
    elif case == "adaptive4":
        act = adaptive_active(t, budget=600)
    else:
        act = uniform_active(t, depth)
    tree = device_tree(depth, act)
    o, d = random_rays(900, 31)
    eo, ed = edge_case_rays()
    to_, td = tilt_rays(500, 7)
    o, d = np.concatenate([o, eo, to_]), np.concatenate([d, ed, td])
    R = o.shape[0]
    rng = np.random.default_rng(depth)
    listed = np.sort(rng.choice(R, size=R // 3, replace=False))
    lst = np.full(R, -1, np.int32)
    lst[:listed.size] = rng.permutation(listed)  # list order is arbitrary
    o_t, d_t = rays_to_device(o, d)
    print(f"{case}: R={R}, listed {listed.size}, active {int(act.sum())}", flush=True)
    g = _listed_traverse(tree, o_t, d_t, 6, torch.from_numpy(lst).cuda())
    r = oracle.traverse(t["node_min"], t["node_max"], oracle.active_ids(act), o[listed], d[listed], 6)
    assert g["status"][0] == 0 and g["status"][1] == 0
    assert np.array_equal(g["niv"][listed], r["ray_num_intervals"])
    assert np.array_equal(g["ns"][listed], r["ray_num_samples"])
    mask = np.ones(R, bool)
    mask[listed] = False
    assert np.all(g["niv"][mask] == 0)
    assert np.array_equal(g["node"], r["iv_node"]) and np.array_equal(g["count"], r["iv_count"])
    assert np.array_equal(g["first"], r["iv_first"])
    assert np.array_equal(g["tmin"].view(np.uint32), r["iv_tmin"].view(np.uint32))
    assert np.array_equal(g["tmax"].view(np.uint32), r["iv_tmax"].view(np.uint32))
    assert np.array_equal(g["ray"], np.repeat(listed, r["ray_num_intervals"]))


def test_create_samples_for_rays_area(oracle, cuda):
    """CreateSamplesForRays_area / CreateIntervalsForRays_area on a depth-0
    tree with a non-root ROI box: samples and intervals bit-exact with the
    oracle's one-box traversal, node id 0 throughout; the projection over
    the box's intervals takes its local coordinates from node 0 (the root)
    and matches the oracle over the same intervals."""
    from test_oracle import ROI_MAX, ROI_MIN, area_traverse
    from adaptivetensordf_amd import native, phantom

    t = oracle.tree_build(0)
    act = np.ones(1, np.int32)
    tree = device_tree(0, act)
    o1, d1 = random_rays(3000, 12, spread=0.9)
    o2, d2 = tilt_rays(2000, 13)
    o3, d3 = edge_case_rays()
    o, d = np.concatenate([o1, o2, o3]), np.concatenate([d1, d2, d3])
    rays = native().RayList(*rays_to_device(o, d))
    K = 9
    sl = tree.CreateSamplesForRays_area(rays, K, False, list(ROI_MIN), list(ROI_MAX))
    r = area_traverse(oracle, o, d, K)
    print(f"area: {sl.size()} samples on {int((r['ray_num_intervals'] > 0).sum())} rays", flush=True)
    for k in ("ray_index", "local_index_in_ray", "node_id"):
        assert np.array_equal(getattr(sl, k).cpu().numpy(), r[k])
    for k in ("ray_t", "weight", "global_coordinate"):
        assert np.array_equal(getattr(sl, k).cpu().numpy().view(np.uint32), r[k].view(np.uint32))
    assert sl.max_samples_per_ray == r["ray_num_samples"].max() <= K
    sp = native().SamplingParams()
    sp.max_samples_per_node = K
    sp.jitter_mode = 2
    sp.jitter_seed = 31
    iv = tree.CreateIntervalsForRays_area(rays, sp, list(ROI_MIN), list(ROI_MAX))
    rj = area_traverse(oracle, o, d, K, jitter_mode=2, seed=31)
    assert np.array_equal(iv.ray_offset.cpu().numpy(), rj["ray_interval_offset"])
    assert np.array_equal(iv.node.cpu().numpy(), rj["iv_node"]) and np.all(rj["iv_node"] == 0)
    assert np.array_equal(iv.count.cpu().numpy(), rj["iv_count"])
    assert np.array_equal(iv.tmin.cpu().numpy().view(np.uint32), rj["iv_tmin"].view(np.uint32))
    assert np.array_equal(iv.tmax.cpu().numpy().view(np.uint32), rj["iv_tmax"].view(np.uint32))
    C_, G = 8, 9
    grid = phantom.init_grid(1, G, C_, std=0.4, seed=2)
    mlp = phantom.init_mlp(C_, 64, seed=2)
    fp = native().FieldParams()
    fp.C, fp.G = C_, G
    P = native().ProjectIntervals(tree, rays, iv, grid.cuda(), mlp.cuda(), sp, fp).cpu().double().numpy()
    fm = oracle.FusedModel(t["node_min"], t["node_max"], act, grid.numpy(), mlp.numpy(), C_, K=K, jitter_mode=2,
                           seed=31)
    fm.set_intervals(rj["ray_interval_offset"], rj["iv_node"], rj["iv_tmin"], rj["iv_tmax"], rj["iv_count"])
    P64 = fm.project(o, d, precision=1)
    scale = np.maximum(np.abs(P64), 1e-3 * np.abs(P64).max())
    err = np.max(np.abs(P - P64) / scale)
    print(f"projection over the ROI box's intervals: max rel err {err:.2e}", flush=True)
    assert err <= 1e-5 and np.all(P[rj["ray_num_intervals"] == 0] == 0)
