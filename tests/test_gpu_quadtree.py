"""The quad-tree representation (use_quad_tree_rep: four children per node
splitting x and z, every node spanning [y_min, y_max]; structure/
HyperTree.cpp:391-406, 461-464) on the GPU against the oracle: traversal
(per-ray and the cooperative kernel over a ray list) bit-exact, projection
and adjoint within 1e-5, point location and the proposal's leaf table
bit-exact, NodeNeighborSamples2D pairs (structure/HyperTree.cu:1815-1913)
bit-exact, and a restructure whose grid transfer equals the oracle's."""
import ctypes as C

import numpy as np
import pytest
import torch

from conftest import edge_case_rays, random_rays
from scene_util import adaptive_active, device_tree, rays_to_device, tilt_rays

pytestmark = pytest.mark.gpu
Y = (-0.4, 0.45)


def _scene(oracle, depth=4, budget=150, seed=0):
    t = oracle.tree_build(depth, 4, *Y)
    act = adaptive_active(t, budget=budget, seed=seed, start_depth=1)
    return t, act, device_tree(depth, act, quad=Y)


def _rays(n, seed):
    o1, d1 = random_rays(n // 2, 90 + seed, spread=1.0)
    o2, d2 = tilt_rays(n - n // 2, 95 + seed)
    o3, d3 = edge_case_rays()
    return np.concatenate([o1, o2, o3]), np.concatenate([d1, d2, d3])


@pytest.mark.parametrize("jitter_mode", [0, 2])
def test_quad_traversal_bit_exact(oracle, native, cuda, jitter_mode):
    t, act, tree = _scene(oracle)
    assert tree.ChildrenPerNode() == 4
    o, d = _rays(3000, 1)
    rays = native.RayList(*rays_to_device(o, d))
    sp = native.SamplingParams()
    sp.max_samples_per_node = 6
    sp.jitter_mode = jitter_mode
    sp.jitter_seed = 17
    iv = tree.CreateIntervalsForRays(rays, sp)  # raises if the front-to-back order check fails
    r = oracle.traverse(t["node_min"], t["node_max"], oracle.active_ids(act), o, d, 6, jitter_mode=jitter_mode,
                        seed=17)
    print(f"quad tree: {int(act.sum())} active nodes, {iv.num_intervals} intervals, {iv.num_samples} samples",
          flush=True)
    assert iv.num_intervals > 10 * o.shape[0] // 2
    assert np.array_equal(iv.ray_offset.cpu().numpy(), r["ray_interval_offset"])
    for k, rk in (("node", "iv_node"), ("count", "iv_count"), ("first", "iv_first")):
        assert np.array_equal(getattr(iv, k).cpu().numpy(), r[rk]), k
    for k, rk in (("tmin", "iv_tmin"), ("tmax", "iv_tmax")):
        assert np.array_equal(getattr(iv, k).cpu().numpy().view(np.uint32), r[rk].view(np.uint32)), k
    if jitter_mode == 0:
        sl = tree.CreateSamplesForRays(rays, 6, False)
        for k in ("ray_t", "weight", "global_coordinate"):
            assert np.array_equal(getattr(sl, k).cpu().numpy().view(np.uint32), r[k].view(np.uint32)), k
        assert np.array_equal(sl.node_id.cpu().numpy(), r["node_id"])


def test_quad_listed_traversal(oracle, native, cuda):
    """The cooperative traversal (eight lanes per ray, four busy on a quad
    tree) over a ray list gives the listed rays the per-ray kernel's lists."""
    from adaptivetensordf_amd import capi

    t, act, tree = _scene(oracle, seed=2)
    o, d = _rays(4000, 2)
    R = o.shape[0]
    o_t, d_t = rays_to_device(o, d)
    rng = np.random.default_rng(5)
    sel = torch.from_numpy((rng.random(R) < 0.2).astype(np.int32)).cuda()
    lst = torch.empty(R, dtype=torch.int32, device="cuda")
    cnt = torch.empty(1, dtype=torch.int32, device="cuda")
    s = capi.stream_ptr()
    capi.call("atdf_compact_rays", capi.ptr(sel), R, capi.ptr(lst), capi.ptr(cnt), s)
    out = []
    for use_list in (False, True):
        sp = capi.Sampling(6, 256, 2, 0, 3, None, sel.data_ptr(), lst.data_ptr() if use_list else None)
        niv = torch.full((R,), -7, dtype=torch.int32, device="cuda")
        ns = torch.full((R,), -7, dtype=torch.int32, device="cuda")
        status = torch.zeros(4, dtype=torch.int32, device="cuda")
        capi.call("atdf_traverse_count", capi.ptr(tree.PackedTree()), tree.NumNodes(), capi.ptr(o_t), capi.ptr(d_t),
                  R, C.byref(sp), capi.ptr(niv), capi.ptr(ns), capi.ptr(status), s)
        out.append((niv.cpu().numpy(), ns.cpu().numpy(), status.cpu().numpy()))
    m = sel.cpu().numpy().astype(bool)
    assert out[0][2][1] == 0 and out[1][2][1] == 0  # order checks
    assert np.array_equal(out[0][0][m], out[1][0][m]) and np.array_equal(out[0][1][m], out[1][1][m])
    r = oracle.traverse(t["node_min"], t["node_max"], oracle.active_ids(act), o, d, 6)
    assert np.array_equal(out[1][0][m], r["ray_num_intervals"][m])


def test_quad_projection_and_adjoint(oracle, native, cuda):
    from adaptivetensordf_amd import phantom

    t, act, tree = _scene(oracle, seed=3)
    o, d = _rays(1024, 3)
    rays = native.RayList(*rays_to_device(o, d))
    C_, G, K = 16, 9, 6
    grid = phantom.init_grid(int(act.sum()), G, C_, std=0.3, seed=3)
    mlp = phantom.init_mlp(C_, 64, seed=3)
    sp = native.SamplingParams()
    sp.max_samples_per_node = K
    sp.jitter_mode = 2
    sp.jitter_seed = 11
    fp = native.FieldParams()
    fp.C, fp.G = C_, G
    g = grid.cuda().requires_grad_()
    m = mlp.cuda().requires_grad_()
    P = native.ProjectRays(tree, rays, g, m, sp, fp)
    fm = oracle.FusedModel(t["node_min"], t["node_max"], act, grid.numpy(), mlp.numpy(), C_, K=K, jitter_mode=2,
                           seed=11)
    P64 = fm.project(o, d, precision=1)
    Pg = P.detach().cpu().double().numpy()
    scale = np.maximum(np.abs(P64), 1e-3 * np.abs(P64).max())
    e = np.max(np.abs(Pg - P64) / scale)
    clear = fm.min_preact(o, d) >= 3e-6
    w = np.random.default_rng(3).normal(size=o.shape[0]) * clear
    gg, gm = torch.autograd.grad(P, [g, m], torch.from_numpy(w).float().cuda())
    ref_g, ref_m = fm.project_bwd(o, d, w)
    eg = np.abs(gg.cpu().double().numpy() - ref_g).max() / np.abs(ref_g).max()
    em = np.abs(gm.cpu().double().numpy() - ref_m).max() / np.abs(ref_m).max()
    print(f"quad tree projection rel err {e:.2e}; adjoint grid {eg:.2e}, decoder {em:.2e} "
          f"({int(clear.sum())} of {clear.size} rays clear of kinks)", flush=True)
    assert e <= 1e-5 and eg <= 1e-5 and em <= 1e-5


def test_quad_point_location_and_leaf_table(oracle, native, cuda):
    t, act, tree = _scene(oracle, seed=4)
    rng = np.random.default_rng(4)
    pts = rng.uniform(-1.05, 1.05, (20000, 3)).astype(np.float32)
    lat = (np.arange(33) / 16 - 1.0).astype(np.float32)
    pts[:1089, 0] = np.repeat(lat, 33)
    pts[:1089, 2] = np.tile(lat, 33)  # every face and edge of the depth-4 quadrants
    pts[:200, 1] = np.float32(Y[1])
    nid, mask = tree.NodeIdForPositionGPU(torch.from_numpy(pts).cuda())
    rn, rm = oracle.node_id_for_position(t["node_min"], t["node_max"], t["children"], act, pts)
    assert np.array_equal(nid.cpu().numpy(), rn) and np.array_equal(mask.cpu().numpy(), rm)
    assert 0 < rm.sum() < rm.size
    table = oracle.leaf_table(t["node_min"], t["node_max"], t["children"], act, 4)
    assert np.array_equal(tree.LeafTable().cpu().numpy(), table)


@pytest.mark.parametrize("use_roi", [False, True])
def test_quad_neighbor_pairs(oracle, native, cuda, use_roi):
    t, act, tree = _scene(oracle, seed=5)
    size = [6, 4, 9]
    roi = ([-0.6, Y[1] + 0.0015, -0.7], [0.8, 2.0, 0.5])
    sl = tree.NodeNeighborSamplesPlane2D(size, 1e-3, 1, roi[0], roi[1], use_roi, True)
    pos, node = oracle.neighbor_pairs(t["node_min"], t["node_max"], t["children"], act, size, eps=1e-3,
                                      roi=roi if use_roi else None, plane=1)
    print(f"use_roi={use_roi}: {node.size // 2} pairs", flush=True)
    assert node.size > 0
    assert np.array_equal(sl.node_id.cpu().numpy(), node)
    assert np.array_equal(sl.global_coordinate.cpu().numpy().view(np.uint32), pos.view(np.uint32))


def test_quad_restructure_grid_transfer(oracle, native, cuda):
    from adaptivetensordf_amd import phantom

    depth, G = 3, 9
    t = oracle.tree_build(depth, 4, *Y)
    act = (t["depth"] == 2).astype(np.int32)
    tree = device_tree(depth, act, quad=Y)
    fp = native.FieldParams()
    fp.G = G
    rng = np.random.default_rng(6)
    grid = torch.from_numpy((0.1 * rng.normal(size=(int(act.sum()), G, G, G, fp.C))).astype(np.float32)).cuda()
    st = native.MakeTrainState(grid, phantom.init_mlp(fp.C, fp.H, 0).cuda(), False, 1e-3, 1e-3)
    err = rng.uniform(0.0, 1.0, t["depth"].size).astype(np.float32)
    ids = np.nonzero(act)[0]
    for n in ids[[1, 6, 14]]:  # three nodes worth splitting
        err[n] = 1.0
        err[t["children"][n]] = 1e-3
    err[t["children"][3]] = 1e-4  # a sibling group (nodes 13-16) cheap to merge
    tree.node_error = torch.from_numpy(err).cuda()
    tree.node_max_density = torch.ones(t["depth"].size, device="cuda")
    old = tree.node_active.cpu().numpy().copy()
    master = st.grid.cpu().numpy().copy()
    p = native.TreeOptimizerParams()
    p.max_active_nodes = int(act.sum()) + 3 * 3 - 3
    changed, dec = native.RestructureTrainState(tree, st, p)
    new = tree.node_active.cpu().numpy()
    print(f"quad restructure: {int(old.sum())} -> {int(new.sum())} ({dec.num_split} split, {dec.num_grp} merged)",
          flush=True)
    assert changed and dec.num_split == 3 and dec.num_grp == 4

    def slots(a):
        s = np.full(a.size, -1, np.int32)
        s[np.nonzero(a)[0]] = np.arange(int(a.sum()), dtype=np.int32)
        return s

    ref = oracle.grid_transfer(t["node_min"], t["node_max"], t["children"], old, slots(old), master,
                               np.nonzero(new)[0], slots(new), np.zeros((int(new.sum()), G, G, G, fp.C), np.float32), 0)
    assert np.array_equal(st.grid.cpu().numpy(), ref)
