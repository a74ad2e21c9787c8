"""Structure optimizer (CPU): the knapsack DP against exhaustive enumeration
of the reference's MIP (structure/HyperTreeStructureOptimizer.cpp:70-332)
on small trees, the ApplySplit semantics, and NodeCulling."""
import itertools

import numpy as np
import pytest

def _tree(native, depth, active, culled, err, dens, quad=False):
    t = native.HyperTreeBase(3, depth)
    t.settree(quad, -0.4, 0.4)
    import torch

    t.node_active = torch.from_numpy(active.astype(np.int32))
    t.node_culled = torch.from_numpy(culled.astype(np.int32))
    t.UpdateActive()
    t.node_error = torch.from_numpy(err.astype(np.float32))
    t.node_max_density = torch.from_numpy(dens.astype(np.float32))
    return t


def _in_roi(mn, mx, lo, hi):
    mid = mn + mx
    return all(((mn[k] > lo[k]) and (mn[k] < hi[k])) or ((mx[k] > lo[k]) and (mx[k] < hi[k])) or
               ((mid[k] > lo[k]) and (mid[k] < hi[k])) for k in range(3))


def _brute(t, p):
    """Minimum of the MIP over every assignment allowed by its bounds, with
    the objective coefficients computed as the reference does (float32
    losses, float64 products)."""
    act = t.node_active.numpy()
    cul = t.node_culled.numpy()
    par = t.node_parent.numpy()
    ch = t.node_children.numpy()
    NS = ch.shape[1]  # 8, or 4 for the quad-tree representation
    err = t.node_error.numpy()
    den = t.node_max_density.numpy()
    mn, mx = t.node_position_min.numpy(), t.node_position_max.numpy()
    ids = [int(i) for i in np.nonzero(act == 1)[0]]
    f32 = np.float32
    info = {}
    for i in ids:
        pa = par[i]
        d = {"none": float(err[i]) * float(den[i]), "elig": False, "nsib": 0, "sibs": []}
        if pa >= 0:
            sib = ch[pa]
            na = int(act[sib].sum())
            nc = int(cul[sib].sum())
            s = f32(0)
            dmax = f32(0)
            nn = 0
            for q in sib:
                if not cul[q] and act[q]:
                    e = max(err[q], f32(0))
                    s = f32(s + e)
                    dmax = max(den[q], dmax)
                    nn += 1
            lg = f32(p.error_merge_factor * float(s))
            lg = f32(lg / f32(nn))
            d.update(grp=float(lg) * float(dmax), elig=(na + nc == NS), nsib=na, sibs=[int(q) for q in sib if act[q]])
        can_split = ch[i, 0] >= 0 and _in_roi(mn[i], mx[i], p.tree_edge_roi_min, p.tree_edge_roi_max)
        if can_split:
            c = ch[i]
            ncc = int(cul[c].sum())
            if p.use_saved_errors and err[c[0]] > 0:
                ls = f32(0)
                for q in c:
                    ls = f32(ls + f32(err[q] * den[q]))
                ds = f32(1)
            else:
                ls = f32(p.error_split_factor * float(err[i]))
                ds = den[i]
            d.update(split=float(ls) * float(ds), wsplit=NS - ncc)
        d["can_split"] = can_split
        d["has_children"] = ch[i, 0] >= 0
        info[i] = d
    opts = []
    for i in ids:
        o = ["none"]
        if info[i]["can_split"]:
            o.append("split")
        if par[i] >= 0:
            o.append("grp")
        opts.append(o)
    best = None
    for combo in itertools.product(*opts):
        a = dict(zip(ids, combo))
        ok = True
        weight = 0.0
        cost = 0.0
        for i in ids:
            d = info[i]
            lhs = (a[i] == "none") + (a[i] == "split" if d["has_children"] else 0)
            if d["elig"]:
                lhs += sum(a[s] == "grp" for s in d["sibs"]) / d["nsib"]
            if abs(lhs - 1.0) > 1e-9:
                ok = False
                break
            weight += (a[i] == "none") + ((a[i] == "grp") / d["nsib"] if d["elig"] else 0)
            if a[i] == "split":
                weight += d["wsplit"]
            cost += d["none"] if a[i] == "none" else (d["split"] if a[i] == "split" else d["grp"])
        if ok and weight <= p.max_active_nodes + 1e-9 and (best is None or cost < best[0] - 1e-15):
            best = (cost, a, weight)
    return best, info


def _check_decision(t, p, dec, info):
    act = t.node_active.numpy()
    ids = [int(i) for i in np.nonzero(act == 1)[0]]
    cmap = {0: "none", 1: "split", 2: "grp"}
    a = {i: cmap[dec.choice[i]] for i in ids}
    cost = sum(info[i]["none"] if a[i] == "none" else (info[i]["split"] if a[i] == "split" else info[i]["grp"])
               for i in ids)
    for i in ids:
        d = info[i]
        if a[i] == "split":
            assert d["can_split"]
        if d["elig"]:
            assert len({a[s] == "grp" for s in d["sibs"]}) == 1  # the group merges as a whole
        else:
            assert a[i] != "grp"
    return cost


def _random_case(native, rng, depth, active_ids, culled_ids, saved, quad=False):
    n = sum((4 if quad else 8) ** k for k in range(depth + 1))
    active = np.zeros(n, np.int32)
    active[list(active_ids)] = 1
    culled = np.zeros(n, np.int32)
    culled[list(culled_ids)] = 1
    err = rng.uniform(0.0, 1.0, n).astype(np.float32)
    if not saved:
        err[~active.astype(bool)] = -1
    dens = rng.uniform(0.0, 2.0, n).astype(np.float32)
    return _tree(native, depth, active, culled, err, dens, quad)


CASES = [
    # (max depth, active node ids, culled node ids)
    (3, range(1, 9), []),                            # one full sibling group at depth 1
    (2, [1, 2, 3, 4] + list(range(9 + 8 * 4, 9 + 8 * 5)), [6, 7, 8]),  # culled siblings, one split node
    (3, [2, 3, 4, 5, 6, 7, 8] + list(range(9, 13)), list(range(13, 17))),  # culled grand-children
]


# the quad-tree representation: children 4i+1 .. 4i+4 of node i
QUAD_CASES = [
    (3, range(1, 5), []),                                  # one full sibling group at depth 1
    (2, [1, 2] + list(range(13, 17)), [4]),                # a culled sibling, one split node
    (3, [2, 3, 4, 5, 6] + list(range(21, 25)), [7, 8]),    # culled grand-children, a split child
]


@pytest.mark.parametrize("case", [("oct", i) for i in range(len(CASES))] + [("quad", i) for i in range(len(QUAD_CASES))])
@pytest.mark.parametrize("saved", [True, False])
def test_dp_equals_exhaustive_mip(native, case, saved):
    kind, ci = case
    quad = kind == "quad"
    depth, act_ids, cul_ids = (QUAD_CASES if quad else CASES)[ci]
    rng = np.random.default_rng(10 * ci + saved + (100 if quad else 0))
    for trial in range(4):
        t = _random_case(native, rng, depth, act_ids, cul_ids, saved, quad)
        na = int(t.node_active.numpy().sum())
        for budget in (na - 3, na, na + 7, 200):
            p = native.TreeOptimizerParams()
            p.max_active_nodes = max(1, budget)
            p.use_saved_errors = saved
            if trial == 3:
                p.tree_edge_roi_min = [-0.2, -1.0, -1.0]  # restrict where nodes may split
            best, info = _brute(t, p)
            dec = native.SolveTreeStructure(t, p)
            if best is None:
                assert not dec.feasible
                continue
            assert dec.feasible
            assert abs(dec.objective - best[0]) <= 1e-12 * max(1.0, abs(best[0])), (case, trial, budget)
            if quad:
                assert t.ChildrenPerNode() == 4
            cost = _check_decision(t, p, dec, info)
            assert abs(cost - dec.objective) <= 1e-12 * max(1.0, abs(cost))
            assert dec.active_weight <= p.max_active_nodes


def test_apply_split_and_merge(native):
    rng = np.random.default_rng(3)
    # one full group at depth 1: cheap merge, or splits where the children are much better
    t = _random_case(native, rng, 2, range(1, 9), [], True)
    err = t.node_error.numpy().copy()
    err[1:9] = 1.0
    err[9:17] = 0.0001  # children of node 1 are near-perfect: node 1 should split
    import torch

    t.node_error = torch.from_numpy(err)
    t.node_max_density = torch.ones_like(t.node_max_density)
    p = native.TreeOptimizerParams()
    p.max_active_nodes = 15
    changed, dec = native.OptimizeTreeStructure(t, p)
    assert changed and dec.choice[1] == 1
    act = t.node_active.numpy()
    assert act[1] == 0 and np.all(act[9:17] == 1) and np.all(act[2:9] == 1)
    assert t.NumActiveNodes() == 15
    # a budget of 1 leaves merging the group as the only feasible move
    t2 = _random_case(native, rng, 2, range(1, 9), [], True)
    p.max_active_nodes = 1
    changed, dec = native.OptimizeTreeStructure(t2, p)
    assert changed and all(dec.choice[i] == 2 for i in range(1, 9))
    assert t2.node_active.numpy()[0] == 1 and t2.NumActiveNodes() == 1


def test_node_culling(native):
    import torch

    t = native.HyperTreeBase(3, 2)
    t.settree()
    t.SetActive(2)
    dens = np.full(t.NumNodes(), 1.0, np.float32)
    ids = np.nonzero(t.node_active.numpy())[0]
    low = ids[[0, 5, 40]]
    dens[low] = 0.001
    t.node_max_density = torch.from_numpy(dens)
    # ROI excluding the first low-density node keeps it from being culled only if it is inside
    mn = t.node_position_min.numpy()[low[0]]
    native.NodeCulling(t, 0.01, [2.0, 2.0, 2.0], [3.0, 3.0, 3.0], 0.0)  # ROI outside the cube: cull all three
    cul = t.node_culled.numpy()
    assert set(np.nonzero(cul)[0]) >= set(low.tolist())
    assert t.NumActiveNodes() == 64 - 3
    assert np.all(t.node_active.numpy()[low] == 0)
    assert mn.shape == (3,)


def test_quad_tree_apply_split_and_culling(native):
    """ApplySplit and UpdateCulling over four children per node."""
    import torch

    rng = np.random.default_rng(5)
    t = _random_case(native, rng, 2, range(1, 5), [], True, quad=True)
    err = t.node_error.numpy().copy()
    err[1:5] = 1.0
    err[5:9] = 0.0001  # the children of node 1 (5..8) are near-perfect: node 1 splits
    t.node_error = torch.from_numpy(err)
    t.node_max_density = torch.ones_like(t.node_max_density)
    p = native.TreeOptimizerParams()
    p.max_active_nodes = 7
    changed, dec = native.OptimizeTreeStructure(t, p)
    act = t.node_active.numpy()
    assert changed and dec.choice[1] == 1
    assert act[1] == 0 and np.all(act[5:9] == 1) and np.all(act[2:5] == 1) and t.NumActiveNodes() == 7
    # culling: every child of node 2 culled -> node 2 is culled too
    t2 = native.HyperTreeBase(3, 2)
    t2.settree(True, -0.4, 0.4)
    t2.SetActive(2)
    cul = np.zeros(t2.NumNodes(), np.int32)
    cul[9:13] = 1
    act = t2.node_active.numpy().copy()
    act[9:13] = 0
    t2.node_culled = torch.from_numpy(cul)
    t2.node_active = torch.from_numpy(act)
    t2.UpdateActive()
    t2.UpdateCulling()
    assert t2.node_culled.numpy()[2] == 1 and t2.node_culled.numpy()[1] == 0
