"""Pins of the point-location and grid-transfer oracles.

Point location is checked against an independent formulation: among the
active nodes, the one whose box, half-open from below, (min, max] on each
axis (closed at the root's lower faces), holds the point.  That is the
partition the reference's first-child descent induces on the shared faces.
The grid transfer is checked against closed forms for a split and a merge
and against a numpy restatement built on that locator, with float32
positions computed one rounding at a time."""
import numpy as np

from scene_util import adaptive_active

G = 11


def _halfopen_locate(t, active, p):
    mn, mx = t["node_min"], t["node_max"]
    rmin = mn[0]
    ids = np.nonzero(active)[0]
    if active[0]:
        return 0
    lo_ok = (p[None, :] > mn[ids]) | ((p[None, :] == mn[ids]) & (mn[ids] == rmin[None, :]))
    hi_ok = p[None, :] <= mx[ids]
    hit = ids[np.all(lo_ok & hi_ok, axis=1)]
    assert hit.size <= 1, hit
    return int(hit[0]) if hit.size else -1


def _points(depth, rng, n=400):
    lat = (np.arange(2 ** (depth + 1) + 1) / 2 ** depth - 1.0).astype(np.float32)  # every face and mid-plane
    grid = np.stack(np.meshgrid(lat, lat, lat, indexing="ij"), -1).reshape(-1, 3)
    pick = grid[rng.choice(grid.shape[0], n, replace=False)]
    mixed = pick.copy()
    mixed[:, 1] = rng.uniform(-1, 1, n).astype(np.float32)  # on a face in some axes only
    rand = rng.uniform(-1.1, 1.1, (n, 3)).astype(np.float32)
    corners = np.array([[-1, -1, -1], [1, 1, 1], [-1, 1, -1], [0, 0, 0], [1.0000001, 0, 0]], np.float32)
    return np.concatenate([pick, mixed, rand, corners])


def _culled_active(t, seed):
    act = adaptive_active(t, budget=120, seed=seed).copy()
    # drop one sibling group without covering it: a culled region
    par = t["parent"]
    ids = np.nonzero(act)[0]
    groups = [p for p in np.unique(par[ids]) if p >= 0 and np.all(act[t["children"][p]] == 1)]
    act[t["children"][groups[0]]] = 0
    return act


def test_node_for_position_matches_halfopen_partition(oracle):
    rng = np.random.default_rng(0)
    for depth, seed in ((3, 0), (3, 1), (2, 2)):
        t = oracle.tree_build(depth)
        for act in (adaptive_active(t, budget=150, seed=seed), _culled_active(t, seed)):
            pts = _points(depth, rng)
            got = [oracle.node_for_position(t["node_min"], t["node_max"], t["children"], act, p) for p in pts]
            want = [_halfopen_locate(t, act, p) for p in pts]
            assert got == want
            assert -1 in want  # the outside points (and, on the culled tree, the hole)
    # root active: every point, even outside the root box, maps to the root
    t = oracle.tree_build(1)
    act = np.zeros(9, np.int32)
    act[0] = 1
    assert oracle.node_for_position(t["node_min"], t["node_max"], t["children"], act, [5.0, 0, 0]) == 0
    node_id, mask = oracle.node_id_for_position(t["node_min"], t["node_max"], t["children"], act,
                                                [[0.1, 0.2, 0.3], [3.0, 0, 0]])
    assert node_id.tolist() == [0, 0] and mask.tolist() == [1.0, 1.0]
    # NodeIdForPositionGPU's fallback: the first active node, mask 0
    act = (t["depth"] == 1).astype(np.int32)
    node_id, mask = oracle.node_id_for_position(t["node_min"], t["node_max"], t["children"], act,
                                                [[0.5, 0.5, 0.5], [1.5, 0, 0]])
    assert node_id.tolist() == [8, 1] and mask.tolist() == [1.0, 0.0]


def test_transfer_split_and_merge_closed_forms(oracle):
    """Node 1 (the lower octant of the root) split into its children and the
    children merged back; its lower faces are root faces, so every lattice
    point stays inside node 1's region.  G = 9 keeps every lattice position
    and tie (i / 2) exact in float32, so the exact-arithmetic closed form
    holds; at G = 11 float rounding moves some ties (the restatement below
    covers that)."""
    G = 9
    rng = np.random.default_rng(1)
    t = oracle.tree_build(2)
    depth1 = (t["depth"] == 1).astype(np.int32)
    kids = t["children"][1]
    split = depth1.copy()
    split[1] = 0
    split[kids] = 1
    Cc = 4
    src = rng.normal(size=(t["depth"].size, G, G, G, Cc)).astype(np.float32)  # slot = node id
    dst = np.zeros_like(src)
    oracle.grid_transfer(t["node_min"], t["node_max"], t["children"], depth1, None, src, kids, None, dst, 0)
    half = (G - 1) // 2
    i = np.arange(G)
    for c, k in enumerate(kids):
        bits = [(c >> a) & 1 for a in range(3)]  # x, y, z
        ix = [half * b + (i + 1) // 2 for b in bits]  # round half away from zero of off*(G-1) + i/2
        want = src[1][np.ix_(ix[2], ix[1], ix[0])]
        assert np.array_equal(dst[k], want)
    # merge: parent lattice i -> lower child (i <= half) at 2i, upper child at 2(i - half)
    dst = np.zeros_like(src)
    oracle.grid_transfer(t["node_min"], t["node_max"], t["children"], split, None, src, [1], None, dst, 0)
    for z in range(G):
        for y in range(G):
            for x in range(G):
                b = [int(v > half) for v in (x, y, z)]
                k = kids[b[0] + 2 * b[1] + 4 * b[2]]
                j = [2 * v - 2 * half * bb for v, bb in zip((x, y, z), b)]
                assert np.array_equal(dst[1, z, y, x], src[k, j[2], j[1], j[0]])


def _transfer_numpy(t, src_active, src_slot, src, dst_nodes, dst_slot, dst, layout):
    f32 = np.float32
    gm1 = f32(G - 1)
    mn, mx = t["node_min"], t["node_max"]
    for node in dst_nodes:
        ds = node if dst_slot is None else dst_slot[node]
        if src_active[node]:  # kept node: its own grid
            dst[ds] = src[node if src_slot is None else src_slot[node]]
            continue
        size = (mx[node] - mn[node]).astype(f32)
        for z in range(G):
            for y in range(G):
                for x in range(G):
                    res = np.array([x, y, z], f32)
                    g = (res / gm1).astype(f32) * size
                    g = (g.astype(f32) + mn[node]).astype(f32)
                    tn = _halfopen_locate(t, src_active, g)
                    if tn < 0:
                        continue
                    ss = tn if src_slot is None else src_slot[tn]
                    ts = (mx[tn] - mn[tn]).astype(f32)
                    c = ((g - mn[tn]).astype(f32) / ts).astype(f32)
                    f = (c * gm1).astype(f32)
                    fi = np.floor(f.astype(np.float64) + 0.5).astype(int)
                    if layout == 0:
                        dst[ds, z, y, x] = src[ss, fi[2], fi[1], fi[0]]
                    else:
                        dst[ds, :, z, y, x] = src[ss, :, fi[2], fi[1], fi[0]]
    return dst


def test_transfer_matches_numpy_restatement(oracle):
    rng = np.random.default_rng(2)
    t = oracle.tree_build(3)
    old = _culled_active(t, 3)
    new = adaptive_active(t, budget=90, seed=4)
    Nn = t["depth"].size
    # compact slots (the fused path's layout 0)
    old_slot = np.where(old == 1, np.cumsum(old) - 1, -1).astype(np.int32)
    new_slot = np.where(new == 1, np.cumsum(new) - 1, -1).astype(np.int32)
    Cc = 3
    src = rng.normal(size=(int(old.sum()), G, G, G, Cc)).astype(np.float32)
    init = rng.normal(size=(int(new.sum()), G, G, G, Cc)).astype(np.float32)
    kept = np.nonzero(new & old)[0][:4]
    assert kept.size > 0
    dst_nodes = np.concatenate([np.nonzero(new & (1 - old))[0][:12], kept])
    got = oracle.grid_transfer(t["node_min"], t["node_max"], t["children"], old, old_slot, src, dst_nodes, new_slot,
                               init.copy(), 0)
    want = _transfer_numpy(t, old, old_slot, src, dst_nodes, new_slot, init.copy(), 0)
    assert np.array_equal(got, want)
    assert not np.array_equal(got, init)
    # the reference layout and operation: inactive nodes filled from the active ones
    full = rng.normal(size=(Nn, Cc, G, G, G)).astype(np.float32)
    inactive = np.nonzero(old == 0)[0][:10]
    got = oracle.grid_transfer(t["node_min"], t["node_max"], t["children"], old, None, full, inactive, None,
                               full.copy(), 1)
    want = _transfer_numpy(t, old, None, full, inactive, None, full.copy(), 1)
    assert np.array_equal(got, want)


def test_quad_tree_point_location_matches_halfopen_partition(oracle):
    """The quad-tree representation (four children splitting x and z, every
    node spanning [y_min, y_max]): the descent equals the half-open
    partition; points outside the y range are outside every node."""
    rng = np.random.default_rng(4)
    for depth, seed in ((3, 0), (4, 1)):
        t = oracle.tree_build(depth, 4, -0.4, 0.4)
        act = adaptive_active(t, budget=90, seed=seed)
        assert t["children"].shape[1] == 4 and act.sum() > 30
        pts = _points(depth, rng)
        pts[::7, 1] = np.float32(0.4)   # on the upper y face (closed)
        pts[1::7, 1] = np.float32(-0.4)  # on the lower y face (a root face: closed)
        got = [oracle.node_for_position(t["node_min"], t["node_max"], t["children"], act, p) for p in pts]
        want = [_halfopen_locate(t, act, p) for p in pts]
        assert got == want
        assert -1 in want and sum(w >= 0 for w in want) > len(want) // 3
