"""Point location and grid transfer on the GPU against the oracle (bit-exact),
and the structure step of training: OptimizeTreeStructure, the move of the
grids to the new active set, the optimizer reset, and a training step on
the restructured tree."""
import numpy as np
import pytest
import torch

from conftest import random_rays
from scene_util import adaptive_active, device_tree, rays_to_device
from test_oracle_restructure import _culled_active, _points

pytestmark = pytest.mark.gpu
G = 11


def test_node_id_for_position(oracle, native, cuda):
    rng = np.random.default_rng(0)
    for depth, seed in ((3, 0), (2, 1), (4, 2)):
        t = oracle.tree_build(depth)
        for act in (adaptive_active(t, budget=150, seed=seed), _culled_active(t, seed)):
            tree = device_tree(depth, act)
            pts = _points(depth, rng)
            node, mask = tree.NodeIdForPositionGPU(torch.from_numpy(pts).cuda())
            rn, rm = oracle.node_id_for_position(t["node_min"], t["node_max"], t["children"], act, pts)
            print(f"depth {depth}: {pts.shape[0]} points, {int((rm == 0).sum())} outside the active set", flush=True)
            assert np.array_equal(node.cpu().numpy(), rn) and np.array_equal(mask.cpu().numpy(), rm)
            assert 0 < (rm == 0).sum() < rm.size


def test_interpolate_grid_for_inactive_nodes(oracle, native, cuda):
    rng = np.random.default_rng(1)
    t = oracle.tree_build(3)
    act = _culled_active(t, 5)
    tree = device_tree(3, act)
    full = rng.normal(size=(t["depth"].size, 8, G, G, G)).astype(np.float32)
    out = tree.InterpolateGridForInactiveNodes(torch.from_numpy(full).cuda()).cpu().numpy()
    inactive = np.nonzero(act == 0)[0]
    ref = oracle.grid_transfer(t["node_min"], t["node_max"], t["children"], act, None, full, inactive, None,
                               full.copy(), 1)
    assert np.array_equal(out, ref)
    assert np.array_equal(out[act == 1], full[act == 1])


def _slots(active):
    return np.where(active == 1, np.cumsum(active) - 1, -1).astype(np.int32)


def test_transfer_node_grids(oracle, native, cuda):
    rng = np.random.default_rng(2)
    t = oracle.tree_build(3)
    old = _culled_active(t, 3)
    new = adaptive_active(t, budget=90, seed=4)
    tree = device_tree(3, old)
    old_packed, old_slot = tree.PackedTree(), tree.node_to_grid
    C_ = 16
    src = rng.normal(size=(int(old.sum()), G, G, G, C_)).astype(np.float32)
    tree.node_active = torch.from_numpy(new).cuda()
    tree.UpdateActive()
    init = rng.normal(size=(int(new.sum()), G, G, G, C_)).astype(np.float32)
    got = native.TransferNodeGrids(tree, old_packed, old_slot, torch.from_numpy(src).cuda(),
                                   torch.from_numpy(init).cuda()).cpu().numpy()
    ref = oracle.grid_transfer(t["node_min"], t["node_max"], t["children"], old, _slots(old), src, np.nonzero(new)[0],
                               _slots(new), init.copy(), 0)
    assert np.array_equal(got, ref)
    zero = native.TransferNodeGrids(tree, old_packed, old_slot, torch.from_numpy(src).cuda()).cpu().numpy()
    ref0 = oracle.grid_transfer(t["node_min"], t["node_max"], t["children"], old, _slots(old), src,
                                np.nonzero(new)[0], _slots(new), np.zeros_like(init), 0)
    assert np.array_equal(zero, ref0)


@pytest.mark.parametrize("bf16", [False, True])
def test_restructure_train_state(oracle, native, cuda, bf16):
    from adaptivetensordf_amd import phantom

    rng = np.random.default_rng(3)
    depth = 3
    t = oracle.tree_build(depth)
    act = (t["depth"] == 2).astype(np.int32)
    tree = device_tree(depth, act)
    fp = native.FieldParams()
    C_ = fp.C
    grid = torch.from_numpy((0.1 * rng.normal(size=(int(act.sum()), G, G, G, C_))).astype(np.float32)).cuda()
    mlp = phantom.init_mlp(C_, fp.H, 0).cuda()
    st = native.MakeTrainState(grid, mlp, bf16, 1e-3, 1e-3)
    o, d = random_rays(2048, 7, spread=0.9)
    rays = native.RayList(*rays_to_device(o, d))
    target = torch.rand(2048, device="cuda")
    sp = native.SamplingParams()
    sp.max_samples_per_node = 4
    for _ in range(3):
        native.TrainStep(tree, st, rays, target, sp, fp)
    # saved child errors that make five nodes worth splitting, and a budget
    # that forces two sibling groups to merge to pay for it
    err = rng.uniform(0.0, 1.0, t["depth"].size).astype(np.float32)
    ids = np.nonzero(act)[0]
    for n in ids[[3, 17, 30, 41, 60]]:
        err[t["children"][n]] = 1e-3
    for g in (2, 5):  # two nearly perfect sibling groups: cheap to merge
        err[t["children"][g]] = 1e-4
    tree.node_error = torch.from_numpy(err).cuda()
    tree.node_max_density = torch.ones(t["depth"].size, device="cuda")
    old_act = tree.node_active.cpu().numpy().copy()
    master = (st.grid_master if bf16 else st.grid).cpu().numpy().copy()
    p = native.TreeOptimizerParams()
    p.max_active_nodes = int(act.sum()) + 5 * 7 - 14
    changed, dec = native.RestructureTrainState(tree, st, p)
    assert changed and dec.num_split == 5 and dec.num_grp == 16  # 64 + 5 * 7 - 2 * 7 active
    new_act = tree.node_active.cpu().numpy()
    print(f"bf16={bf16}: {int(old_act.sum())} -> {int(new_act.sum())} active nodes "
          f"({dec.num_split} split, {dec.num_grp} merged)", flush=True)
    ref = oracle.grid_transfer(t["node_min"], t["node_max"], t["children"], old_act, _slots(old_act), master,
                               np.nonzero(new_act)[0], _slots(new_act),
                               np.zeros((int(new_act.sum()), G, G, G, C_), np.float32), 0)
    moved = (st.grid_master if bf16 else st.grid).cpu().numpy()
    assert np.array_equal(moved, ref)
    if bf16:
        assert torch.equal(st.grid, st.grid_master.to(torch.bfloat16))
    assert st.step == 0 and float(st.m_grid.abs().sum()) == 0 and float(st.v_mlp.abs().sum()) == 0
    for _ in range(3):
        native.TrainStep(tree, st, rays, target, sp, fp)
    assert st.step == 3 and np.isfinite(st.loss.item())
