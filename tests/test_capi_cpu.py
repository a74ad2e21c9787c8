"""CPU-only checks of the boundary: libatdf.so loads and exports every entry
point include/atdf.h declares, argument validation fails cleanly before any
device work, and the libtorch host layer's tree logic matches the oracle."""
import os
import re

import numpy as np
import pytest
import torch

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(REPO, "include", "atdf.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(atdf_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from adaptivetensordf_amd import capi

    lib = capi.lib()
    decl = declared_symbols()
    assert len(decl) >= 20
    for name in decl:
        assert hasattr(lib, name), name
    assert set(decl) == set(capi.exported_symbols())
    assert b"gfx950" in lib.atdf_version()


def test_validation_without_device_work():
    import ctypes as C

    from adaptivetensordf_amd import capi

    lib = capi.lib()
    assert lib.atdf_tree_pack(None, None, None) == capi.ERR_INVALID_ARGUMENT
    assert b"null" in lib.atdf_last_error()
    sp = capi.Sampling(0, 256, 0, 0, 0, None)
    st = lib.atdf_traverse_count(C.c_void_p(1), 9, None, None, 4, C.byref(sp), None, None, None, None)
    assert st == capi.ERR_INVALID_ARGUMENT and b"max_samples_per_node" in lib.atdf_last_error()
    f = capi.Field(12, 64, 11, 0, 2.0, 0, None, 0, None, None)
    sp = capi.Sampling(4, 256, 0, 0, 0, None)
    st = lib.atdf_project_fwd(C.c_void_p(1), 9, None, None, 4, C.byref(sp), None, None, C.byref(f), None, None)
    assert st == capi.ERR_INVALID_ARGUMENT and b"channels" in lib.atdf_last_error()
    assert lib.atdf_tree_packed_bytes(10) == 480
    lo, hi = (C.c_float * 3)(-0.5, 0.2, -1.0), (C.c_float * 3)(0.5, 0.2, 1.0)  # empty along y
    st = lib.atdf_tree_pack_box(lo, hi, C.c_void_p(1), None)
    assert st == capi.ERR_INVALID_ARGUMENT and b"axis 1" in lib.atdf_last_error()
    assert lib.atdf_tree_pack_box(None, hi, C.c_void_p(1), None) == capi.ERR_INVALID_ARGUMENT
    assert lib.atdf_scan_workspace_bytes(1024) == 8


def test_tensorial_field_validation_without_device_work():
    """atdf_vm checks and the entry points that reject a tensorial field,
    all decided before any device work (fake, aligned addresses)."""
    import ctypes as C

    from adaptivetensordf_amd import capi

    lib = capi.lib()
    A = [C.c_void_p(256 * (i + 1)) for i in range(16)]
    vm = capi.VM(A[0], A[1], 40, A[2], 9, 17, -0.5, 0.5, A[3], A[4])
    f = capi.Field(16, 64, 5, capi.GRID_F32, 2.0, 0, A[5], 30, A[6], A[7], capi.DECODER_F32, 0, C.pointer(vm))
    sp = capi.Sampling(4, 256, 0, 0, 0, None)
    iv = capi.Intervals(A[8], A[9], A[10], A[11], A[12], A[13], 100)
    args = (A[14], 9, A[15], A[15], 0, C.byref(sp), A[8], C.byref(iv), C.byref(f))
    assert lib.atdf_project_fwd(*args, A[0], None) == capi.OK  # no rays: checks only
    ws_vm = lib.atdf_project_bwd_workspace_bytes(1000, 100, C.byref(f))
    f_plain = capi.Field(16, 64, 5, capi.GRID_F32, 2.0, 0, A[5], 30 + 40, A[6], A[7], capi.DECODER_F32, 0)
    assert ws_vm > lib.atdf_project_bwd_workspace_bytes(1000, 100, C.byref(f_plain)) > 0  # + line partial rows
    bad = [(capi.Field(16, 64, 5, capi.GRID_BF16, 2.0, 0, A[5], 30, A[6], A[7], 0, 0, C.pointer(vm)), b"f32 only"),
           (capi.Field(16, 64, 5, 0, 2.0, 0, A[5], 30, A[6], A[7], 0, 0,
                       C.pointer(capi.VM(A[0], A[1], 40, A[2], 1, 17, -0.5, 0.5))), b">= 2"),
           (capi.Field(16, 64, 5, 0, 2.0, 0, A[5], 30, A[6], A[7], 0, 0,
                       C.pointer(capi.VM(A[0], A[1], 40, A[2], 9, 17, 0.5, 0.5))), b"ROI x range"),
           (capi.Field(16, 64, 5, 0, 2.0, 0, A[5], 30, A[6], A[7], 0, 0,
                       C.pointer(capi.VM(C.c_void_p(260), A[1], 40, A[2], 9, 17, -0.5, 0.5))), b"aligned"),
           (capi.Field(16, 64, 5, 0, 2.0, 0, A[5], 30, A[6], A[7], 0, 0,
                       C.pointer(capi.VM(A[0], None, 40, A[2], 9, 17, -0.5, 0.5))), b"incomplete tensorial")]
    for fb, msg in bad:
        st = lib.atdf_project_fwd(A[14], 9, A[15], A[15], 0, C.byref(sp), A[8], C.byref(iv), C.byref(fb), A[0], None)
        assert st == capi.ERR_INVALID_ARGUMENT and msg in lib.atdf_last_error(), msg
    # one ray: the checks pass, then these entry points refuse the field
    # (atdf_project_fwd_prepared takes it: the node-major tensorial forward)
    args1 = (A[14], 9, A[15], A[15], 1, C.byref(sp), A[8], C.byref(iv), C.byref(f))
    st = lib.atdf_project_fwd_bins(*args1, A[0], A[1], None)
    assert st == capi.ERR_INVALID_ARGUMENT and b"tensorial" in lib.atdf_last_error()
    ga = capi.GridAdam(A[5], None, A[2], A[3], 1e-3, 0.9, 0.999, 1e-8, 1)
    st = lib.atdf_project_bwd_adam(*args1, A[0], A[5], A[7], C.byref(ga), A[1], 1 << 30, None)
    assert st == capi.ERR_INVALID_ARGUMENT and b"tensorial" in lib.atdf_last_error()
    vm_nograd = capi.VM(A[0], A[1], 40, A[2], 9, 17, -0.5, 0.5, None, None)
    f_ng = capi.Field(16, 64, 5, 0, 2.0, 0, A[5], 30, A[6], A[7], 0, 0, C.pointer(vm_nograd))
    st = lib.atdf_project_bwd(A[14], 9, A[15], A[15], 1, C.byref(sp), A[8], C.byref(iv), C.byref(f_ng), A[0], A[5],
                              A[7], A[1], 1 << 30, None)
    assert st == capi.ERR_INVALID_ARGUMENT and b"tensorial gradient" in lib.atdf_last_error()


def test_proposal_boundary_without_device_work(oracle):
    import ctypes as C

    from adaptivetensordf_amd import capi, phantom

    lib = capi.lib()
    assert lib.atdf_proposal_num_params() == oracle.prop_num_params() == phantom.proposal_num_params()
    assert lib.atdf_leaf_table_numel(5) == 32 ** 3 and lib.atdf_leaf_table_numel(0) == 0
    assert lib.atdf_leaf_table(C.c_void_p(1), 9, 9, C.c_void_p(1), None) == capi.ERR_INVALID_ARGUMENT
    assert b"depth" in lib.atdf_last_error()
    p = capi.Proposal(None, 5, 0.25, 0.5, 0, 0, None, None)
    st = lib.atdf_proposal_sample(C.c_void_p(1), 9, None, None, 4, C.byref(p), *([None] * 7))
    assert st == capi.ERR_INVALID_ARGUMENT and b"incomplete proposal" in lib.atdf_last_error()
    p = capi.Proposal(C.c_void_p(1), 5, 1.5, 0.5, 0, 0, C.c_void_p(1), C.c_void_p(1))
    st = lib.atdf_proposal_features(C.c_void_p(1), 9, None, None, 4, C.byref(p), None, None, None)
    assert st == capi.ERR_INVALID_ARGUMENT and b"mix" in lib.atdf_last_error()
    assert lib.atdf_proposal_grad(None, None, None, 0, 1.0, None, None, None, 0, None) == capi.ERR_INVALID_ARGUMENT
    assert lib.atdf_proposal_grad_workspace_bytes(64) > 64 * 455 * 4


def test_restructure_boundary_without_device_work():
    import ctypes as C

    from adaptivetensordf_amd import capi

    lib = capi.lib()
    one = C.c_void_p(1)
    st = lib.atdf_grid_transfer(one, 9, None, one, 9, one, 3, None, C.c_void_p(2), 9, 8, 1, 0, None)
    assert st == capi.ERR_INVALID_ARGUMENT and b"bad C" in lib.atdf_last_error()
    st = lib.atdf_grid_transfer(one, 9, None, one, 9, one, 3, None, one, 9, 8, 11, 0, None)
    assert st == capi.ERR_INVALID_ARGUMENT and b"alias" in lib.atdf_last_error()
    assert lib.atdf_grid_transfer(one, 9, None, None, 9, None, 0, None, None, 9, 8, 11, 1, None) == capi.OK
    assert lib.atdf_node_id_for_position(one, 9, None, None, 0, None, None, None) == capi.OK
    assert lib.atdf_compact_rays(None, -1, None, one, None) == capi.ERR_INVALID_ARGUMENT
    assert lib.atdf_compact_rays(None, 5, None, one, None) == capi.ERR_INVALID_ARGUMENT
    st = lib.atdf_node_id_for_position(one, 9, None, one, 5, one, one, None)
    assert st == capi.ERR_INVALID_ARGUMENT and b"null buffer" in lib.atdf_last_error()


def test_host_tree_matches_oracle(oracle, native):
    t = native.HyperTreeBase(3, 3)
    t.settree()
    ref = oracle.tree_build(3)
    assert np.array_equal(t.node_position_min.numpy(), ref["node_min"])
    assert np.array_equal(t.node_position_max.numpy(), ref["node_max"])
    assert np.array_equal(t.node_children.numpy(), ref["children"])
    assert np.array_equal(t.node_parent.numpy(), ref["parent"])
    assert np.array_equal(t.node_depth.numpy(), ref["depth"])
    assert t.NumActiveNodes() == 1  # SetActive(0) after settree, as the reference
    t.SetActive(3)
    assert np.array_equal(t.active_node_ids.numpy(), np.arange(73, 585))
    pre = t.node_active_prefix_sum.numpy()
    assert pre[72] == -1 and pre[73] == 0 and pre[584] == 511
    assert np.array_equal(t.node_to_grid.numpy(), pre.astype(np.int32))


def test_host_quad_tree_matches_oracle(oracle, native):
    t = native.HyperTreeBase(3, 3)
    t.settree(True, -0.3, 0.45)
    ref = oracle.tree_build(3, 4, -0.3, 0.45)
    assert t.ChildrenPerNode() == 4
    for a, b in (("node_position_min", "node_min"), ("node_position_max", "node_max"), ("node_children", "children"),
                 ("node_parent", "parent"), ("node_depth", "depth")):
        assert np.array_equal(getattr(t, a).numpy(), ref[b]), a
    t.SetActive(3)
    assert np.array_equal(t.active_node_ids.numpy(), np.arange(21, 85))


def test_host_error_and_culling_semantics(native):
    t = native.HyperTreeBase(3, 2)
    t.settree()
    t.SetActive(2)
    n = t.NumNodes()
    err = torch.arange(n, dtype=torch.float32)
    t.SetErrorForActiveNodes(err, "override")
    e = t.node_error.numpy()
    assert np.all(e[:9] == -1) and np.array_equal(e[9:], np.arange(9, n, dtype=np.float32))
    t.SetErrorForActiveNodes(err * 0.5, "min")
    assert np.array_equal(t.node_error.numpy()[9:], 0.5 * np.arange(9, n, dtype=np.float32))
    # cull all children of node 1 -> node 1 becomes culled; culled errors are zeroed
    c = t.node_culled.clone()
    c[t.node_children[1].long()] = 1
    t.node_culled = c
    a = t.node_active.clone()
    a[t.node_children[1].long()] = 0
    t.node_active = a
    t.UpdateCulling()
    assert t.node_culled[1].item() == 1 and t.node_culled[2].item() == 0
    assert t.node_error[t.node_children[1].long()].abs().sum().item() == 0
    with pytest.raises(RuntimeError):
        t.SetErrorForActiveNodes(torch.zeros(3), "override")


def _np_in_roi(mn, mx, lo, hi, eps):
    """geometry/geometry.h:426-449 restated: on every axis the min, the max or
    min + max (the reference's "mid") lies strictly inside the interval; the
    y interval widened by eps."""
    lo = np.array(lo, np.float32).copy()
    hi = np.array(hi, np.float32).copy()
    lo[1] -= np.float32(eps)
    hi[1] += np.float32(eps)
    mid = mn + mx
    ok = ((mn > lo) & (mn < hi)) | ((mx > lo) & (mx < hi)) | ((mid > lo) & (mid < hi))
    return ok.all(axis=1)


def test_roi_split_host(native):
    t = native.HyperTreeBase(3, 3)
    t.settree()
    t.SetActive(2)
    lo, hi = [-0.5, -0.6, -1.0], [0.5, 0.2, 0.1]
    m = t.RoiSplit(lo, hi, 0.0)
    act = t.node_active.numpy().astype(bool)
    mn, mx = t.node_position_min.numpy(), t.node_position_max.numpy()
    ref = _np_in_roi(mn, mx, lo, hi, 0.0) & act
    assert np.array_equal(m.in_roi_flag.numpy().astype(bool), ref)
    assert np.array_equal(m.in_roi_node_index.numpy(), np.nonzero(ref)[0])
    assert np.array_equal(m.out_roi_node_index.numpy(), np.nonzero(act & ~ref)[0])
    pin = m.prefix_inroi.numpy()
    assert np.array_equal(pin[ref], np.arange(ref.sum())) and np.all(pin[~ref] == -1)
    pout = m.prefix_outroi.numpy()
    assert np.array_equal(pout[act & ~ref], np.arange((act & ~ref).sum())) and np.all(pout[~(act & ~ref)] == -1)
    # hand-checked nodes: the depth-2 node [-0.5, 0]^3 is inside; [0.5, 1] on x is not (min 0.5 is not < 0.5,
    # max 1 is outside, and min + max = 1.5 is outside)
    ids = np.nonzero(act)[0]
    inside = [i for i in ids if np.allclose(mn[i], [-0.5, -0.5, -0.5])][0]
    outside = [i for i in ids if np.isclose(mn[i][0], 0.5)][0]
    assert ref[inside] and not ref[outside]
    # the widened y interval admits a node whose y-min sits on the ROI bound
    m2 = t.RoiSplit([-1, 0.5, -1], [1, 1, 1], 1e-6)
    assert np.array_equal(m2.in_roi_flag.numpy().astype(bool), _np_in_roi(mn, mx, [-1, 0.5, -1], [1, 1, 1], 1e-6) & act)
