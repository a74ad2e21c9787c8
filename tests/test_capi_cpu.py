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
    assert lib.atdf_scan_workspace_bytes(1024) == 8


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
