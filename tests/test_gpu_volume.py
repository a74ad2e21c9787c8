"""Volume export on the GPU: lattice points, point location and mask
bit-exact against the oracle; the field on the lattice within 1e-5 of the
oracle's composition (point location -> local coordinates -> trilinear
gather -> decoder -> softplus), zero outside the active set."""
import numpy as np
import pytest
import torch

from scene_util import adaptive_active, device_tree
from test_oracle_restructure import _culled_active

pytestmark = pytest.mark.gpu


def _field_oracle(oracle, t, act, grid_cl, mlp, pos, beta=2.0):
    node, mask = oracle.node_id_for_position(t["node_min"], t["node_max"], t["children"], act, pos)
    slots = np.where(act == 1, np.cumsum(act) - 1, -1)
    loc = oracle.local_samples(t["node_min"], t["node_max"], pos, node)
    grid_cf = np.ascontiguousarray(np.transpose(grid_cl, (0, 4, 1, 2, 3)))  # [N, C, Z, Y, X]
    feat = oracle.grid_sample3d_fwd(grid_cf, slots[node], loc)
    _, dens = oracle.mlp_fwd(mlp, grid_cl.shape[4], 64, beta, feat)
    return np.where(mask > 0, dens, 0.0), node, mask


def test_uniform_phantom_samples(oracle, native, cuda):
    t = oracle.tree_build(3)
    act = _culled_active(t, 1)
    tree = device_tree(3, act)
    for size, swap in (((9, 17, 5), False), ((8, 8, 8), True)):
        s = tree.UniformPhantomSamplesGPU(list(size), swap)
        pos = oracle.volume_points(0, size, swap_xy=swap)
        node, mask = oracle.node_id_for_position(t["node_min"], t["node_max"], t["children"], act, pos)
        assert np.array_equal(s.global_coordinate.cpu().numpy(), pos)
        assert np.array_equal(s.node_id.cpu().numpy(), node) and np.array_equal(s.weight.cpu().numpy(), mask)
        assert np.array_equal(s.ray_index.cpu().numpy(), np.arange(pos.shape[0]))
        assert 0 < (mask == 0).sum() < mask.size
    roi, step = [-0.8, -0.5, -0.9], [0.05, 0.07, 0.03]
    s = tree.UniformPhantomSamplesGPUbySlice([20, 10, 30], False, roi, step)
    pos = oracle.volume_points(1, (20, 10, 30), roi, step)
    assert np.array_equal(s.global_coordinate.cpu().numpy(), pos)


@pytest.mark.parametrize("C_", [16, 8])
def test_uniform_sampled_volume(oracle, native, cuda, C_):
    from adaptivetensordf_amd import phantom

    t = oracle.tree_build(3)
    act = _culled_active(t, 2)
    tree = device_tree(3, act)
    na = int(act.sum())
    grid = phantom.init_grid(na, 11, C_, std=0.3, seed=4)
    mlp = phantom.init_mlp(C_, 64, seed=4)
    fp = native.FieldParams()
    fp.C, fp.G = C_, 11
    shape = [33, 17, 25]
    vol, node, valid = native.UniformSampledVolume(tree, grid.cuda(), mlp.cuda(), fp, shape)
    assert vol.shape == (1, 33, 17, 25)
    pos = oracle.volume_points(0, shape)
    ref, rnode, rmask = _field_oracle(oracle, t, act, grid.numpy(), mlp.numpy(), pos)
    got = vol.cpu().numpy().ravel()  # point i at flat index i
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"C={C_}: {pos.shape[0]} points, {int((rmask == 0).sum())} outside, max rel err {err:.2e}", flush=True)
    assert err <= 1e-5
    assert np.all(got[rmask == 0] == 0)
    assert np.array_equal(node.cpu().numpy().ravel(), rnode)
    assert np.array_equal(valid.cpu().numpy().ravel(), rmask.astype(np.int64))
    # the by-slice lattice
    roi, step = [-0.9, -0.6, -0.95], [0.04, 0.05, 0.06]
    vs = native.SampleVolumeBySlice(tree, grid.cuda(), mlp.cuda(), fp, [40, 20, 30], roi, step)
    pos = oracle.volume_points(1, (40, 20, 30), roi, step)
    ref, _, _ = _field_oracle(oracle, t, act, grid.numpy(), mlp.numpy(), pos)
    assert np.abs(vs.cpu().numpy().ravel() - ref).max() <= 1e-5 * np.abs(ref).max()


def test_volume_through_analyze(oracle, native, cuda, tmp_path):
    from adaptivetensordf_amd import phantom

    t = oracle.tree_build(2)
    act = adaptive_active(t, budget=40, seed=0)
    tree = device_tree(2, act)
    grid = phantom.init_grid(int(act.sum()), 11, 16, std=0.3, seed=1).cuda()
    mlp = phantom.init_mlp(16, 64, seed=1).cuda()
    fp = native.FieldParams()
    vol, _, _ = native.UniformSampledVolume(tree, grid, mlp, fp, [24, 24, 24])
    path = str(tmp_path / "recon.hdr")
    native.SaveHDRImageTensor(vol, path)
    back = native.LoadHDRImageTensor(path)
    assert torch.equal(back, vol.cpu())


def test_bf16_volume_error_on_crop(oracle, native, cuda):
    """The BF16 field (BF16 grids, BF16-MFMA decoder, FP32 accumulation) on a
    64^3 lattice over the crop [-0.5, 0.5]^3 (the centre eighth of the cube,
    a 256^3 crop of the 512^3 volume sampled every fourth voxel) against the
    FP64 oracle: of the f32 master grid (the precision cost of BF16, stated
    and bounded) and of the BF16-rounded grid (the kernel's own arithmetic)."""
    from scene_util import adaptive_active, device_tree
    from adaptivetensordf_amd import phantom

    t = oracle.tree_build(4)
    act = adaptive_active(t, budget=400, seed=2)
    tree = device_tree(4, act)
    na = int(act.sum())
    grid = phantom.init_grid(na, 11, 16, std=0.3, seed=5)
    mlp = phantom.init_mlp(16, 64, seed=5)
    fp = native.FieldParams()
    fp.C, fp.G = 16, 11
    fp.bf16_decoder = True
    gb = grid.cuda().to(torch.bfloat16)
    n = 64
    roi, step = [-0.5] * 3, [1.0 / (n - 1)] * 3
    vol = native.SampleVolumeBySlice(tree, gb, mlp.cuda(), fp, [n, n, n], roi, step).cpu().numpy().reshape(-1)
    pos = oracle.volume_points(1, (n, n, n), roi, step)
    node, mask = oracle.node_id_for_position(t["node_min"], t["node_max"], t["children"], act, pos)
    node = np.where(mask > 0, node, -1)
    f32 = oracle.FusedModel(t["node_min"], t["node_max"], act, grid.numpy(), mlp.numpy(), 16)
    fbf = oracle.FusedModel(t["node_min"], t["node_max"], act, gb.float().cpu().numpy(), mlp.numpy(), 16,
                            decoder_bf16=True)
    d64 = oracle.field_density(f32, pos, node, precision=1)
    dbf = oracle.field_density(fbf, pos, node, precision=1)
    scale = np.abs(d64).max()
    e_bf16 = np.abs(vol - d64).max() / scale
    rms_bf16 = np.sqrt(np.mean((vol - d64) ** 2)) / np.sqrt(np.mean(d64 ** 2))
    e_self = np.abs(vol - dbf).max() / scale
    print(f"64^3 crop: BF16 field vs FP64(f32 grid) max {e_bf16:.2e} rms {rms_bf16:.2e}; "
          f"vs FP64(bf16 grid, bf16 decoder inputs) max {e_self:.2e}", flush=True)
    assert np.all(vol[node < 0] == 0)
    assert e_self <= 1e-3 and e_bf16 <= 3e-2 and rms_bf16 <= 1e-2
