"""DecodeFeatures on feature rows and the Fourier regulariser on the GPU
against the oracle: the decoder forward and adjoint (FP32 MFMA within 1e-5
of FP64, BF16 MFMA against the BF16-emulating FP64 decoder), determinism,
the ring means (FP64 sums, within 1e-5 of the pointwise oracle), the whole
term and its gradient (torch::fft on the device, against the numpy FFT
oracle and its finite differences), the term's lattice, and the term inside
VolumeRegularizer on a tensorial field."""
import numpy as np
import pytest
import torch

from test_gpu_vm import _vm_setup

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)


def _fp(native, Cc, bf16=False):
    fp = native.FieldParams()
    fp.C, fp.H, fp.softplus_beta, fp.bf16_decoder = Cc, 64, 2.0, bf16
    return fp


def _mlp(rng, Cc, H=64, scale=0.4):
    return (rng.normal(size=H * Cc + H + H * H + H + H + 1) * scale).astype(np.float32)


@pytest.mark.parametrize("Cc,bf16", [(16, False), (8, False), (16, True)])
def test_decode_features_matches_oracle(oracle, native, cuda, Cc, bf16):
    rng = np.random.default_rng(Cc + bf16)
    S = 4099
    x = rng.normal(size=(S, Cc)).astype(np.float32)
    mlp = _mlp(rng, Cc)
    dd = rng.normal(size=S).astype(np.float32)
    fp = _fp(native, Cc, bf16)
    print(f"decode C={Cc} bf16={bf16} S={S}", flush=True)
    xt = torch.from_numpy(x).to(cuda).requires_grad_()
    mt = torch.from_numpy(mlp).to(cuda).requires_grad_()
    y = native.DecodeFeatures(xt, mt, fp)
    assert y.shape == (S, 1)
    gx, gm = torch.autograd.grad(y[:, 0], (xt, mt), torch.from_numpy(dd).to(cuda))
    dens, ogx, ogm = oracle.decode_features(mlp, Cc, 64, 2.0, x, dd.astype(np.float64), bf=bf16)
    got = y[:, 0].detach().cpu().double().numpy()
    tol = 1e-5 if not bf16 else 1e-4
    print(f"  density rel {_rel(got, dens):.2e}", flush=True)
    assert _rel(got, dens) <= tol
    if not bf16:
        assert _rel(gx.cpu().double().numpy(), ogx) <= 1e-5
        assert _rel(gm.cpu().double().numpy(), ogm) <= 1e-5
    else:
        # the device rounds the backward's MFMA operands too; FP32 dy and
        # accumulation order decide near-ties of the rounding
        assert np.linalg.norm(gx.cpu().double().numpy() - ogx) <= 2e-2 * np.linalg.norm(ogx)
        assert np.linalg.norm(gm.cpu().double().numpy() - ogm) <= 2e-2 * np.linalg.norm(ogm)
    # deterministic: the same call gives the same bits
    y2 = native.DecodeFeatures(xt, mt, fp)
    gx2, gm2 = torch.autograd.grad(y2[:, 0], (xt, mt), torch.from_numpy(dd).to(cuda))
    assert torch.equal(y, y2) and torch.equal(gx, gx2) and torch.equal(gm, gm2)
    assert native.DecodeFeatures(xt[:0], mt, fp).shape == (0, 1)


@pytest.mark.parametrize("Y,X,Z", [(5, 17, 17), (3, 33, 21), (2, 9, 9), (4, 65, 65)])
def test_fourier_ring_means(oracle, native, cuda, Y, X, Z):
    rng = np.random.default_rng(X * Z)
    img = rng.uniform(0, 1, (Y, X, Z)).astype(np.float32)
    print(f"rings Y={Y} X={X} Z={Z}", flush=True)
    got = native.FourierRingMeans(torch.from_numpy(img).to(cuda)).cpu().numpy()
    want = oracle.fourier_ring_means(img)
    assert got.shape == want.shape == (X // 2, Y)
    assert np.all(np.isinf(got[0]))
    if X // 2 > 1:
        assert _rel(got[1:], want[1:]) <= 1e-5


def _loss(native, d, cuda, rate=0.2, rate_w=0.01, lf=1.0):
    fp = native.FourierParams()
    fp.loss_fourier, fp.rate, fp.rate_w = lf, rate, rate_w
    t = torch.from_numpy(np.ascontiguousarray(d, np.float32)).to(cuda).requires_grad_()
    loss = native.FourierLoss(t, fp)
    (g,) = torch.autograd.grad(loss, t)
    return loss.item(), g.cpu().double().numpy()


@pytest.mark.parametrize("rate_w", [0.01, 0.2])
def test_fourier_loss_and_gradient(oracle, native, cuda, rate_w):
    rng = np.random.default_rng(5)
    d = rng.uniform(0.0, 1.0, (33, 7, 33)).astype(np.float32)
    print(f"fourier loss 33x7x33 rate_w={rate_w}", flush=True)
    got, g = _loss(native, d, cuda, rate_w=rate_w, lf=0.5)
    want, parts = oracle.fourier_terms(d.astype(np.float64), rate_w=rate_w)
    want *= 0.5
    print(f"  loss {got:.8e} oracle {want:.8e}", flush=True)
    assert abs(got - want) <= 1e-4 * abs(want)
    # central differences of the FP64 oracle at a few entries, the ring
    # profile held fixed (it is detached)
    ring = parts["ring"]
    h = 1e-4
    idx = [tuple(rng.integers(0, s) for s in d.shape) for _ in range(12)]
    for ix in idx:
        dp = d.astype(np.float64)
        dm = dp.copy()
        dp[ix] += h
        dm[ix] -= h
        fd = 0.5 * (oracle.fourier_terms(dp, rate_w=rate_w, ring=ring)[0] -
                    oracle.fourier_terms(dm, rate_w=rate_w, ring=ring)[0]) / (2 * h)
        assert abs(g[ix] - fd) <= 1e-3 * np.abs(g).max() + 1e-7, ix
    # invariances: a positive scale, a circular shift along x
    s2, _ = _loss(native, 2.0 * d, cuda, rate_w=rate_w, lf=0.5)
    sh, _ = _loss(native, np.roll(d, 5, axis=0), cuda, rate_w=rate_w, lf=0.5)
    assert abs(s2 - 2 * got) <= 1e-5 * abs(got) and abs(sh - got) <= 1e-5 * abs(got)


def test_fourier_loss_rejects_non_square(native, cuda):
    fp = native.FourierParams()
    fp.loss_fourier = 1.0
    with pytest.raises(RuntimeError, match="must equal Z"):
        native.FourierLoss(torch.zeros(9, 3, 11, device=cuda), fp)


def test_make_fourier_grid(oracle, native, cuda):
    t, act, tree, roi, f, arr, tr, lohi = _vm_setup(native, oracle, "add", seed=2)
    fp = native.FourierParams()
    fp.shift, fp.scale, fp.scale_y = 0.05, 1.0, 0.5
    lo, hi = [-0.5, -0.6, -0.5], [0.5, 0.4, 0.5]
    g = native.MakeFourierGrid(tree, 5, 3, lo, hi, fp)
    # (grid_size - 1) 2^depth = 32; x = round(32 * 0.5) = 16 -> odd 17; y = (int)(0.5 * round(32 * 0.9 / 2)) = 7; z like x
    assert (g.X, g.Y, g.Z) == (17, 7, 17)
    c = g.coords.cpu().numpy().reshape(17, 7, 17, 3)
    assert np.allclose(c[:, 0, 0, 0], np.linspace(-0.5, 0.5, 17), atol=1e-7)
    assert np.allclose(c[0, :, 0, 1], np.linspace(-0.55, 0.35, 7), atol=1e-7)
    assert np.allclose(c[0, 0, :, 2], np.linspace(-0.5, 0.5, 17), atol=1e-7)
    nid, _ = tree.NodeIdForPositionGPU(g.coords)
    assert torch.equal(nid, g.node_id)


def test_volume_regularizer_fourier_term(oracle, native, cuda):
    t, act, tree, roi, f, arr, tr, lohi = _vm_setup(native, oracle, "add", seed=4)
    rng = np.random.default_rng(9)
    fpp = native.FourierParams()
    fpp.loss_fourier = 0.25
    g = native.MakeFourierGrid(tree, 5, 3, [-0.5, -0.5, -0.5], [0.5, 0.5, 0.5], fpp)
    print(f"fourier lattice {g.X}x{g.Y}x{g.Z}", flush=True)
    p = native.RegularizerParams()
    p.loss_tv, p.loss_l2, p.loss_edge = 0.0, -1.0, 0.0
    p.fourier = fpp
    fp = _fp(native, 8)
    mlp = torch.from_numpy(_mlp(rng, 8, scale=0.3)).to(cuda).requires_grad_()
    loss = native.VolumeRegularizer(tree, roi, f, p, None, g, mlp, fp)
    # the composition: decoded densities of the lattice, then the term
    feat = native.SampleVolumeIndirectFeature(tree, roi, f, g.coords, g.node_id)
    dens = native.DecodeFeatures(feat, mlp, fp).reshape(g.X, g.Y, g.Z)
    ref = native.FourierLoss(dens, fpp)
    assert loss.item() == ref.item()
    # against the oracle on the same features
    od, _, _ = oracle.decode_features(mlp.detach().cpu().numpy(), 8, 64, 2.0, feat.detach().cpu().numpy())
    want = 0.25 * oracle.fourier_terms(od.reshape(g.X, g.Y, g.Z))[0]
    print(f"  loss {loss.item():.8e} oracle {want:.8e}", flush=True)
    assert abs(loss.item() - want) <= 1e-4 * abs(want)
    names = ["plane_yz", "plane_xz", "plane_xy", "line_x", "line_y", "line_z"]
    grads = torch.autograd.grad(loss, [getattr(f, k) for k in names] + [mlp], allow_unused=True)
    for k, gr in zip(names + ["mlp"], grads):
        assert gr is not None and torch.isfinite(gr).all() and gr.abs().max() > 0, k
    # the term is off without its lattice or with loss_fourier <= 0
    assert native.VolumeRegularizer(tree, roi, f, p) is None


def test_compute_image_vm(oracle, native, cuda):
    """ComputeImage of the tensorial geometry: traversal samples, VM features,
    decoder, line integral; against the oracle composition in FP64, and the
    decoder-parameter gradient against the oracle's adjoint."""
    from test_gpu_vm import _vm_oracle
    from conftest import random_rays

    t, act, tree, roi, f, arr, tr, lohi = _vm_setup(native, oracle, "add", seed=6)
    o, d = random_rays(300, 46, spread=0.9)
    rays = native.RayList(torch.from_numpy(o).to(cuda), torch.from_numpy(d).to(cuda))
    samples = tree.CreateSamplesForRays(rays, 4, False)
    rng = np.random.default_rng(12)
    mlp_np = _mlp(rng, 8, scale=0.3)
    mlp = torch.from_numpy(mlp_np).to(cuda).requires_grad_()
    fp = _fp(native, 8)
    img = native.ComputeImageVM(tree, roi, f, mlp, fp, samples, 300)
    assert img.shape == (1, 300)
    feats, _ = _vm_oracle(oracle, t, act, roi, arr, tr, lohi, "add")
    assert np.array_equal(samples.node_id.cpu().numpy(), tr["node_id"])
    dens, _, _ = oracle.decode_features(mlp_np, 8, 64, 2.0, feats)
    want = np.zeros(300)
    np.add.at(want, tr["ray_index"], dens * tr["weight"].astype(np.float64))
    got = img[0].detach().cpu().double().numpy()
    print(f"VM image: {tr['node_id'].size} samples, rel {_rel(got, want):.2e}", flush=True)
    assert _rel(got, want) <= 1e-5
    go = rng.normal(size=300)
    (gm,) = torch.autograd.grad(img[0], mlp, torch.from_numpy(go.astype(np.float32)).to(cuda))
    dd = go[tr["ray_index"]] * tr["weight"].astype(np.float64)
    _, _, ogm = oracle.decode_features(mlp_np, 8, 64, 2.0, feats, dd)
    assert _rel(gm.cpu().double().numpy(), ogm) <= 1e-5
