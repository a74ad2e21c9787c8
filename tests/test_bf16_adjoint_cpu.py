"""The oracle's BF16-decoder mode (forward and adjoint) pinned on the CPU.

The device's BF16 decoder (v_mfma_f32_16x16x16_bf16 / 16x16x32_bf16 with
FP32 accumulation) consumes bf16 operands at exactly these points:

  forward   h1 = relu(bf(W1) bf(x) + b1)         h2 = relu(bf(W2) bf(h1) + b2)
            y  = W3 h2 + b3 (FP32), density = softplus_beta(y)
  adjoint   dy = dL/ddensity * softplus'(y)      dh2 = [h2 > 0] dy W3
            dW3 += dy h2, db3 += dy, db2 += dh2   (no rounding)
            dW2 += bf(dh2) bf(h1)^T               dh1 = [h1 > 0] bf(W2)^T bf(dh2)
            db1 += dh1                            dx  = bf(W1)^T bf(dh1)
            dW1 += bf(dh1) bf(x)^T

This file restates that in numpy FP64 (vectorised over samples, written
independently of the C oracle) and requires the oracle's decoder_bf16 mode
to agree with it to 1e-6, both on feature rows (atdfo_decode_features) and
composed with traversal, gather and the trilinear scatter
(atdfo_project_fwd / atdfo_project_bwd with decoder_bf16).  Each rounding
point is then removed from the restatement in turn (a planted mistake, on
the CPU): every such variant must move the result by far more than the
tolerance, so the check fails if the oracle rounds in the wrong place.
"""
import numpy as np
import pytest

from bf16_emulation import (DEC_GRAD_L2_TOL, GRID_GRAD_MAX_TOL, GRID_GRAD_Q90_TOL, ROUND_POINTS, composition,
                            block_errors, decoder_bf16, gpu_case_arrays, gpu_case_dLdP, grad_errors)


def _params(rng, Cc, H, scale=0.4):
    n = H * Cc + H + H * H + H + H + 1
    return (rng.normal(size=n) * scale).astype(np.float32)


def _maxrel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


@pytest.mark.parametrize("Cc", [16, 8])
def test_decode_features_bf16_adjoint_pinned(oracle, Cc):
    rng = np.random.default_rng(100 + Cc)
    H, beta = 64, 2.0
    params = _params(rng, Cc, H)
    x = rng.normal(size=(512, Cc)).astype(np.float32)
    dd = rng.normal(size=512)
    dens_o, gx_o, gm_o = oracle.decode_features(params, Cc, H, beta, x, dd, bf=True)
    dens, gx, gm = decoder_bf16(params, Cc, H, beta, x, dd)
    e = (_maxrel(dens_o, dens), _maxrel(gx_o, gx), _maxrel(gm_o, gm))
    print(f"C={Cc}: oracle bf16 mode vs restatement: density {e[0]:.1e}, dx {e[1]:.1e}, dparams {e[2]:.1e}")
    assert max(e) <= 1e-6
    # the mode is not the FP64 decoder
    _, gx64, gm64 = oracle.decode_features(params, Cc, H, beta, x, dd)
    assert _maxrel(gx64, gx) > 1e-3 and _maxrel(gm64, gm) > 1e-3
    # each rounding point matters: leaving one out moves dx or dparams by
    # far more than the 1e-6 the oracle is held to
    for pt in ROUND_POINTS:
        _, gxm, gmm = decoder_bf16(params, Cc, H, beta, x, dd, skip=(pt,))
        moved = max(_maxrel(gxm, gx), _maxrel(gmm, gm))
        print(f"  without rounding {pt:7s}: moves the adjoint by {moved:.1e}")
        assert moved > 1e-4, pt


def _bf16_case(oracle, seed=0, Cc=16, depth=3, G=11, K=6, nrays=96):
    from scene_util import adaptive_active, tilt_rays
    from adaptivetensordf_amd import phantom

    t = oracle.tree_build(depth)
    act = adaptive_active(t, budget=300, seed=seed)
    na = int(act.sum())
    grid = phantom.init_grid(na, G, Cc, std=0.3, seed=seed).numpy()
    mlp = phantom.init_mlp(Cc, 64, seed=seed).numpy()
    o, d = tilt_rays(nrays, 7 + seed)
    fm = oracle.FusedModel(t["node_min"], t["node_max"], act, grid, mlp, Cc, K=K, jitter_mode=2, seed=99 + seed,
                           decoder_bf16=True)
    return t, act, grid, mlp, o, d, fm


@pytest.mark.parametrize("seed", [0, 1])
def test_project_bf16_oracle_pinned(oracle, seed):
    t, act, grid, mlp, o, d, fm = _bf16_case(oracle, seed=seed)
    dLdP = np.random.default_rng(seed).normal(size=o.shape[0])
    P_o = fm.project(o, d, precision=1)
    gg_o, gm_o = fm.project_bwd(o, d, dLdP)
    P, gg, gm = composition(oracle, t, act, grid, mlp, o, d, 6, 99 + seed, dLdP)
    e = (_maxrel(P_o, P), _maxrel(gg_o, gg), _maxrel(gm_o, gm))
    print(f"seed {seed}: fused oracle (decoder_bf16) vs composition: P {e[0]:.1e}, grid grad {e[1]:.1e}, "
          f"decoder grad {e[2]:.1e}")
    assert max(e) <= 1e-6
    for pt in ("dh2", "dh1", "W2T", "W1T"):
        _, ggm, gmm = composition(oracle, t, act, grid, mlp, o, d, 6, 99 + seed, dLdP, skip=(pt,))
        moved = max(_maxrel(ggm, gg), _maxrel(gmm, gm))
        print(f"  without rounding {pt}: moves the adjoint by {moved:.1e}")
        assert moved > 1e-4, pt


@pytest.mark.parametrize("C_,grid_bf16", [(16, False), (16, True), (8, True)])
def test_planted_mistakes_exceed_gpu_tolerances(oracle, C_, grid_bf16):
    """On the exact inputs of the GPU BF16-decoder check, leaving out any one
    adjoint rounding point moves the decoder gradient or the grid gradient
    by several times the tolerance the GPU is held to against the oracle."""
    t, act, grid, mlp, o, d, K, js = gpu_case_arrays(oracle, C_, grid_bf16)
    dLdP = gpu_case_dLdP(C_, o.shape[0])
    _, gg, gm = composition(oracle, t, act, grid, mlp, o, d, K, js, dLdP)
    fm = oracle.FusedModel(t["node_min"], t["node_max"], act, grid, mlp, C_, K=K, jitter_mode=2, seed=js,
                           decoder_bf16=True)
    gg_o, gm_o = fm.project_bwd(o, d, dLdP)
    assert max(_maxrel(gg_o, gg), _maxrel(gm_o, gm)) <= 1e-6
    for pt in ("dh2", "dh1", "W2T", "W1T", "h1_gW2", "x_gW1"):
        _, ggm, gmm = composition(oracle, t, act, grid, mlp, o, d, K, js, dLdP, skip=(pt,))
        l2 = block_errors(gmm, gm, C_)[1]
        e_dec = max(e / tol for e, tol in zip(l2, DEC_GRAD_L2_TOL))
        e_max, e_q90 = grad_errors(ggm, gg)
        print(f"C={C_} bf16 grid={grid_bf16} without {pt:6s}: decoder grad L2 / tolerance (worst block) {e_dec:.1f}, "
              f"grid grad max {e_max:.1e} q90 {e_q90:.1e}")
        assert e_dec > 3 or (e_max > 3 * GRID_GRAD_MAX_TOL and e_q90 > 3 * GRID_GRAD_Q90_TOL), pt
