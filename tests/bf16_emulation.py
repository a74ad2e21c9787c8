"""BF16-decoder emulation shared by the CPU pin of the oracle's BF16 mode
and the GPU checks against it: an FP64 numpy restatement of the decoder and
its adjoint that rounds to bf16 exactly the operands the device's BF16 MFMA
consumes, the composition of traversal, gather, that decoder and the
trilinear scatter, and the tolerances the GPU is held to."""
import numpy as np
import torch

# GPU BF16 decoder against the emulating oracle.  An FP32/FP64 difference
# can flip a single bf16 rounding in a rare sample, which moves a few
# entries by up to ~1e-4 of their block; a misplaced rounding point moves
# every entry.  So the decoder gradient is held per parameter block to an
# L2 bound (tight) and a max-norm bound (loose), and the grid gradient to a
# 90th-percentile bound (tight) and a max-norm bound (loose).
# per block W1, b1, W2, b2, W3, b3 (the W1 and W2 sums see the most flips)
DEC_GRAD_L2_TOL = (2e-4, 5e-5, 1e-4, 5e-5, 5e-5, 5e-5)
DEC_GRAD_MAX_TOL = 5e-4
GRID_GRAD_MAX_TOL = 4e-4
GRID_GRAD_Q90_TOL = 5e-6


ROUND_POINTS = ("x", "W1", "h1", "W2", "dh2", "W2T", "dh1", "W1T", "h1_gW2", "x_gW1")


def bf(a):
    """Round to bfloat16 (nearest even) through float32, as the device does."""
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(torch.bfloat16).double().numpy()


def _split(params, Cc, H):
    p = params.astype(np.float64)
    o = 0
    W1 = p[o:o + H * Cc].reshape(H, Cc); o += H * Cc
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, W3, b3


def decoder_bf16(params, Cc, H, beta, x, dd=None, skip=()):
    """density [S], dL/dx [S, C], dL/dparams [NP] of the BF16 decoder;
    `skip` names rounding points to leave out (planted mistakes)."""
    r = lambda name, a: np.asarray(a, np.float64) if name in skip else bf(a)
    W1, b1, W2, b2, W3, b3 = _split(params, Cc, H)
    x = np.asarray(x, np.float32)
    h1p = r("x", x) @ r("W1", W1).T + b1
    h1 = np.maximum(h1p, 0)
    h2p = r("h1", h1) @ r("W2", W2).T + b2
    h2 = np.maximum(h2p, 0)
    y = h2 @ W3 + b3
    dens = np.where(y * beta > 20, y, np.log1p(np.exp(np.minimum(y * beta, 20))) / beta)
    if dd is None:
        return dens, None, None
    sig = np.where(y * beta > 20, 1.0, 1.0 / (1.0 + np.exp(-np.minimum(y * beta, 20))))
    dy = dd * sig
    dh2 = (h2 > 0) * dy[:, None] * W3[None, :]
    gW3 = dy @ h2
    gb3 = dy.sum()
    gb2 = dh2.sum(0)
    gW2 = r("dh2", dh2).T @ r("h1_gW2", h1)
    dh1 = (h1 > 0) * (r("dh2", dh2) @ r("W2T", W2))
    gb1 = dh1.sum(0)
    gW1 = r("dh1", dh1).T @ r("x_gW1", x)
    dx = r("dh1", dh1) @ r("W1T", W1)
    gm = np.concatenate([gW1.reshape(-1), gb1, gW2.reshape(-1), gb2, gW3, [gb3]])
    return dens, dx, gm


def composition(oracle, t, act, grid, mlp, o, d, K, jitter_seed, dLdP, skip=()):
    """traversal -> FP32 local coordinates and gather -> BF16 decoder
    (restated above) -> line integral, and the adjoint back through the
    trilinear weights (FP64, from FP64 local coordinates) into the
    channels-last grid."""
    Cc, G = grid.shape[-1], grid.shape[1]
    ids = oracle.active_ids(act)
    smp = oracle.traverse(t["node_min"], t["node_max"], ids, o, d, K, jitter_mode=2, seed=jitter_seed)
    pos, node, w, ray = smp["global_coordinate"], smp["node_id"], smp["weight"], smp["ray_index"]
    loc32 = oracle.local_samples(t["node_min"], t["node_max"], pos, node)
    slot = oracle.prefix_sum(act)[node]
    ref_layout = np.ascontiguousarray(np.transpose(grid, (0, 4, 1, 2, 3)))  # [N, C, Z, Y, X]
    feat = oracle.grid_sample3d_fwd(ref_layout, slot, loc32)
    R = o.shape[0]
    dd = dLdP[ray] * w.astype(np.float64)
    dens, dx, gm = decoder_bf16(mlp, Cc, 64, 2.0, feat, dd, skip=skip)
    P = np.zeros(R)
    np.add.at(P, ray, dens * w.astype(np.float64))
    # scatter: FP64 trilinear weights, clamped corner indices
    mn, mx = t["node_min"][node].astype(np.float64), t["node_max"][node].astype(np.float64)
    loc = ((pos.astype(np.float64) - mn) / (mx - mn)) * 2 - 1
    fi = (loc + 1) / 2 * (G - 1)
    i0 = np.floor(fi).astype(np.int64)
    f1 = fi - i0
    gg = np.zeros(grid.shape, np.float64).reshape(-1, Cc)
    for k in range(8):
        b = np.array([k & 1, (k >> 1) & 1, (k >> 2) & 1])
        wk = np.prod(np.where(b[None, :] == 1, f1, 1 - f1), axis=1)
        ix = np.clip(i0 + b[None, :], 0, G - 1)
        cell = ((slot * G + ix[:, 2]) * G + ix[:, 1]) * G + ix[:, 0]
        np.add.at(gg, cell, wk[:, None] * dx)
    return P, gg.reshape(grid.shape), gm




def gpu_case_arrays(oracle, C_, grid_bf16, nrays=512, depth=3, G=11, K=6):
    """The oracle-side arrays of the GPU BF16-decoder check (the same tree,
    field, rays and jitter seed as that test builds): tree arrays, active
    mask, channels-last grid (bf16-rounded when grid_bf16), decoder
    parameters, rays, K and the jitter seed."""
    from conftest import random_rays
    from scene_util import adaptive_active, tilt_rays
    from adaptivetensordf_amd import phantom

    seed = 31 + C_
    t = oracle.tree_build(depth)
    act = adaptive_active(t, budget=300, seed=seed)
    grid = phantom.init_grid(int(act.sum()), G, C_, std=0.3, seed=seed)
    if grid_bf16:
        grid = grid.to(torch.bfloat16).float()
    mlp = phantom.init_mlp(C_, 64, seed=seed)
    o1, d1 = random_rays(nrays // 2, 30 + seed)
    o2, d2 = tilt_rays(nrays - nrays // 2, 40 + seed)
    o, d = np.concatenate([o1, o2]), np.concatenate([d1, d2])
    return t, act, grid.numpy(), mlp.numpy(), o, d, K, 1234 + seed


def gpu_case_dLdP(C_, R):
    return np.random.default_rng(C_).normal(size=R)


def grad_errors(g, ref):
    """(max |g - ref| / max|ref|, 90th percentile of the same over the
    entries where ref != 0)."""
    g, ref = np.asarray(g, np.float64), np.asarray(ref, np.float64)
    scale = np.abs(ref).max()
    err = np.abs(g - ref) / scale
    touched = ref != 0
    return float(err.max()), float(np.quantile(err[touched], 0.9)) if touched.any() else 0.0


def block_errors(g, ref, C_, H=64):
    """Per decoder parameter block (W1, b1, W2, b2, W3, b3): the lists of
    max |g - ref| / max |ref| and ||g - ref||_2 / ||ref||_2."""
    sizes = [H * C_, H, H * H, H, H, 1]
    mx, l2, o = [], [], 0
    for n in sizes:
        a, b = np.asarray(g[o:o + n], np.float64), np.asarray(ref[o:o + n], np.float64)
        mx.append(float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)))
        l2.append(float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)))
        o += n
    return mx, l2
