"""Synthetic inputs of the reconstruction benchmarks: ellipsoid phantoms,
parallel-beam tilt-series rays, analytic line-integral targets, adaptive
octree selection and field initialisation.

This is the input producer around the hot path (the reference reads a scene
directory and generates rays in SceneBase::GetRays,
hyperacorn/csrc/data/SceneBase.cpp:369-449); here everything is generated
from a fixed splitmix64 stream so every run and the oracle see the same
inputs.

Conventions (orthographic branch of SceneBase.cpp:392-437):
  tilt k: theta_k = -60 + 2 k degrees, rotation about the y axis
  pixel (x, y) of a W x H detector: p = (2 x / (W - 1) - 1, 2 y / (H - 1) - 1, 0)
  direction = R (0, 0, 1), origin = R p - sqrt(2) direction   (FP64, then f32)
"""
from __future__ import annotations

import heapq
import math

import numpy as np

MASK64 = (1 << 64) - 1


class SplitMix64:
    """splitmix64 generator; uniform() returns 53-bit doubles in [0, 1)."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, n: int) -> np.ndarray:
        return np.array([(self.next() >> 11) * (1.0 / (1 << 53)) for _ in range(n)], dtype=np.float64)

    def normal(self, n: int) -> np.ndarray:
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log1p(-u1))
        z = np.concatenate([r * np.cos(2 * math.pi * u2), r * np.sin(2 * math.pi * u2)])
        return z[:n]


def ellipsoids(num: int, seed: int) -> np.ndarray:
    """[num, 7] = centre (3), semi-axes (3), density; centres U[-0.8,0.8]^3,
    semi-axes U(0.05,0.3), density U(0.2,1.0), axis-aligned."""
    rng = SplitMix64(seed)
    out = np.empty((num, 7), dtype=np.float64)
    for i in range(num):
        u = rng.uniform(7)
        out[i, 0:3] = -0.8 + 1.6 * u[0:3]
        out[i, 3:6] = 0.05 + 0.25 * u[3:6]
        out[i, 6] = 0.2 + 0.8 * u[6]
    return out


def tilt_angles(num_tilts: int = 61, first_deg: float = -60.0, step_deg: float = 2.0) -> np.ndarray:
    return np.deg2rad(first_deg + step_deg * np.arange(num_tilts, dtype=np.float64))


def rays_for_pixels(tilt: np.ndarray, px: np.ndarray, py: np.ndarray, W: int, H: int, angles: np.ndarray):
    """Parallel-beam rays for pixel indices; FP64 math, returned as float32 [R,3]."""
    th = angles[tilt]
    c, s = np.cos(th), np.sin(th)
    p0 = 2.0 * px / (W - 1) - 1.0
    p1 = 2.0 * py / (H - 1) - 1.0
    d = np.stack([s, np.zeros_like(s), c], axis=1)
    rp = np.stack([c * p0, p1, -s * p0], axis=1)
    o = rp - math.sqrt(2.0) * d
    return o.astype(np.float32), d.astype(np.float32)


def _interleave_bits(v: np.ndarray, bits: int) -> np.ndarray:
    out = np.zeros_like(v)
    for b in range(bits):
        out |= ((v >> b) & 1) << (2 * b)
    return out


def sample_pixels(num_rays: int, num_tilts: int, W: int, H: int, seed: int, order: str = "row"):
    """A random batch of detector pixels (like the reference's random-pixel
    loader).  Sampling is with replacement.  order "row": sorted by (tilt,
    y, x), so adjacent rays are adjacent pixels of a row; "tile": sorted by
    tilt, then by the Z-order (bit-interleaved y, x) of the pixel, so the 64
    rays of one wave cover a compact patch of the detector and walk mostly
    the same octree nodes.  The same pixels are drawn in either order."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lin = rng.integers(0, num_tilts * H * W, size=num_rays, dtype=np.int64)
    lin.sort()
    tilt = lin // (H * W)
    rem = lin % (H * W)
    px, py = rem % W, rem // W
    if order == "tile":
        bits = int(max(W, H) - 1).bit_length()
        key = (tilt << (2 * bits)) | (_interleave_bits(py, bits) << 1) | _interleave_bits(px, bits)
        idx = np.argsort(key, kind="stable")
        tilt, px, py = tilt[idx], px[idx], py[idx]
    elif order != "row":
        raise ValueError(f"unknown pixel order {order!r}")
    return tilt, px, py


def analytic_projection(o: np.ndarray, d: np.ndarray, ell: np.ndarray) -> np.ndarray:
    """Line integrals of the ellipsoid phantom clipped to [-1,1]^3 (FP64).
    Overlapping ellipsoids add."""
    o = o.astype(np.float64)
    d = d.astype(np.float64)
    dn = np.linalg.norm(d, axis=1)
    # ray interval inside the cube
    with np.errstate(divide="ignore", invalid="ignore"):
        t1 = (-1.0 - o) / d
        t2 = (1.0 - o) / d
    tlo = np.where(d == 0, np.where(np.abs(o) <= 1, -np.inf, np.inf), np.minimum(t1, t2))
    thi = np.where(d == 0, np.where(np.abs(o) <= 1, np.inf, -np.inf), np.maximum(t1, t2))
    c0 = tlo.max(axis=1)
    c1 = thi.min(axis=1)
    out = np.zeros(o.shape[0], dtype=np.float64)
    for e in ell:
        ctr, ax, rho = e[0:3], e[3:6], e[6]
        q0 = (o - ctr) / ax
        q1 = d / ax
        A = (q1 * q1).sum(1)
        B = 2 * (q0 * q1).sum(1)
        Cc = (q0 * q0).sum(1) - 1
        disc = B * B - 4 * A * Cc
        ok = disc > 0
        sq = np.sqrt(np.where(ok, disc, 0))
        e0 = (-B - sq) / (2 * A)
        e1 = (-B + sq) / (2 * A)
        lo = np.maximum(e0, c0)
        hi = np.minimum(e1, c1)
        seg = np.where(ok & (hi > lo), hi - lo, 0.0)
        out += rho * seg * dn
    return out


def analytic_projection_torch(o, d, ell, chunk: int = 1 << 16):
    """Same as analytic_projection on torch tensors (any device), FP64."""
    import torch

    o = o.to(torch.float64)
    d = d.to(torch.float64)
    e = torch.as_tensor(ell, dtype=torch.float64, device=o.device)
    out = torch.empty(o.shape[0], dtype=torch.float64, device=o.device)
    for s in range(0, o.shape[0], chunk):
        oo, dd = o[s:s + chunk], d[s:s + chunk]
        dn = dd.norm(dim=1)
        inv = torch.where(dd == 0, torch.full_like(dd, float("inf")), 1.0 / dd)
        t1 = (-1.0 - oo) * inv
        t2 = (1.0 - oo) * inv
        par_in = (dd == 0) & (oo.abs() <= 1)
        par_out = (dd == 0) & (oo.abs() > 1)
        tlo = torch.minimum(t1, t2)
        thi = torch.maximum(t1, t2)
        tlo = torch.where(par_in, torch.full_like(tlo, -float("inf")), tlo)
        thi = torch.where(par_in, torch.full_like(thi, float("inf")), thi)
        tlo = torch.where(par_out, torch.full_like(tlo, float("inf")), tlo)
        thi = torch.where(par_out, torch.full_like(thi, -float("inf")), thi)
        c0 = tlo.amax(1, keepdim=True)
        c1 = thi.amin(1, keepdim=True)
        ctr, ax, rho = e[:, 0:3], e[:, 3:6], e[:, 6]
        q0 = (oo[:, None, :] - ctr[None]) / ax[None]
        q1 = dd[:, None, :] / ax[None]
        A = (q1 * q1).sum(2)
        B = 2 * (q0 * q1).sum(2)
        Cc = (q0 * q0).sum(2) - 1
        disc = B * B - 4 * A * Cc
        ok = disc > 0
        sq = torch.sqrt(torch.clamp(disc, min=0))
        lo = torch.maximum((-B - sq) / (2 * A), c0)
        hi = torch.minimum((-B + sq) / (2 * A), c1)
        seg = torch.where(ok & (hi > lo), hi - lo, torch.zeros_like(hi))
        out[s:s + chunk] = (seg * rho[None]).sum(1) * dn
    return out


def adaptive_active_mask(node_min: np.ndarray, node_max: np.ndarray, children: np.ndarray, depth: np.ndarray,
                         ell: np.ndarray, budget: int, start_depth: int, max_depth: int) -> np.ndarray:
    """Choose an adaptive active-node set: start uniform at start_depth and
    repeatedly split the node with the largest structure score (number of
    ellipsoid bounding boxes cut by the node, times node volume) while the
    active count stays within budget and depth <= max_depth."""
    lo = ell[:, 0:3] - ell[:, 3:6]
    hi = ell[:, 0:3] + ell[:, 3:6]

    def score(n):
        mn, mx = node_min[n].astype(np.float64), node_max[n].astype(np.float64)
        overlap = np.all((lo < mx) & (hi > mn), axis=1)
        inside = np.all((lo <= mn) & (hi >= mx), axis=1)
        cut = int(np.count_nonzero(overlap & ~inside))
        vol = float(np.prod(mx - mn))
        return cut * vol

    active = np.zeros(node_min.shape[0], dtype=np.int32)
    heap = []
    for n in np.nonzero(depth == start_depth)[0]:
        active[n] = 1
        s = score(n)
        if s > 0 and depth[n] < max_depth:
            heapq.heappush(heap, (-s, int(n)))
    count = int(active.sum())
    grow = children.shape[1] - 1  # nodes added by one split
    while heap and count + grow <= budget:
        _, n = heapq.heappop(heap)
        active[n] = 0
        for c in children[n]:
            active[c] = 1
            s = score(c)
            if s > 0 and depth[c] < max_depth:
                heapq.heappush(heap, (-s, int(c)))
        count += grow
    return active


def mlp_num_params(C: int, H: int = 64) -> int:
    return H * C + H + H * H + H + H + 1


def init_mlp(C: int, H: int = 64, seed: int = 0):
    """Decoder parameters [W1, b1, W2, b2, W3, b3]: Kaiming-normal weights
    (fan_in, ReLU gain; Modules.h:33) and torch Linear default biases
    U(-1/sqrt(fan_in), 1/sqrt(fan_in))."""
    import torch

    g = torch.Generator().manual_seed(seed)
    parts = []
    for fan_in, fan_out in ((C, H), (H, H), (H, 1)):
        w = torch.randn(fan_out, fan_in, generator=g, dtype=torch.float32) * math.sqrt(2.0 / fan_in)
        b = (torch.rand(fan_out, generator=g, dtype=torch.float32) * 2 - 1) / math.sqrt(fan_in)
        parts += [w.reshape(-1), b]
    return torch.cat(parts)


PROPOSAL_BINS, PROPOSAL_INPUTS, PROPOSAL_HIDDEN = 32, 38, 64


def proposal_num_params() -> int:
    n, h, b = PROPOSAL_INPUTS, PROPOSAL_HIDDEN, PROPOSAL_BINS
    return h * n + h + 2 * (h * h + h) + b * h + b


def init_proposal(seed: int = 0):
    """Proposal MLP parameters [W1, b1, W2, b2, W3, b3, W4, b4] for
    38 -> 64 -> 64 -> 64 -> 32, initialised like the decoder."""
    import torch

    g = torch.Generator().manual_seed(seed)
    n, h, b = PROPOSAL_INPUTS, PROPOSAL_HIDDEN, PROPOSAL_BINS
    parts = []
    for fan_in, fan_out in ((n, h), (h, h), (h, h), (h, b)):
        w = torch.randn(fan_out, fan_in, generator=g, dtype=torch.float32) * math.sqrt(2.0 / fan_in)
        bb = (torch.rand(fan_out, generator=g, dtype=torch.float32) * 2 - 1) / math.sqrt(fan_in)
        parts += [w.reshape(-1), bb]
    return torch.cat(parts)


def init_grid(num_grids: int, G: int, C: int, std: float = 0.01, seed: int = 0, device="cpu"):
    """Channels-last node grids [N, G, G, G, C] ~ N(0, std^2)."""
    import torch

    g = torch.Generator(device=device).manual_seed(seed)
    return torch.randn((num_grids, G, G, G, C), generator=g, dtype=torch.float32, device=device) * std
