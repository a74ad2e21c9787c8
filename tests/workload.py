"""Shapes of the benchmark configurations for the parity tests that run at
them: the 512^3 workload's depth-5 adaptive octree (8192-node budget, 8191
active nodes, refined toward a 200-ellipsoid phantom, seed 1), 61 tilts
from -60 to +60 degrees on a 512^2 (or 768^2) detector, K = 12 (about 128
stratified samples per ray), C = 16 channels, 11^3 node grids and the
16-64-64-1 decoder."""
import numpy as np

from adaptivetensordf_amd import phantom

DEPTH, BUDGET, C, G, H, K = 5, 8192, 16, 11, 64, 12
TILTS = 61


def tree_arrays(oracle, budget=BUDGET, seed=1):
    t = oracle.tree_build(DEPTH)
    ell = phantom.ellipsoids(200, seed)
    act = phantom.adaptive_active_mask(t["node_min"], t["node_max"], t["children"], t["depth"], ell, budget, 3,
                                       DEPTH).astype(np.int32)
    return t, act, ell


def rays(n, seed, det=512):
    """n random detector pixels of the tilt series (sorted by tilt, row,
    column) and their rays (f32)."""
    ang = phantom.tilt_angles(TILTS)
    tl, px, py = phantom.sample_pixels(n, TILTS, det, det, seed)
    return phantom.rays_for_pixels(tl, px, py, det, det, ang)


def window_rays(n, seed, det=512, lo=128, hi=384):
    """n random pixels of the detector window [lo, hi)^2 over all tilts: the
    rays through a centred crop of the volume."""
    rng = np.random.default_rng(seed)
    tl = rng.integers(0, TILTS, n)
    px = rng.integers(lo, hi, n)
    py = rng.integers(lo, hi, n)
    return phantom.rays_for_pixels(tl, px, py, det, det, phantom.tilt_angles(TILTS))


def chord_through_root(o, d):
    """FP64 length of each ray's chord through [-1, 1]^3 (|d| = 1 units).
    A ray parallel to an axis is inside the slab iff -1 <= o < 1, the
    half-open rule of the reference's slab test (a ray on the upper face
    misses)."""
    o, d = o.astype(np.float64), d.astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        t1 = (-1.0 - o) / d
        t2 = (1.0 - o) / d
    inside = (o >= -1.0) & (o < 1.0)
    lo = np.where(d == 0, np.where(inside, -np.inf, np.inf), np.minimum(t1, t2)).max(1)
    hi = np.where(d == 0, np.where(inside, np.inf, -np.inf), np.maximum(t1, t2)).min(1)
    return np.maximum(hi - lo, 0.0) * np.linalg.norm(d, axis=1)
