"""numpy front end of the CPU oracle (liboracle.so, built from atdf_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke() check and by
bench.py's cpu_baseline leg.  The product path never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_lib = None

_P = C.c_void_p


def build():
    subprocess.run(["make", "-C", HERE, "-B"], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        _lib.atdfo_tree_num_nodes.restype = C.c_int64
        _lib.atdfo_tree_num_nodes.argtypes = [C.c_int]
        _lib.atdfo_intersect_box_ray.restype = C.c_int
        _lib.atdfo_jitter.restype = C.c_float
        _lib.atdfo_jitter.argtypes = [C.c_uint64, C.c_int64, C.c_int64]
        _lib.atdfo_traverse.restype = C.c_int64
        _lib.atdfo_mlp_num_params.restype = C.c_int64
        _lib.atdfo_project_fwd.restype = C.c_int
        _lib.atdfo_project_bwd.restype = C.c_int
        _lib.atdfo_train_step_f32.restype = C.c_double
        _lib.atdfo_project_fwd_bins.restype = C.c_int
        _lib.atdfo_prop_num_params.restype = C.c_int64
        _lib.atdfo_prop_exp.restype = C.c_float
        _lib.atdfo_prop_exp.argtypes = [C.c_float]
        _lib.atdfo_prop_log.restype = C.c_float
        _lib.atdfo_prop_log.argtypes = [C.c_float]
        _lib.atdfo_locate.restype = C.c_int32
        _lib.atdfo_prop_grad.restype = C.c_double
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_P)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class Model(C.Structure):
    _fields_ = [("node_min", _P), ("node_max", _P), ("active_ids", _P), ("num_active", C.c_int64),
                ("node_to_grid", _P), ("grid", _P), ("G", C.c_int), ("C", C.c_int), ("H", C.c_int),
                ("mlp", _P), ("beta", C.c_float), ("K", C.c_int), ("max_intersections", C.c_int),
                ("jitter_mode", C.c_int), ("seed", C.c_uint64), ("decoder_bf16", C.c_int),
                ("iv_off", _P), ("iv_node", _P), ("iv_tmin", _P), ("iv_tmax", _P), ("iv_count", _P)]


# ------------------------------------------------------------------ tree
def tree_build(max_depth: int, ns: int = 8, y_min: float = -1.0, y_max: float = 1.0):
    """ns = 8: the octree; ns = 4: the quad-tree representation (children
    split x and z, every node spans [y_min, y_max] in y)."""
    L = lib()
    L.atdfo_tree_num_nodes_ns.restype = C.c_int64
    n = L.atdfo_tree_num_nodes_ns(C.c_int(max_depth), C.c_int(ns))
    nmin = np.zeros((n, 3), np.float32)
    nmax = np.zeros((n, 3), np.float32)
    ch = np.zeros((n, ns), np.int32)
    par = np.zeros(n, np.int32)
    dep = np.zeros(n, np.int32)
    L.atdfo_tree_build_ns(C.c_int(max_depth), C.c_int(ns), C.c_float(y_min), C.c_float(y_max), _p(nmin), _p(nmax),
                          _p(ch), _p(par), _p(dep))
    return dict(node_min=nmin, node_max=nmax, children=ch, parent=par, depth=dep)


def active_ids(active: np.ndarray) -> np.ndarray:
    return np.nonzero(active)[0].astype(np.int64)


def prefix_sum(active: np.ndarray) -> np.ndarray:
    out = np.full(active.shape[0], -1, np.int32)
    ids = np.nonzero(active)[0]
    out[ids] = np.arange(ids.shape[0], dtype=np.int32)
    return out


def intersect(bmin, bmax, o, d):
    tmin = C.c_float()
    tmax = C.c_float()
    arrs = [_c(x, np.float32) for x in (bmin, bmax, o, d)]
    hit = lib().atdfo_intersect_box_ray(*[_p(a) for a in arrs], C.byref(tmin), C.byref(tmax))
    return bool(hit), tmin.value, tmax.value


def jitter(seed: int, ray: int, local: int) -> float:
    return lib().atdfo_jitter(C.c_uint64(seed), C.c_int64(ray), C.c_int64(local))


# ------------------------------------------------------------- traversal
def traverse(node_min, node_max, act_ids, o, d, K, max_intersections=256, jitter_mode=0, jitter_arr=None, seed=0):
    """Canonical samples and intervals of a ray batch.  Returns a dict, or
    raises ValueError when a ray exceeds max_intersections."""
    L = lib()
    node_min, node_max = _c(node_min, np.float32), _c(node_max, np.float32)
    act_ids = _c(act_ids, np.int64)
    o, d = _c(o, np.float32), _c(d, np.float32)
    R = o.shape[0]
    ns = np.zeros(R, np.int32)
    nh = np.zeros(R, np.int32)
    common = [_p(node_min), _p(node_max), _p(act_ids), C.c_int64(act_ids.shape[0]), _p(o), _p(d), C.c_int64(R),
              C.c_int(K), C.c_int(max_intersections), C.c_int(jitter_mode),
              _p(None if jitter_arr is None else _c(jitter_arr, np.float32)), C.c_uint64(seed)]
    jarr = None if jitter_arr is None else _c(jitter_arr, np.float32)
    common[10] = _p(jarr)
    tot = L.atdfo_traverse(*common, _p(ns), _p(nh), *([None] * 7), *([None] * 6))
    if tot < 0:
        raise ValueError("too many intersections")
    soff = np.zeros(R + 1, np.int64)
    soff[1:] = np.cumsum(ns)
    ioff = np.zeros(R + 1, np.int64)
    ioff[1:] = np.cumsum(nh)
    S, H = int(soff[-1]), int(ioff[-1])
    out = dict(ray_num_samples=ns, ray_num_intervals=nh, ray_sample_offset=soff, ray_interval_offset=ioff,
               global_coordinate=np.zeros((S, 3), np.float32), weight=np.zeros(S, np.float32),
               ray_t=np.zeros(S, np.float32), ray_index=np.zeros(S, np.int64),
               local_index_in_ray=np.zeros(S, np.int64), node_id=np.zeros(S, np.int64),
               iv_node=np.zeros(H, np.int32), iv_tmin=np.zeros(H, np.float32), iv_tmax=np.zeros(H, np.float32),
               iv_count=np.zeros(H, np.int32), iv_first=np.zeros(H, np.int32))
    L.atdfo_traverse(*common, None, None, _p(soff), _p(out["global_coordinate"]), _p(out["weight"]),
                     _p(out["ray_t"]), _p(out["ray_index"]), _p(out["local_index_in_ray"]), _p(out["node_id"]),
                     _p(ioff), _p(out["iv_node"]), _p(out["iv_tmin"]), _p(out["iv_tmax"]), _p(out["iv_count"]),
                     _p(out["iv_first"]))
    return out


def local_samples(node_min, node_max, g, node):
    g = _c(g, np.float32)
    node = _c(node, np.int64)
    out = np.zeros_like(g)
    lib().atdfo_local_samples(_p(_c(node_min, np.float32)), _p(_c(node_max, np.float32)), _p(g), _p(node),
                              C.c_int64(node.shape[0]), _p(out))
    return out


# ---------------------------------------------------------------- gather
def grid_sample3d_fwd(grid, index, uv):
    grid = _c(grid, np.float32)
    N, Cc, Z, Y, X = grid.shape
    index = _c(index, np.int64)
    uv = _c(uv, np.float32)
    out = np.zeros((index.shape[0], Cc), np.float32)
    lib().atdfo_grid_sample3d_fwd(_p(grid), C.c_int64(N), C.c_int(Cc), C.c_int(Z), C.c_int(Y), C.c_int(X),
                                  _p(index), _p(uv), C.c_int64(index.shape[0]), _p(out))
    return out


def grid_sample3d_bwd(grid, index, uv, grad_out):
    grid = _c(grid, np.float32)
    N, Cc, Z, Y, X = grid.shape
    index = _c(index, np.int64)
    uv = _c(uv, np.float32)
    go = _c(grad_out, np.float32)
    gg = np.zeros(grid.shape, np.float64)
    gu = np.zeros(uv.shape, np.float64)
    lib().atdfo_grid_sample3d_bwd(_p(grid), C.c_int64(N), C.c_int(Cc), C.c_int(Z), C.c_int(Y), C.c_int(X),
                                  _p(index), _p(uv), C.c_int64(index.shape[0]), _p(go), _p(gg), _p(gu))
    return gg, gu


def grid_sample2d_fwd(grid, index, uv):
    grid = _c(grid, np.float32)
    N, Cc, H, W = grid.shape
    index, uv = _c(index, np.int64), _c(uv, np.float32)
    out = np.zeros((index.shape[0], Cc), np.float32)
    lib().atdfo_grid_sample2d_fwd(_p(grid), C.c_int64(N), C.c_int(Cc), C.c_int(H), C.c_int(W), _p(index), _p(uv),
                                  C.c_int64(index.shape[0]), _p(out))
    return out


def grid_sample2d_bwd(grid, index, uv, grad_out):
    grid = _c(grid, np.float32)
    N, Cc, H, W = grid.shape
    index, uv, go = _c(index, np.int64), _c(uv, np.float32), _c(grad_out, np.float32)
    gg = np.zeros(grid.shape, np.float64)
    gu = np.zeros(uv.shape, np.float64)
    lib().atdfo_grid_sample2d_bwd(_p(grid), C.c_int64(N), C.c_int(Cc), C.c_int(H), C.c_int(W), _p(index), _p(uv),
                                  C.c_int64(index.shape[0]), _p(go), _p(gg), _p(gu))
    return gg, gu


def grid_sample1d_fwd(grid, index, uv):
    grid = _c(grid, np.float32)
    N, Cc, W = grid.shape[:3]
    index, uv = _c(index, np.int64), _c(uv, np.float32).reshape(-1)
    out = np.zeros((index.shape[0], Cc), np.float32)
    lib().atdfo_grid_sample1d_fwd(_p(grid), C.c_int64(N), C.c_int(Cc), C.c_int(W), _p(index), _p(uv),
                                  C.c_int64(index.shape[0]), _p(out))
    return out


def grid_sample1d_bwd(grid, index, uv, grad_out):
    grid = _c(grid, np.float32)
    N, Cc, W = grid.shape[:3]
    index, uv, go = _c(index, np.int64), _c(uv, np.float32).reshape(-1), _c(grad_out, np.float32)
    gg = np.zeros(grid.shape, np.float64)
    gu = np.zeros(uv.shape, np.float64)
    lib().atdfo_grid_sample1d_bwd(_p(grid), C.c_int64(N), C.c_int(Cc), C.c_int(W), _p(index), _p(uv),
                                  C.c_int64(index.shape[0]), _p(go), _p(gg), _p(gu))
    return gg, gu


def gradient_correction(direction, ray_index, grad_out):
    direction, ray_index, go = _c(direction, np.float32), _c(ray_index, np.int64), _c(grad_out, np.float32)
    out = np.zeros_like(go)
    lib().atdfo_gradient_correction(_p(direction), _p(ray_index), _p(go), C.c_int64(go.shape[0]), _p(out))
    return out


def compute_bound_weight(node_id, in_roi_node):
    node_id, in_roi_node = _c(node_id, np.int64), _c(in_roi_node, np.int64)
    out = np.zeros(node_id.shape[0], np.int64)
    lib().atdfo_compute_bound_weight(_p(node_id), C.c_int64(node_id.shape[0]), _p(in_roi_node),
                                     C.c_int64(in_roi_node.shape[0]), _p(out))
    return out


# ------------------------------------------------------------------- mlp
def mlp_fwd(params, Cc, H, beta, x):
    params = _c(params, np.float32)
    x = _c(x, np.float32)
    S = x.shape[0]
    y = np.zeros(S, np.float32)
    dens = np.zeros(S, np.float32)
    lib().atdfo_mlp_fwd(_p(params), C.c_int(Cc), C.c_int(H), C.c_float(beta), _p(x), C.c_int64(S), _p(y), _p(dens))
    return y, dens


def integrate(v, w, ray, R):
    v = _c(v, np.float32)
    Cc = v.shape[1]
    P = np.zeros((Cc, R), np.float32)
    lib().atdfo_integrate(_p(v), _p(_c(w, np.float32)), _p(_c(ray, np.int64)), C.c_int64(v.shape[0]),
                          C.c_int(Cc), C.c_int64(R), _p(P))
    return P


def loss_per_node(per_ray_loss, ray_index, weight, node_id, num_nodes):
    prl = _c(per_ray_loss, np.float32)
    ls = np.zeros(num_nodes, np.float64)
    wsum = np.zeros(num_nodes, np.float64)
    lib().atdfo_accumulate_loss_per_node(_p(prl), C.c_int(prl.shape[0]), C.c_int64(prl.shape[1]),
                                         _p(_c(ray_index, np.int64)), _p(_c(weight, np.float32)),
                                         _p(_c(node_id, np.int64)), C.c_int64(len(ray_index)),
                                         C.c_int64(num_nodes), _p(ls), _p(wsum))
    return ls, wsum


# ----------------------------------------------------------- fused path
class FusedModel:
    """Holds the arrays the fused oracle reads (keeps them alive)."""

    def __init__(self, node_min, node_max, active, grid_cl, mlp, C_, H=64, beta=2.0, K=8, max_intersections=256,
                 jitter_mode=0, seed=0, decoder_bf16=False):
        self.node_min = _c(node_min, np.float32)
        self.node_max = _c(node_max, np.float32)
        self.active_ids = active_ids(active)
        self.node_to_grid = prefix_sum(active)
        self.grid = _c(grid_cl, np.float32)
        self.mlp = _c(mlp, np.float32)
        self.G = self.grid.shape[1]
        self.m = Model(_p(self.node_min), _p(self.node_max), _p(self.active_ids), self.active_ids.shape[0],
                       _p(self.node_to_grid), _p(self.grid), self.G, C_, H, _p(self.mlp), beta, K,
                       max_intersections, jitter_mode, seed, 1 if decoder_bf16 else 0)

    def set_intervals(self, iv_off=None, node=None, tmin=None, tmax=None, count=None):
        """Use explicit per-ray intervals instead of the traversal (None resets)."""
        if iv_off is None:
            self._iv = None
            self.m.iv_off = self.m.iv_node = self.m.iv_tmin = self.m.iv_tmax = self.m.iv_count = None
            return
        self._iv = (_c(iv_off, np.int64), _c(node, np.int32), _c(tmin, np.float32), _c(tmax, np.float32),
                    _c(count, np.int32))
        self.m.iv_off, self.m.iv_node, self.m.iv_tmin, self.m.iv_tmax, self.m.iv_count = [_p(a) for a in self._iv]

    def project(self, o, d, precision=0, threads=0, bins=0):
        """Projection [R] (float64 array); with bins > 0 also the per-ray mass
        histogram [R, bins] over the root chord."""
        o, d = _c(o, np.float32), _c(d, np.float32)
        P = np.zeros(o.shape[0], np.float64)
        B = np.zeros((o.shape[0], bins), np.float64) if bins else None
        st = lib().atdfo_project_fwd_bins(C.byref(self.m), _p(o), _p(d), C.c_int64(o.shape[0]), C.c_int(precision),
                                          _p(P), _p(B), C.c_int(bins), C.c_int(threads))
        if st != 0:
            raise ValueError("too many intersections")
        return (P, B) if bins else P

    def min_preact(self, o, d):
        """Per-ray smallest |hidden pre-activation| (FP64); rays below ~1e-5
        sit on a ReLU kink where FP32 and FP64 gradients may differ."""
        o, d = _c(o, np.float32), _c(d, np.float32)
        out = np.zeros(o.shape[0], np.float64)
        lib().atdfo_project_min_preact(C.byref(self.m), _p(o), _p(d), C.c_int64(o.shape[0]), _p(out))
        return out

    def project_bwd(self, o, d, dLdP):
        o, d = _c(o, np.float32), _c(d, np.float32)
        dLdP = _c(dLdP, np.float64)
        gg = np.zeros(self.grid.shape, np.float64)
        gm = np.zeros(self.mlp.shape, np.float64)
        st = lib().atdfo_project_bwd(C.byref(self.m), _p(o), _p(d), C.c_int64(o.shape[0]), _p(dLdP), _p(gg), _p(gm))
        if st != 0:
            raise ValueError("too many intersections")
        return gg, gm


def point_uniform(seed, node, point, axis):
    f = lib().atdfo_point_uniform
    f.restype = C.c_float
    f.argtypes = [C.c_uint64, C.c_int64, C.c_int64, C.c_int]
    return f(seed, node, point, axis)


def node_density_stats(fm, mode, n, rounds=1, seed=0, precision=0, want_densities=False):
    na = fm.active_ids.shape[0]
    per = n ** 3 * (rounds if mode == 1 else 1)
    mx = np.zeros(na, np.float32)
    s1 = np.zeros(na, np.float64)
    s2 = np.zeros(na, np.float64)
    dens = np.zeros(na * per, np.float32) if want_densities else None
    lib().atdfo_node_density_stats(C.byref(fm.m), C.c_int(mode), C.c_int(n), C.c_int(rounds), C.c_uint64(seed),
                                   C.c_int(precision), _p(mx), _p(s1), _p(s2), _p(dens))
    return (mx, s1, s2, dens.reshape(na, per) if want_densities else None)


class TrainBaseline:
    """FP32 CPU training step (the timed CPU baseline)."""

    def __init__(self, fused: FusedModel, lr=1e-3):
        self.f = fused
        self.grid = fused.grid.copy()
        self.mlp = fused.mlp.copy()
        self.gg = np.zeros_like(self.grid)
        self.gm = np.zeros_like(self.mlp)
        self.mg = np.zeros_like(self.grid)
        self.vg = np.zeros_like(self.grid)
        self.mm = np.zeros_like(self.mlp)
        self.vm = np.zeros_like(self.mlp)
        self.step = 0
        self.lr = lr

    def train_step(self, o, d, target, threads=0, do_adam=True):
        o, d = _c(o, np.float32), _c(d, np.float32)
        t = _c(target, np.float32)
        self.step += 1
        return lib().atdfo_train_step_f32(C.byref(self.f.m), _p(self.grid), _p(self.mlp),
                                          C.c_int64(self.grid.size), _p(o), _p(d), _p(t), C.c_int64(o.shape[0]),
                                          _p(self.gg), _p(self.gm), _p(self.mg), _p(self.vg), _p(self.mm),
                                          _p(self.vm), C.c_int(self.step), C.c_float(self.lr), C.c_int(threads),
                                          C.c_int(1 if do_adam else 0))


# ------------------------------------------------------ learned proposal
PROP_BINS, PROP_IN, PROP_H = 32, 38, 64


class PropStruct(C.Structure):
    _fields_ = [("root_min", _P), ("root_max", _P), ("table", _P), ("depth", C.c_int), ("coarse", _P),
                ("params", _P), ("mix", C.c_float), ("kl_threshold", C.c_float), ("jitter_mode", C.c_int),
                ("jitter_seed", C.c_uint64)]


def prop_num_params():
    return int(lib().atdfo_prop_num_params())


def prop_exp(x):
    return lib().atdfo_prop_exp(float(x))


def prop_log(x):
    return lib().atdfo_prop_log(float(x))


def leaf_table(node_min, node_max, children, active, depth):
    res = 1 << depth
    table = np.zeros(res ** 3, np.int32)
    ch = _c(children, np.int32)
    lib().atdfo_leaf_table_ns(_p(_c(node_min, np.float32)), _p(_c(node_max, np.float32)), _p(ch), C.c_int(ch.shape[1]),
                              _p(_c(active, np.int32)), C.c_int(depth), _p(table))
    return table


def locate(node_min, node_max, table, depth, p):
    return int(lib().atdfo_locate(_p(_c(node_min[0], np.float32)), _p(_c(node_max[0], np.float32)),
                                  _p(_c(table, np.int32)), C.c_int(depth), _p(_c(p, np.float32))))


class Proposal:
    """Holds the arrays the proposal oracle reads."""

    def __init__(self, node_min, node_max, table, depth, coarse, params, mix=0.25, kl_threshold=0.5, jitter_mode=0,
                 jitter_seed=0):
        self.root_min = _c(np.asarray(node_min)[0], np.float32)
        self.root_max = _c(np.asarray(node_max)[0], np.float32)
        self.table = _c(table, np.int32)
        self.coarse = _c(coarse, np.float32)
        self.params = _c(params, np.float32)
        self.s = PropStruct(_p(self.root_min), _p(self.root_max), _p(self.table), depth, _p(self.coarse),
                            _p(self.params), mix, kl_threshold, jitter_mode, jitter_seed)

    def features(self, o, d):
        o, d = _c(o, np.float32), _c(d, np.float32)
        R = o.shape[0]
        f = np.zeros((R, PROP_IN), np.float32)
        ch = np.zeros((R, 2), np.float32)
        lib().atdfo_prop_features(C.byref(self.s), _p(o), _p(d), C.c_int64(R), _p(f), _p(ch))
        return f, ch

    def forward(self, feat):
        feat = _c(feat, np.float32)
        R = feat.shape[0]
        z = np.zeros((R, PROP_BINS), np.float32)
        p = np.zeros((R, PROP_BINS), np.float32)
        lib().atdfo_prop_forward(_p(self.params), _p(feat), C.c_int64(R), _p(z), _p(p))
        return z, p

    def sample(self, o, d):
        o, d = _c(o, np.float32), _c(d, np.float32)
        R = o.shape[0]
        out = dict(fallback=np.zeros(R, np.int32), n_learned=np.zeros(R, np.int32),
                   node=np.full((PROP_BINS, R), -7, np.int32), tmin=np.zeros((PROP_BINS, R), np.float32),
                   tmax=np.zeros((PROP_BINS, R), np.float32), kl=np.zeros(R, np.float32))
        lib().atdfo_prop_sample(C.byref(self.s), _p(o), _p(d), C.c_int64(R), _p(out["fallback"]),
                                _p(out["n_learned"]), _p(out["node"]), _p(out["tmin"]), _p(out["tmax"]),
                                _p(out["kl"]))
        return out

    def allocate(self, o, d, extra, iv_off, tmin, tmax, count):
        """Learned allocation over a stratified interval list: (count, first,
        fallback, ray sample counts, kl), count and first rewritten."""
        o, d = _c(o, np.float32), _c(d, np.float32)
        R = o.shape[0]
        iv_off = _c(iv_off, np.int64)
        tmin, tmax = _c(tmin, np.float32), _c(tmax, np.float32)
        cnt = np.array(count, np.int32, copy=True)
        first = np.zeros_like(cnt)
        fb = np.zeros(R, np.int32)
        ns = np.zeros(R, np.int32)
        kl = np.zeros(R, np.float32)
        lib().atdfo_prop_allocate(C.byref(self.s), _p(o), _p(d), C.c_int64(R), C.c_int32(extra), _p(iv_off),
                                  _p(tmin), _p(tmax), _p(cnt), _p(first), _p(fb), _p(ns), _p(kl))
        return cnt, first, fb, ns, kl


def prop_grad(params, feat, bins, power=1.0):
    params = _c(params, np.float32)
    feat = _c(feat, np.float32)
    bins = _c(bins, np.float32)
    g = np.zeros(params.shape[0], np.float64)
    loss = lib().atdfo_prop_grad(_p(params), _p(feat), _p(bins), C.c_int64(feat.shape[0]), C.c_float(power), _p(g))
    return loss, g


# ------------------------------------------------------- point location / grid transfer
def node_for_position(node_min, node_max, children, active, p):
    ch = _c(children, np.int32)
    fn = lib().atdfo_active_node_for_position_ns
    fn.restype = C.c_int32
    return int(fn(_p(_c(node_min, np.float32)), _p(_c(node_max, np.float32)), _p(ch), C.c_int(ch.shape[1]),
                  _p(_c(active, np.int32)), _p(_c(p, np.float32))))


def node_id_for_position(node_min, node_max, children, active, pos):
    """(node_id i64, mask f32) of NodeIdForPositionGPU."""
    active = _c(active, np.int32)
    pos = _c(pos, np.float32).reshape(-1, 3)
    n = pos.shape[0]
    node_id = np.zeros(n, np.int64)
    mask = np.zeros(n, np.float32)
    first = int(np.nonzero(active)[0][0])
    ch = _c(children, np.int32)
    lib().atdfo_node_id_for_position_ns(_p(_c(node_min, np.float32)), _p(_c(node_max, np.float32)), _p(ch),
                                        C.c_int(ch.shape[1]), _p(active), C.c_int64(first), _p(pos), C.c_int64(n),
                                        _p(node_id), _p(mask))
    return node_id, mask


def grid_transfer(node_min, node_max, children, src_active, src_slot, src_grid, dst_nodes, dst_slot, dst_grid,
                  layout):
    """Writes into dst_grid (a contiguous f32 array) and returns it.  Slots
    None = identity.  layout 0: [slots, G, G, G, C]; 1: [slots, C, G, G, G]."""
    src_grid = _c(src_grid, np.float32)
    assert dst_grid.dtype == np.float32 and dst_grid.flags.c_contiguous
    if layout == 0:
        G, Cc = src_grid.shape[1], src_grid.shape[4]
    else:
        Cc, G = src_grid.shape[1], src_grid.shape[2]
    dst_nodes = _c(dst_nodes, np.int64)
    ss = None if src_slot is None else _c(src_slot, np.int32)
    ds = None if dst_slot is None else _c(dst_slot, np.int32)
    ch = _c(children, np.int32)
    lib().atdfo_grid_transfer_ns(_p(_c(node_min, np.float32)), _p(_c(node_max, np.float32)), _p(ch),
                                 C.c_int(ch.shape[1]),
                                 _p(_c(src_active, np.int32)), None if ss is None else _p(ss), _p(src_grid),
                              _p(dst_nodes), C.c_int64(dst_nodes.shape[0]), None if ds is None else _p(ds),
                              _p(dst_grid), C.c_int(Cc), C.c_int(G), C.c_int(layout))
    return dst_grid


def volume_points(mode, size, roi_min=(0, 0, 0), step=(0, 0, 0), swap_xy=False):
    size = np.asarray(size, np.int32)
    n = int(np.prod(size.astype(np.int64)))
    pos = np.zeros((n, 3), np.float32)
    lib().atdfo_volume_points(C.c_int(mode), C.c_int(int(swap_xy)), _p(_c(size, np.int32)),
                              _p(_c(roi_min, np.float32)), _p(_c(step, np.float32)), _p(pos))
    return pos


# --------------------------------------------------------------------------
# scene data path
# --------------------------------------------------------------------------
def rotate_point(q, v):
    out = np.zeros(3, np.float64)
    lib().atdfo_rotate_point(_p(_c(q, np.float64)), _p(_c(v, np.float64)), _p(out))
    return out


def generate_rays(camera_model, w, h, intrinsics, rotation, translation, uv, image_id, camera_id,
                  vol_translate=(0.0, 0.0, 0.0), align_corners=True):
    uv = _c(uv, np.float32).reshape(-1, 2)
    n = uv.shape[0]
    o = np.zeros((n, 3), np.float32)
    d = np.zeros((n, 3), np.float32)
    tr = None if translation is None else _c(translation, np.float64)
    lib().atdfo_generate_rays(C.c_int(camera_model), C.c_int(w), C.c_int(h), C.c_int(int(align_corners)),
                              _p(_c(intrinsics, np.float64)), _p(_c(rotation, np.float64)), _p(tr),
                              _p(_c(vol_translate, np.float32)), _p(uv), _p(_c(image_id, np.int32)),
                              _p(_c(camera_id, np.int32)), C.c_int64(n), _p(o), _p(d))
    return o, d


def log_convert(raw, xray_min, xray_max, projection_factor=1.0, log10=True, log_space_input=False, mask=None):
    raw = _c(raw, np.float32)
    out = np.zeros(raw.shape, np.float32)
    m = None if mask is None else _c(mask, np.float32)
    lib().atdfo_log_convert(_p(raw), _p(m), C.c_int64(raw.size), C.c_double(xray_min), C.c_double(xray_max),
                            C.c_double(projection_factor), C.c_int(int(log10)), C.c_int(int(log_space_input)),
                            _p(out))
    return out


def sample_projection(projection, uv, image_id, mask=None, align_corners=True):
    projection = _c(projection, np.float32)
    h, w = projection.shape[-2:]
    uv = _c(uv, np.float32).reshape(-1, 2)
    n = uv.shape[0]
    t = np.zeros(n, np.float32)
    tm = np.zeros(n, np.float32)
    m = None if mask is None else _c(mask, np.float32)
    lib().atdfo_sample_projection(_p(projection), _p(m), C.c_int(w), C.c_int(h), C.c_int(int(align_corners)),
                                  _p(uv), _p(_c(image_id, np.int32)), C.c_int64(n), _p(t), _p(tm))
    return t, tm


def random_pixels(indices, image_camera, w, h, n, seed, interpolate=True, align_corners=True):
    indices = _c(indices, np.int32)
    uv = np.zeros((n, 2), np.float32)
    img = np.zeros(n, np.int32)
    cam = np.zeros(n, np.int32)
    ic = None if image_camera is None else _c(image_camera, np.int32)
    lib().atdfo_random_pixels(_p(indices), C.c_int64(indices.size), _p(ic), C.c_int(w), C.c_int(h),
                              C.c_int(int(align_corners)), C.c_int(int(interpolate)), C.c_uint64(seed),
                              C.c_int64(n), _p(uv), _p(img), _p(cam))
    return uv, img, cam


def point_in_any_image(camera_model, w, h, intrinsics, rotation, translation, roi_min, roi_max, points):
    points = _c(points, np.float32).reshape(-1, 3)
    rotation = _c(rotation, np.float64)
    mask = np.zeros(points.shape[0], np.float32)
    lib().atdfo_point_in_any_image(C.c_int(camera_model), C.c_int(w), C.c_int(h), _p(_c(intrinsics, np.float64)),
                                   _p(rotation), _p(_c(translation, np.float64)), C.c_int64(rotation.shape[0]),
                                   _p(_c(roi_min, np.float32)), _p(_c(roi_max, np.float32)), _p(points),
                                   C.c_int64(points.shape[0]), _p(mask))
    return mask


def node_lattice_points(node_min, node_max, active_ids, n):
    ids = _c(active_ids, np.int64)
    pos = np.zeros((ids.size, n * n * n, 3), np.float32)
    lib().atdfo_node_lattice_points(_p(_c(node_min, np.float32)), _p(_c(node_max, np.float32)), _p(ids),
                                    C.c_int64(ids.size), C.c_int(n), _p(pos))
    return pos


def tv_loss(x, channels_last=False, weight=None, scale=1.0, grad=None):
    """TVLoss of x [B, C, n...] (channels_last: [B, n..., C]); returns
    (loss, grad) with the gradient added to `grad` (zeros by default)."""
    x = _c(x, np.float32)
    if channels_last:
        nb, nc = x.shape[0], x.shape[-1]
        n = list(x.shape[1:-1])
        st = [s // 4 for s in x.strides]
        sb, sc, ss = st[0], st[-1], st[1:-1]
    else:
        nb, nc = x.shape[0], x.shape[1]
        n = list(x.shape[2:])
        st = [s // 4 for s in x.strides]
        sb, sc, ss = st[0], st[1], st[2:]
    D = len(n)
    g = np.zeros_like(x) if grad is None else _c(grad, np.float32).copy()
    w = None if weight is None else _c(weight, np.float32)
    fn = lib().atdfo_tv_loss
    fn.restype = C.c_double
    loss = fn(_p(x), C.c_int64(nb), C.c_int64(nc), C.c_int(D), _p(_c(n + [1] * (3 - D), np.int32)), C.c_int64(sb),
              C.c_int64(sc), _p(_c(list(ss) + [0] * (3 - D), np.int64)), _p(w), C.c_double(scale), _p(g))
    return loss, g


def neighbor_pairs(node_min, node_max, children, active, size, eps=1e-3, roi=None, quad=False, plane=None):
    """NodeNeighborSamples (quad: NodeNeighborSamplesQuad; plane given:
    NodeNeighborSamples2D of the quad-tree representation on that plane):
    (pos [2P, 3], node [2P]) in candidate order."""
    ids = active_ids(np.asarray(active))
    ch = _c(children, np.int32)
    variant = 2 if plane is not None else (1 if quad else 0)
    args = [_p(_c(node_min, np.float32)), _p(_c(node_max, np.float32)), _p(ch), C.c_int(ch.shape[1]),
            _p(_c(active, np.int32)), _p(ids), C.c_int64(ids.size), _p(_c(size, np.int32)), C.c_float(eps),
            C.c_int(roi is not None), _p(_c(roi[0] if roi else (0, 0, 0), np.float32)),
            _p(_c(roi[1] if roi else (0, 0, 0), np.float32)), C.c_int(variant), C.c_int(plane or 0)]
    fn = lib().atdfo_neighbor_pairs_ns
    fn.restype = C.c_int64
    n = fn(*args, None, None, C.c_int64(0))
    pos = np.zeros((2 * n, 3), np.float32)
    node = np.zeros(2 * n, np.int64)
    fn(*args, _p(pos), _p(node), C.c_int64(n))
    return pos, node


def field_density(fm, pos, node, precision=1):
    """Density of the FusedModel field at points pos [n, 3] in node [n]."""
    pos = _c(pos, np.float32).reshape(-1, 3)
    node = _c(node, np.int64)
    out = np.zeros(pos.shape[0], np.float64)
    lib().atdfo_field_density(C.byref(fm.m), _p(pos), _p(node), C.c_int64(pos.shape[0]), C.c_int(precision), _p(out))
    return out


def decode_features(params, Cc, H, beta, x, dd=None, bf=False):
    """DecodeFeatures in FP64: (density [S], grad_x [S, C], grad_mlp [NP]);
    the gradients are None without dd (dL/ddensity [S])."""
    x = _c(x, np.float32).reshape(-1, Cc)
    S = x.shape[0]
    dens = np.zeros(S, np.float64)
    gx = gm = None
    if dd is not None:
        dd = _c(dd, np.float64).reshape(S)
        gx = np.zeros((S, Cc), np.float64)
        gm = np.zeros(H * Cc + H + H * H + H + H + 1, np.float64)
    lib().atdfo_decode_features(_p(_c(params, np.float32)), C.c_int(Cc), C.c_int(H), C.c_double(beta), _p(x),
                                C.c_int64(S), C.c_int(int(bf)), _p(dd), _p(dens), _p(gx), _p(gm))
    return dens, gx, gm


def fourier_ring_means(img):
    """Ring means of the Fourier regulariser on img [Y, X, Z] f32: [X // 2, Y]
    (row 0 +inf), every meshgrid pair interpolated pointwise."""
    img = _c(img, np.float32)
    Y, X, Z = img.shape
    out = np.zeros((X // 2, Y), np.float32)
    lib().atdfo_fourier_ring_means(_p(img), C.c_int64(Y), C.c_int(X), C.c_int(Z), _p(out))
    return out


def fourier_terms(densities, rate=0.2, rate_w=0.01, ring=None):
    """The Fourier regulariser of GeometryTensorQO::VolumeRegularizer
    (geometry_tensor_qo.cpp:758-850, the masked branch) on densities
    [X, Y, Z], unscaled: (fx_term + fy_term, parts).  FP64 FFT (numpy), the
    band and ring values as the reference forms them.  `ring` [X // 2, Y]
    replaces the ring means (the reference detaches them: a fixed ring gives
    the function whose derivative its autograd takes)."""
    d = np.transpose(np.asarray(densities, np.float64), (1, 0, 2))
    Y, X, Z = d.shape
    cx, cy = X // 2, Z // 2
    rh = max(int(np.float32(rate_w) * np.float32(cy)), 1)
    rw = max(int(np.float32(rate_w) * np.float32(cx)), 1)
    F = np.abs(np.fft.fftshift(np.fft.fft2(d, axes=(-2, -1), norm="forward"), axes=(-2, -1)))
    fx = F[:, cy - rh + 1:cy + rh, :].mean(1)
    fy = F[:, :, cx - rw + 1:cx + rw].mean(2)
    total = X // 2
    if ring is None:
        ring = fourier_ring_means(F.astype(np.float32))
    ring = np.asarray(ring, np.float64)
    circle = np.concatenate([ring[::-1], np.zeros((1, Y)), ring], 0).T
    xs = np.arange(total, dtype=np.float32)
    yv = np.float32(1) - np.float32(1) / (np.exp(xs - np.float32(np.float32(rate) * np.float32(total))) + np.float32(1))
    factor = np.broadcast_to(np.concatenate([yv[::-1], [0.0], yv]).astype(np.float64)[None, :], circle.shape)
    mx, my = fx > circle, fy > circle
    tx = ((fx[mx] - circle[mx]) * factor[mx]).mean() if mx.any() else np.nan
    ty = ((fy[my] - circle[my]) * factor[my]).mean() if my.any() else np.nan
    return tx + ty, dict(fx=fx, fy=fy, circle=circle, factor=factor, mask_x=mx, mask_y=my, F=F, ring=ring)
