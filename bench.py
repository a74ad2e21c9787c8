#!/usr/bin/env python3
"""Benchmark: fwd+bwd rays/s of one training step of the reconstruction hot
path on one MI355X.

Workload (the 512^3 configuration): a 512^3 FP32 volume of 200 random
ellipsoids (seed 1) plus Gaussian noise (sigma 0.01), 61 tilts from -60 to
+60 degrees on a 512 x 512 detector, targets = line integrals of the
trilinearly interpolated noisy volume (step 1/4 voxel, FP64 sums), batches
of 2^18 rays, 11^3 x 16-channel node grids, the 16-64-64-1 decoder.

The octree is grown by the method's own structure loop in an untimed
warm-up: it starts uniform at depth 3 (512 nodes), trains the plain FP32
field, and every 100 iterations computes the per-node FP64 error
(AccumulateSampleLossPerNode of the per-ray MSE), the per-node max density
(ComputeMaxDensity), and restructures (exact split/merge solve, grid
transfer, optimizer reset) up to depth 5 and 8192 active nodes.  Every leg
then starts from that grown state.

One step = sampling -> fused forward -> MSE -> fused adjoint -> Adam on every
grid and decoder parameter, all on the device; inputs (rays, targets) are
resident in HBM before the timed region.

Legs (--mode learned, the default, runs all four in one process):
  plain    FP32 grids and decoder, stratified samples (~128 per ray)
  bf16     BF16 grids (FP32 masters), BF16-MFMA decoder, stratified ~128
  strat32  BF16 as above, stratified at the learned leg's sample count (~39):
           the ablation of the learned proposal
  learned  BF16 with the learned proposal.  Default (--proposal-mode alloc):
           every node interval of the K=1 stratified list keeps one sample
           and 16 more per ray follow the proposal's CDF (~38 per ray); the
           KL fallback rays keep the K=1 list.  --proposal-mode cells: 32
           one-sample cells per ray, stratified K=12 on the fallback rays.
           The proposal is trained (on K=12 binned masses) before the leg's
           first step and retrained every 100 iterations (50 Adam steps on
           2^15 rays); that cost is amortised into ms_per_step
Each leg trains --train-iters iterations over --batches distinct batches
(default 32 x 2^18 rays, half of the tilt series' pixels: with fewer the
174 M-parameter field mostly memorises its training rays and the held-out
error measures that, not the sampler); every 100 iterations the held-out
projection MSE is judged with the stratified ~128-sample sampler
(midpoints) against the held-out targets, each field in its own
arithmetic.  `value` is the learned leg's rays/s; the other legs sit beside
it, and "equal_mse" names the legs whose held-out MSE is within 1 % of the
plain leg's (or below it) with their speed-up over plain.

Prints one JSON line (see the keys below).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import statistics
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "fwd+bwd rays/s at 512³, 61 tilts; learned+bf16 vs plain at equal proj. MSE"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0
C, G, H, DEPTH = 16, 11, 64, 5


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--mode", choices=["plain", "bf16", "learned", "vm"], default="learned",
                    help="learned: all four legs; plain / bf16: that leg alone; vm: the tensorial field through "
                         "the composed operators")
    ap.add_argument("--train-iters", type=int, default=1500,
                    help="iterations each leg trains (MSE judged every 100)")
    ap.add_argument("--grow-iters", type=int, default=400, help="shared warm-up iterations growing the octree")
    ap.add_argument("--restructure-every", type=int, default=100)
    ap.add_argument("--mix", type=float, default=0.3, help="uniform mixing of the proposal pdf")
    ap.add_argument("--kl-threshold", type=float, default=0.3, help="proposal fallback threshold")
    ap.add_argument("--proposal-rays", type=int, default=1 << 15, help="rays per proposal training pass")
    ap.add_argument("--proposal-mode", choices=["alloc", "cells"], default="alloc",
                    help="alloc: the samples of the K=--alloc-k stratified list plus --alloc-extra "
                         "samples along the proposal's CDF; cells: 32 one-sample cells per ray")
    ap.add_argument("--alloc-extra", type=int, default=9, help="learned samples per ray beyond the stratified list's")
    ap.add_argument("--proposal-target", choices=["mass", "variation"], default="mass",
                    help="what the proposal learns: the binned density mass, or its change across each bin")
    ap.add_argument("--target-power", type=float, default=0.5, help="proposal target = (binned quantity)^power")
    ap.add_argument("--alloc-k", type=int, default=2, help="stratified K of the allocation's base list (and fallback)")
    ap.add_argument("--seed", type=int, default=0,
                    help="offset of the per-iteration jitter seeds (repeat runs of the same workload)")
    ap.add_argument("--volume", type=int, default=512, help="phantom and detector size (512 or 768)")
    ap.add_argument("--rays", type=int, default=1 << 18)
    ap.add_argument("--batches", type=int, default=32,
                    help="distinct training batches (32 x 2^18 rays = half of the 61 x 512^2 detector pixels)")
    ap.add_argument("--budget", type=int, default=8192, help="max active octree nodes")
    ap.add_argument("--ray-order", choices=["row", "tile"], default="tile",
                    help="order of a batch's rays: by (tilt, row, column), or by tilt and the Z-order of the pixel")
    ap.add_argument("--samples-per-ray", type=float, default=128.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-runs", type=int, default=5, help="timed CPU baseline runs (median reported)")
    ap.add_argument("--ray-major-forward", action="store_true",
                    help="training step with the ray-major forward (default: the node-major forward over the "
                         "adjoint's sort)")
    ap.add_argument("--save-state", default="", help="write the grown workload state to this file and exit")
    ap.add_argument("--load-state", default="", help="start from a grown workload state written by --save-state")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="only run this many steps of the plain and learned legs (profiling)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy a 512 MiB buffer 3 times (scale factor for HBM byte counters)")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# --------------------------------------------------------------------------
# synthetic data: voxelised noisy phantom and its line integrals
# --------------------------------------------------------------------------
def voxel_phantom(ell, n, noise_sigma, noise_seed, dev):
    """[n, n, n] float32 volume (index z, y, x) on the lattice -1 + 2 i / (n - 1)
    (align_corners): the sum of the ellipsoid densities at each voxel plus
    N(0, noise_sigma^2) noise."""
    import torch

    lin = torch.linspace(-1.0, 1.0, n, dtype=torch.float64, device=dev)
    vol = torch.zeros((n, n, n), dtype=torch.float32, device=dev)
    step = 2.0 / (n - 1)
    for e in ell:
        c, a, rho = e[0:3], e[3:6], float(e[6])
        lo = [max(0, int(math.floor((c[k] - a[k] + 1.0) / step))) for k in range(3)]
        hi = [min(n, int(math.ceil((c[k] + a[k] + 1.0) / step)) + 1) for k in range(3)]
        x = ((lin[lo[0]:hi[0]] - c[0]) / a[0]) ** 2
        y = ((lin[lo[1]:hi[1]] - c[1]) / a[1]) ** 2
        z = ((lin[lo[2]:hi[2]] - c[2]) / a[2]) ** 2
        inside = (z[:, None, None] + y[None, :, None] + x[None, None, :]) <= 1.0
        vol[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]] += inside.to(torch.float32) * rho
    if noise_sigma > 0:
        g = torch.Generator(device=dev).manual_seed(noise_seed)
        vol += torch.randn((n, n, n), generator=g, device=dev, dtype=torch.float32) * noise_sigma
    return vol


def volume_projection(vol, o, d, chunk=4096):
    """Line integrals of the trilinearly interpolated volume along the rays
    (|d| = 1), midpoint rule with steps of at most 1/4 voxel over each
    ray's chord through [-1, 1]^3, FP64 sums.  o, d: [R, 3] tensors."""
    import torch
    import torch.nn.functional as F

    n = vol.shape[0]
    h = 2.0 / (n - 1) / 4.0
    o64, d64 = o.double(), d.double()
    with torch.no_grad():
        inv = torch.where(d64 == 0, torch.full_like(d64, float("inf")), 1.0 / d64)
        t1, t2 = (-1.0 - o64) * inv, (1.0 - o64) * inv
        par = d64 == 0
        inside = (o64 >= -1.0) & (o64 < 1.0)
        tlo = torch.where(par, torch.where(inside, -math.inf, math.inf), torch.minimum(t1, t2)).amax(1)
        thi = torch.where(par, torch.where(inside, math.inf, -math.inf), torch.maximum(t1, t2)).amin(1)
        L = torch.clamp(thi - tlo, min=0.0)
        nsteps = torch.ceil(L / h).long()
        out = torch.zeros(o.shape[0], dtype=torch.float64, device=o.device)
        src = vol[None, None]
        for s in range(0, o.shape[0], chunk):
            e = min(s + chunk, o.shape[0])
            ns = nsteps[s:e]
            m = int(ns.max().item()) if e > s else 0
            if m == 0:
                continue
            stp = L[s:e] / ns.clamp(min=1)
            k = torch.arange(m, device=o.device, dtype=torch.float64)
            t = tlo[s:e, None] + (k[None, :] + 0.5) * stp[:, None]
            pts = (o64[s:e, None, :] + d64[s:e, None, :] * t[:, :, None]).float()
            v = F.grid_sample(src, pts.view(1, e - s, m, 1, 3), mode="bilinear", padding_mode="zeros",
                              align_corners=True)[0, 0, :, :, 0]
            valid = k[None, :] < ns[:, None]
            out[s:e] = (v.double() * valid).sum(1) * stp
        return out


# --------------------------------------------------------------------------
# workload: data, grown tree, initial field
# --------------------------------------------------------------------------
def build_data(N, args, dev):
    import torch

    from adaptivetensordf_amd import phantom

    n = args.volume
    t0 = time.time()
    ell = phantom.ellipsoids(200, 1)
    vol = voxel_phantom(ell, n, 0.01, 11, dev)
    angles = phantom.tilt_angles(61)

    def make_batch(nrays, seed):
        tl, px, py = phantom.sample_pixels(nrays, 61, n, n, seed, args.ray_order)
        o, d = phantom.rays_for_pixels(tl, px, py, n, n, angles)
        o_t, d_t = torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev)
        return N.RayList(o_t, d_t), volume_projection(vol, o_t, d_t).float().contiguous()

    batches = [make_batch(args.rays, 100 + b) for b in range(args.batches)]
    ev = make_batch(1 << 16, 999)
    del vol
    torch.cuda.synchronize()
    log(f"data {time.time() - t0:.1f}s: {n}^3 noisy phantom, {args.batches} x {args.rays} training rays, "
        f"{ev[0].size()} held-out rays")
    return batches, ev


def field_params(N, bf16, node_forward=True):
    fp = N.FieldParams()
    fp.C, fp.G, fp.H = C, G, H
    fp.bf16_decoder = bf16
    fp.node_forward = node_forward
    return fp


def grow_tree(N, args, batches, dev):
    """The method's structure loop in the untimed warm-up: plain FP32 training
    from a uniform depth-3 tree, restructured every --restructure-every
    iterations by the per-node FP64 error.  Returns (tree, grid, mlp, log)."""
    import torch

    from adaptivetensordf_amd import phantom

    tree = N.HyperTreeBase(3, DEPTH)
    tree.settree()
    tree.SetActive(3)
    tree.to("cuda")
    st = N.MakeTrainState(phantom.init_grid(int(tree.NumActiveNodes()), G, C, std=0.01, seed=0, device=dev),
                          phantom.init_mlp(C, H, seed=0).to(dev), False, 1e-3, 1e-3)
    fp = field_params(N, False)
    sp = N.SamplingParams()
    sp.max_samples_per_node = 12
    sp.jitter_mode = 2
    tp = N.TreeOptimizerParams()
    tp.max_active_nodes = args.budget
    rays_e, tgt_e = batches[0]
    sub = torch.arange(0, rays_e.size(), 4, device=dev)
    err_rays = N.RayList(rays_e.origin.index_select(0, sub).contiguous(),
                         rays_e.direction.index_select(0, sub).contiguous())
    err_tgt = tgt_e.index_select(0, sub).contiguous()
    history = []
    t0 = time.time()
    for it in range(args.grow_iters):
        if it > 0 and it % args.restructure_every == 0 and it <= args.grow_iters - args.restructure_every:
            samples = tree.CreateSamplesForRays(err_rays, 12, False)
            with torch.no_grad():
                P = N.ProjectRays(tree, err_rays, st.grid, st.mlp, sp, fp)
            per_ray = ((P - err_tgt) ** 2)[None].float()
            loss_sum, _ = N.AccumulateSampleLossPerNode(samples, per_ray, tree.NumNodes())
            tree.SetErrorForActiveNodes(loss_sum.to(tree.node_error.device), "override")
            N.ComputeMaxDensity(tree, st.grid, st.mlp, fp, 20)
            before = int(tree.NumActiveNodes())
            changed, dec = N.RestructureTrainState(tree, st, tp)
            depth = int(tree.node_depth[tree.ActiveNodeTensor().to(tree.node_depth.device)].max().item())
            history.append({"iteration": it, "active_before": before, "active_after": int(tree.NumActiveNodes()),
                            "split": int(dec.num_split), "merged_groups": int(dec.num_grp), "max_depth": depth,
                            "err_rays": err_rays.size()})
            log(f"  restructure at {it}: {before} -> {tree.NumActiveNodes()} active nodes (split {dec.num_split}, "
                f"merged groups {dec.num_grp}), depth {depth}")
        sp.jitter_seed = 1000 + it
        b = it % len(batches)
        N.TrainStep(tree, st, batches[b][0], batches[b][1], sp, fp)
    torch.cuda.synchronize()
    log(f"grow {time.time() - t0:.1f}s: {args.grow_iters} iterations, {tree.NumActiveNodes()} active nodes, "
        f"loss {float(st.loss.item()):.4e}")
    return tree, st.grid.clone(), st.mlp.clone(), history


def save_state(path, tree, grid, mlp, history):
    import torch

    torch.save({"node_active": tree.node_active.cpu(), "grid": grid.cpu(), "mlp": mlp.cpu(), "history": history},
               path)


def load_state(N, path, dev):
    import torch

    s = torch.load(path, weights_only=True)
    tree = N.HyperTreeBase(3, DEPTH)
    tree.settree()
    tree.node_active = s["node_active"].to(torch.int32)
    tree.UpdateActive()
    tree.to("cuda")
    return tree, s["grid"].to(dev), s["mlp"].to(dev), s["history"]


def choose_K(N, tree, rays, target):
    sp = N.SamplingParams()
    best = None
    for K in range(1, 65):
        sp.max_samples_per_node = K
        s_eff = tree.CreateIntervalsForRays(rays, sp).num_samples / rays.size()
        if best is None or abs(s_eff - target) < abs(best[1] - target):
            best = (K, s_eff)
        if s_eff > target:
            break
    return best


# --------------------------------------------------------------------------
# legs
# --------------------------------------------------------------------------
class Judge:
    """Held-out projection MSE with the stratified ~128-sample sampler."""

    def __init__(self, N, tree, ev, K):
        self.N, self.tree = N, tree
        self.rays, self.tgt = ev
        self.sp = N.SamplingParams()
        self.sp.max_samples_per_node = K
        self.iv = tree.CreateIntervalsForRays(self.rays, self.sp)

    def __call__(self, st, fp):
        P = self.N.ProjectIntervals(self.tree, self.rays, self.iv, st.grid, st.mlp, self.sp, fp)
        return float(((P.double() - self.tgt.double()) ** 2).mean().item())


def run_leg(N, name, tree, batches, K, fp, grid0, mlp0, bf16, ps, judge, args, steps, prop_rays=None, K_train=None):
    """Train one leg for --train-iters iterations (or `steps` when
    profiling), timing iterations [warmup, warmup + steps).  With a proposal
    state, the proposal is trained before the first step (200 Adam steps)
    and every 100 iterations (50 steps) on masses of the K_train stratified
    binned forward; K is the leg's sampling (for the proposal: its fallback /
    base list)."""
    import torch

    st = N.MakeTrainState(grid0.clone(), mlp0.clone(), bf16, 1e-3, 1e-3)
    sp = N.SamplingParams()
    sp.max_samples_per_node = K
    sp.jitter_mode = 2
    sp_train = N.SamplingParams()
    sp_train.max_samples_per_node = K_train or K
    sp_train.jitter_mode = 2
    updates = []  # (iteration, ms, proposal loss)
    curve = [] if judge is not None else None
    total = args.warmup + steps if args.profile_steps > 0 else max(args.train_iters, args.warmup + steps)

    def update(i):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b = (i // 100) % len(batches)
        sp_train.jitter_seed = 5000 + i + 100000 * args.seed
        loss = N.ProposalTrain(tree, ps, prop_rays[b], sp_train, st.grid, st.mlp, fp, 200 if i == 0 else 50)
        torch.cuda.synchronize()
        updates.append((i, (time.perf_counter() - t0) * 1e3, loss))

    t_start = elapsed = None
    in_window_update_ms = 0.0
    for i in range(total):
        # judged every 100 iterations, and every 25 over the last 500 (the
        # points the equal-MSE comparison takes its median over); a judgement
        # inside the timed steps is taken out of their time
        if curve is not None and (i % 100 == 0 or (i >= total - 500 and i % 25 == 0)):
            timed = args.warmup <= i < args.warmup + steps
            if timed:
                torch.cuda.synchronize()
                tj = time.perf_counter()
            curve.append((i, judge(st, fp)))
            if timed:
                torch.cuda.synchronize()
                in_window_update_ms += (time.perf_counter() - tj) * 1e3
        if i == args.warmup:
            torch.cuda.synchronize()
            first_loss = float(st.loss.item()) if i > 0 else float("nan")
            t_start = time.perf_counter()
        if ps is not None and i % 100 == 0:
            n_before = len(updates)
            update(i)
            if args.warmup <= i < args.warmup + steps:
                in_window_update_ms += updates[n_before][1]
        b = i % len(batches)
        rays, tgt = batches[b]
        sp.jitter_seed = 7 + i + 100000 * args.seed
        if ps is None:
            N.TrainStep(tree, st, rays, tgt, sp, fp)
        else:
            N.TrainStepLearned(tree, st, ps, rays, tgt, sp, fp)
        if i == args.warmup + steps - 1:
            torch.cuda.synchronize()
            elapsed = time.perf_counter() - t_start
            last_loss = float(st.loss.item())
    torch.cuda.synchronize()
    if curve is not None:
        curve.append((total, judge(st, fp)))
    R = batches[0][0].size()
    raw_ms = elapsed * 1e3 / steps
    leg = {"name": name, "state": st, "raw_ms_per_step": raw_ms, "loss_first_last": [first_loss, last_loss],
           "iterations": total, "mse_curve": curve}
    rays0 = batches[0][0]
    sp.jitter_seed = 7
    if ps is not None:
        iv, fb = N.CreateProposalIntervals(tree, ps, rays0, sp)
        retrain = [u[1] for u in updates if u[0] > 0]
        amort = (statistics.mean(retrain) if retrain else updates[0][1]) / 100.0
        ms = raw_ms - in_window_update_ms / steps + amort
        mode = "alloc" if ps.pp.alloc_extra > 0 else "cells"
        leg["proposal"] = {
            "mode": mode, "mix": args.mix, "kl_threshold": args.kl_threshold, "bins": 32,
            **({"base_K": K, "extra_samples": ps.pp.alloc_extra} if mode == "alloc" else {"samples": 32}),
            "train_rays": args.proposal_rays, "updates": len(updates),
            "updates_in_timed_region": sum(1 for u in updates if args.warmup <= u[0] < args.warmup + steps),
            "initial_update_ms": updates[0][1], "retrain_ms_mean": statistics.mean(retrain) if retrain else None,
            "amortised_ms_per_step": amort, "proposal_loss_first_last": [updates[0][2], updates[-1][2]],
            "fallback_frac": float(fb.float().mean().item())}
    else:
        iv = tree.CreateIntervalsForRays(rays0, sp)
        ms = raw_ms
    leg["K"] = K
    leg["ms_per_step"] = ms
    leg["rays_per_s"] = R / (ms * 1e-3)
    leg["samples_per_ray"] = iv.num_samples / rays0.size()
    log(f"{name}: {steps} timed steps {raw_ms:.2f} ms/step raw, {ms:.2f} ms/step"
        f"{' with the proposal retraining amortised' if ps is not None else ''}, {leg['rays_per_s'] / 1e6:.3f} M rays/s, "
        f"{leg['samples_per_ray']:.1f} samples/ray"
        + (f", final held-out MSE {curve[-1][1]:.4e} after {total} iterations" if curve else ""))
    return leg


def leg_sampling(N, K):
    """The sampling of a leg's kernel timing: stratified K (for the learned
    leg its fallback / base list), hash jitter, a fixed seed."""
    sp = N.SamplingParams()
    sp.max_samples_per_node = K
    sp.jitter_mode = 2
    sp.jitter_seed = 7
    return sp


def kernel_times(N, tree, rays, sp, fp, st, bf16, ps):
    """HIP-event durations of the fused forward and adjoint over one batch's
    interval list (the leg's own sampler), on the stream they run on, and the
    roofline of the dominant one."""
    import ctypes as Ct

    import torch

    from adaptivetensordf_amd import capi

    if ps is not None:
        iv, _ = N.CreateProposalIntervals(tree, ps, rays, sp)
    else:
        iv = tree.CreateIntervalsForRays(rays, sp)
    dev = rays.origin.device
    S_total = iv.num_samples
    R = rays.size()
    f = capi.Field(C, H, G, capi.GRID_BF16 if bf16 else capi.GRID_F32, 2.0, 0, st.grid.data_ptr(),
                   st.grid.shape[0], tree.node_to_grid.data_ptr(), st.mlp.data_ptr(),
                   capi.DECODER_BF16 if bf16 else capi.DECODER_F32, 0)
    csp = capi.Sampling(sp.max_samples_per_node, 256, sp.jitter_mode, 0, sp.jitter_seed, None, None)
    civ = capi.Intervals(iv.node.data_ptr(), iv.tmin.data_ptr(), iv.tmax.data_ptr(), iv.count.data_ptr(),
                         iv.first.data_ptr(), iv.ray.data_ptr(), iv.num_intervals)
    packed = tree.PackedTree()
    s = capi.stream_ptr()
    P = torch.empty(R, device=dev)
    dP = torch.randn(R, device=dev) * 1e-6
    gg = torch.zeros(st.grid.shape, dtype=torch.float32, device=dev)
    gm = torch.zeros_like(st.mlp)
    ws_bytes = capi.lib().atdf_project_bwd_workspace_bytes(R, iv.num_intervals, Ct.byref(f))
    if fp.two_pass_adjoint:
        ws_bytes += capi.lib().atdf_project_bwd_staging_bytes(iv.num_intervals, S_total, Ct.byref(f))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    args = (capi.ptr(packed), tree.NumNodes(), capi.ptr(rays.origin), capi.ptr(rays.direction), R, Ct.byref(csp),
            capi.ptr(iv.ray_offset), Ct.byref(civ), Ct.byref(f))
    node_fwd = fp.node_forward

    def prep():
        capi.call("atdf_project_prepare", *args, capi.ptr(ws), ws.numel(), s)

    def fwd():
        if node_fwd:
            capi.call("atdf_project_fwd_prepared", *args, capi.ptr(P), capi.ptr(ws), ws.numel(), s)
        else:
            capi.call("atdf_project_fwd", *args, capi.ptr(P), s)

    def bwd():
        # the adjoint over the counting sort made by prep(): the node kernel
        # and its decoder-partial reduction (atdf_project_bwd = prep + this)
        capi.call("atdf_project_bwd_prepared", *args, capi.ptr(dP), capi.ptr(gg), capi.ptr(gm), None,
                  capi.ptr(ws), ws.numel(), s)

    def timed(fn, reps=5):
        fn()
        torch.cuda.synchronize()
        cur = torch.cuda.current_stream()  # the stream every call above launches on
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(cur)
        for _ in range(reps):
            fn()
        e1.record(cur)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps / 1e3

    t_prep = timed(prep)
    prep()
    t_fwd = timed(fwd)
    t_bwd = timed(bwd)
    b = 2 if bf16 else 4
    # algorithmic bytes per launch (no cache reuse): per sample the gather of
    # 8 corners x C channels (fwd), plus the scatter of 8 x C f32 (bwd);
    # per ray origin + direction (24 B) and the projection or dL/dP (4 B)
    fwd_bytes = S_total * 8 * C * b + R * 28
    bwd_bytes = S_total * (8 * C * b + 8 * C * 4) + R * 28
    mlp_fwd_flop = 2 * (C * H + H * H + H)
    fwd_flop = S_total * mlp_fwd_flop
    bwd_flop = S_total * 3 * mlp_fwd_flop  # recomputed forward + backward
    dom = "bwd" if t_bwd >= t_fwd else "fwd"
    t_dom, bytes_dom, flop_dom = (t_bwd, bwd_bytes, bwd_flop) if dom == "bwd" else (t_fwd, fwd_bytes, fwd_flop)
    achieved_gbs = bytes_dom / t_dom / 1e9
    achieved_tflops = flop_dom / t_dom / 1e12
    mfma_peak = BF16_PEAK_TFLOPS if bf16 else FP32_PEAK_TFLOPS
    log(f"  project_prepare (counting sort) {t_prep * 1e3:.2f} ms, "
        f"project_fwd {t_fwd * 1e3:.2f} ms, project_bwd {t_bwd * 1e3:.2f} ms ({S_total / R:.1f} samples/ray, "
        f"{iv.num_intervals} intervals); dominant {dom}: {achieved_gbs:.0f} GB/s algorithmic, "
        f"{achieved_tflops:.1f} TFLOP/s decoder")
    mode = ("learned" if ps is not None else "bf16") if bf16 else "plain"
    traffic = measured_traffic(mode, f"project_{dom}")
    return {
        "kernel_ms": {"project_prepare": t_prep * 1e3,
                      ("project_fwd_prepared" if node_fwd else "project_fwd"): t_fwd * 1e3,
                      "project_bwd_prepared": t_bwd * 1e3},
        "roofline": {
            "bound": "hbm", "kernel": f"project_{dom}", "achieved": achieved_gbs, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved_gbs / HBM_PEAK_GBS, "traffic": traffic,
            "algorithmic_bytes_per_launch": bytes_dom, "samples_per_launch": S_total,
            "mfma": {"achieved": achieved_tflops, "peak": mfma_peak, "unit": "TFLOP/s",
                     "frac": achieved_tflops / mfma_peak},
        },
    }


def measured_traffic(mode, kernel):
    """HBM bytes per call of `kernel` (project_fwd, or project_bwd: the
    adjoint's kernels summed) for this mode's interval lists, from the
    committed rocprofv3 PMC summary of the 512^3 workload (FETCH_SIZE +
    WRITE_SIZE, copy-calibrated), or None when that combination was not
    measured."""
    path = os.path.join(HERE, "profiles", "traffic_512cube.json")
    if mode not in ("plain", "learned", "bf16") or not os.path.exists(path):
        return None
    with open(path) as fh:
        data = json.load(fh)
    kernels = data.get("kernels")
    if not isinstance(kernels, list):
        return None
    parts = [rec for rec in kernels if rec.get("leg") == mode and rec.get("call") == kernel]
    return sum(rec["hbm_bytes_per_launch"] for rec in parts) if parts else None


# --------------------------------------------------------------------------
# CPU baseline: the oracle's FP32 training step on the host cores
# --------------------------------------------------------------------------
def host_cpu():
    model = None
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except OSError:
        pass
    affinity = len(os.sched_getaffinity(0))
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or affinity
    return model, affinity, min(threads, affinity)


def _median_runs(fn, runs):
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    return statistics.median(times), times


def cpu_baseline(N, tree, grid0, mlp0, rays, tgt, K, runs, dev):
    """The oracle's FP32 training step (forward, adjoint, Adam; OpenMP over
    rays) on the host cores: at the 512^3 workload's shape (2048 rays of a
    batch, the grown tree and field) and at the 128^3 CPU configuration's
    own shape (uniform depth-3 octree, C = 8, 4096 rays x ~64 samples).
    Median of `runs` timed runs after one warm-up."""
    import torch

    from adaptivetensordf_amd import phantom

    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import oracle as O

    model, affinity, threads = host_cpu()
    # 512^3 workload shape
    active = tree.node_active.cpu().numpy()
    nr = 2048
    o, d, t = (x[:nr].cpu().numpy() for x in (rays.origin, rays.direction, tgt))
    fm = O.FusedModel(tree.node_position_min.cpu().numpy(), tree.node_position_max.cpu().numpy(), active,
                      grid0.cpu().numpy(), mlp0.cpu().numpy(), C, K=K, jitter_mode=2, seed=7)
    tb = O.TrainBaseline(fm, lr=1e-3)
    tb.train_step(o, d, t, threads, do_adam=False)
    t1, _ = _median_runs(lambda: tb.train_step(o, d, t, threads, do_adam=False), runs)
    t_adam, _ = _median_runs(lambda: tb.train_step(o[:1], d[:1], t[:1], threads, do_adam=True), 3)
    R = rays.size()
    c1_value = R / (t1 / nr * R + t_adam)
    # the 128^3 CPU configuration: 20 ellipsoids (seed 0) + noise, 128^2
    # detector, uniform depth 3, C = 8, 4096 rays, ~64 samples per ray
    t0 = O.tree_build(3)
    act0 = (t0["depth"] == 3).astype(np.int32)
    ang = phantom.tilt_angles(61)
    tl, px, py = phantom.sample_pixels(4096, 61, 128, 128, 0)
    o0, d0 = phantom.rays_for_pixels(tl, px, py, 128, 128, ang)
    vol0 = voxel_phantom(phantom.ellipsoids(20, 0), 128, 0.01, 3, dev)
    tg0 = volume_projection(vol0, torch.from_numpy(o0).to(dev), torch.from_numpy(d0).to(dev)).float().cpu().numpy()
    K0, s0 = None, 0.0
    for k in range(1, 40):
        r = O.traverse(t0["node_min"], t0["node_max"], O.active_ids(act0), o0, d0, k)
        s0 = r["weight"].shape[0] / 4096
        if s0 >= 64:
            K0 = k
            break
    fm0 = O.FusedModel(t0["node_min"], t0["node_max"], act0, phantom.init_grid(512, G, 8, std=0.01, seed=0).numpy(),
                       phantom.init_mlp(8, H, seed=0).numpy(), 8, K=K0, jitter_mode=2, seed=7)
    tb0 = O.TrainBaseline(fm0, lr=1e-3)
    tb0.train_step(o0, d0, tg0, threads)
    t_c0, _ = _median_runs(lambda: tb0.train_step(o0, d0, tg0, threads), runs)
    return {"value": c1_value, "unit": "rays/s", "cores": threads, "kind": "port",
            "cpu_model": model, "affinity_cpus": affinity, "runs": runs, "statistic": "median",
            "sample": f"512^3 workload shape: median of {runs} FP32 fwd+bwd passes over {nr} rays of a 2^18-ray batch "
                      f"(grown tree, K={K}) = {t1:.2f} s each, plus one dense Adam pass over all "
                      f"{grid0.numel() + mlp0.numel()} parameters ({t_adam:.2f} s), scaled to one {R}-ray step; "
                      f"oracle restatement, OpenMP over rays on {threads} threads of {model}",
            "c0_128cube": {"value": 4096 / t_c0, "unit": "rays/s", "s_per_iteration": t_c0,
                           "s_for_200_iterations": 200 * t_c0, "samples_per_ray": round(s0, 1), "K": K0,
                           "sample": f"median of {runs} full training iterations (4096 rays, fwd+bwd, Adam on "
                                     f"all parameters) at the 128^3 CPU configuration's shape, {threads} threads"}}


# --------------------------------------------------------------------------
def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "run it with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    if args.copy_calibration:
        x = torch.empty(1 << 27, dtype=torch.float32, device="cuda:0").fill_(1.0)
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        log("copied 3 x 512 MiB (read 512 MiB + write 512 MiB per copy)")
        return

    from adaptivetensordf_amd import native, phantom

    N = native()
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    n = args.volume
    batches, ev = build_data(N, args, dev)
    if args.load_state:
        tree, grid0, mlp0, grow_log = load_state(N, args.load_state, dev)
        log(f"loaded grown state: {tree.NumActiveNodes()} active nodes")
    else:
        tree, grid0, mlp0, grow_log = grow_tree(N, args, batches, dev)
    if args.save_state:
        save_state(args.save_state, tree, grid0, mlp0, grow_log)
        log(f"saved the grown state to {args.save_state}")
        return
    num_active = int(tree.NumActiveNodes())
    K, s_eff = choose_K(N, tree, batches[0][0], args.samples_per_ray)
    judge = Judge(N, tree, ev, K) if args.profile_steps == 0 else None
    steps = args.profile_steps if args.profile_steps > 0 else args.steps
    log(f"workload: {num_active} active nodes, K {K} ({s_eff:.1f} samples/ray)")

    if args.mode == "vm":
        vm_leg(N, tree, batches, K, s_eff, n, args, steps)
        return

    pp = N.ProposalParams()
    pp.mix = args.mix
    pp.kl_threshold = args.kl_threshold
    pp.alloc_extra = args.alloc_extra if args.proposal_mode == "alloc" else 0
    pp.target_mode = 1 if args.proposal_target == "variation" else 0
    pp.target_power = args.target_power
    K_learned = args.alloc_k if args.proposal_mode == "alloc" else K
    prop_rays = []
    for rays, _ in batches:
        R = rays.size()
        idx = torch.arange(0, R, max(1, R // args.proposal_rays), device=dev)[: args.proposal_rays]
        prop_rays.append(N.RayList(rays.origin.index_select(0, idx).contiguous(),
                                   rays.direction.index_select(0, idx).contiguous()))
    legs = {}
    if args.mode in ("plain", "learned"):
        legs["plain"] = run_leg(N, "plain", tree, batches, K, field_params(N, False, not args.ray_major_forward), grid0,
                                mlp0, False, None, judge, args, steps)
    if args.mode == "learned":
        ps = N.MakeProposalState(phantom.init_proposal(0).to(dev), 3e-3, pp)
        legs["learned"] = run_leg(N, "learned+bf16", tree, batches, K_learned,
                                  field_params(N, True, not args.ray_major_forward), grid0, mlp0, True, ps, judge, args,
                                  steps, prop_rays, K_train=K)
    if args.profile_steps > 0:
        if args.mode == "learned":
            legs["bf16"] = run_leg(N, "bf16", tree, batches, K, field_params(N, True, not args.ray_major_forward),
                                   grid0, mlp0, True, None, judge, args, steps)
        # the launches the full run's roofline times: 1 + 5 of project_fwd,
        # then 1 + 5 of project_bwd per leg, last in the run
        for key, leg in legs.items():
            bf = key != "plain"
            kernel_times(N, tree, batches[0][0], leg_sampling(N, leg["K"]), field_params(N, bf, not args.ray_major_forward),
                         leg["state"], bf, ps if key == "learned" else None)
        return
    if args.mode in ("bf16", "learned"):
        legs["bf16"] = run_leg(N, "bf16", tree, batches, K, field_params(N, True, not args.ray_major_forward), grid0,
                               mlp0, True, None, judge,
                               args, steps)
    if args.mode == "learned":
        K32, s32 = choose_K(N, tree, batches[0][0], legs["learned"]["samples_per_ray"])
        legs["strat32"] = run_leg(N, "bf16 stratified at the learned sample count", tree, batches, K32,
                                  field_params(N, True, not args.ray_major_forward), grid0, mlp0, True, None, judge,
                                  args, steps)
        legs["strat32"]["K"] = K32

    head = legs["learned" if args.mode == "learned" else args.mode]
    bf16_head = args.mode != "plain"
    kern = kernel_times(N, tree, batches[0][0], leg_sampling(N, head["K"]),
                        field_params(N, bf16_head, not args.ray_major_forward), head["state"], bf16_head,
                        ps if args.mode == "learned" else None)
    summary = {}
    for key, leg in legs.items():
        if leg is head:
            continue
        summary[key] = {"value": leg["rays_per_s"], "ms_per_step": leg["ms_per_step"],
                        "samples_per_ray": leg["samples_per_ray"], "final_mse": leg["mse_curve"][-1][1],
                        "loss_first_last": leg["loss_first_last"]}
        if key == "strat32":
            summary[key]["K"] = leg["K"]
    if "plain" in legs and head is not legs["plain"]:
        kp = kernel_times(N, tree, batches[0][0], leg_sampling(N, legs["plain"]["K"]),
                          field_params(N, False, not args.ray_major_forward), legs["plain"]["state"], False, None)
        summary["plain"].update({"kernel_ms": kp["kernel_ms"], "roofline": kp["roofline"],
                                 "speedup_learned_over_plain": head["rays_per_s"] / legs["plain"]["rays_per_s"]})
    mse = None
    if args.mode == "learned":
        final = {k: leg["mse_curve"][-1][1] for k, leg in legs.items()}
        # the median of the judge points over the last 500 iterations (21 by
        # default): single points carry transient spikes of the training noise
        med5 = {k: statistics.median(v for it, v in leg["mse_curve"] if it >= leg["iterations"] - 500)
                for k, leg in legs.items()}
        sp_own = N.SamplingParams()
        sp_own.max_samples_per_node = legs["learned"]["K"]
        ivl, fb = N.CreateProposalIntervals(tree, ps, ev[0], sp_own)
        P_own = N.ProjectIntervals(tree, ev[0], ivl, legs["learned"]["state"].grid, legs["learned"]["state"].mlp,
                                   sp_own, field_params(N, True))
        mse = {
            "judge": f"stratified K={K} (~{s_eff:.0f} samples/ray, midpoints) on {ev[0].size()} held-out rays vs their "
                     f"targets (noisy-volume line integrals), every 100 iterations of each leg",
            "final": final,
            "median_last500": med5,
            "curve": {k: leg["mse_curve"] for k, leg in legs.items()},
            "ratio_learned_over_plain": med5["learned"] / med5["plain"],
            "ratio_learned_over_plain_final_point": final["learned"] / final["plain"],
            "learned_within_1pct_or_better": med5["learned"] / med5["plain"] <= 1.01,
            "ratio_learned_over_strat32": med5["learned"] / med5["strat32"],
            "ratio_learned_over_strat32_final_point": final["learned"] / final["strat32"],
            "learned_own_sampler": float(((P_own.double() - ev[1].double()) ** 2).mean().item()),
            "learned_own_sampler_samples_per_ray": ivl.num_samples / ev[0].size(),
            "learned_own_sampler_fallback_frac": float(fb.float().mean().item()),
            # legs whose field is as good as the plain leg's (median of the
            # judge points of the last 500 iterations within 1 %, or below), with their speed
            "equal_mse": {k: {"mse_over_plain": med5[k] / med5["plain"],
                              "speedup_over_plain": legs[k]["rays_per_s"] / legs["plain"]["rays_per_s"]}
                          for k in legs if k != "plain" and med5[k] / med5["plain"] <= 1.01},
        }
        eq = mse["equal_mse"]
        best = max(eq, key=lambda k: eq[k]["speedup_over_plain"]) if eq else None
        r_ls = mse["ratio_learned_over_strat32"]
        mse["finding"] = (
            (f"at equal projection MSE (within 1 %) the fastest leg is {best}, "
             f"{eq[best]['speedup_over_plain']:.2f}x the plain FP32 leg's rays/s" if best
             else "no leg reaches the plain FP32 leg's projection MSE within 1 %")
            + f"; the learned leg's MSE is {mse['ratio_learned_over_plain']:.2f}x the plain leg's and {r_ls:.3f}x "
              f"that of BF16 stratified sampling at the same sample count ("
            + ("the learned allocation gains over stratified sampling)" if r_ls < 0.99
               else "the learned allocation gains nothing over stratified sampling)"))
        log(f"held-out MSE (median of the judge points of the last 500 iterations): "
            + ", ".join(f"{k} {v:.4e}" for k, v in med5.items()) + f"; {mse['finding']}")

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(N, tree, grid0, mlp0, batches[0][0], batches[0][1], K, args.cpu_runs, dev)
        log(f"cpu baseline: {cpu['value']:.0f} rays/s (512^3 shape), {cpu['c0_128cube']['value']:.0f} rays/s "
            f"(128^3 shape) on {cpu['cores']} threads of {cpu['cpu_model']}")
    b = 2 if bf16_head else 4
    if args.mode == "learned":
        pr = head["proposal"]
        how = (f"allocation: the samples of the K={pr['base_K']} stratified list plus "
               f"{pr['extra_samples']} along the proposal's CDF" if pr["mode"] == "alloc" else "32 learned cells")
        workload = (f"learned proposal ({how}) + BF16 grids training step, {n}^3 noisy ellipsoid phantom "
                    f"(200, sigma 0.01), "
                    f"61 tilts, {n}x{n} detector, octree grown by the error-driven structure loop to {num_active} "
                    f"nodes (depth<=5); the proposal retraining (every 100 iterations, "
                    f"{pr['retrain_ms_mean'] or pr['initial_update_ms']:.1f} ms each) is amortised into ms_per_step "
                    f"({pr['updates_in_timed_region']} retrainings fell inside the timed steps); "
                    f"{args.batches} x {args.rays} distinct training rays; plain FP32, BF16 "
                    f"and BF16-stratified-at-equal-samples legs trained in the same run")
        dtype = "bf16 grids / f32 accumulate"
    else:
        workload = (f"{'plain FP32' if not bf16_head else 'BF16 grids'} training step, {n}^3 noisy ellipsoid phantom, "
                    f"61 tilts, {n}x{n} detector, octree grown by the error-driven structure loop to {num_active} "
                    f"nodes (depth<=5)")
        dtype = "f32" if not bf16_head else "bf16 grids / f32 accumulate"
    out = {
        "metric": METRIC,
        "value": head["rays_per_s"],
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (voxelised noisy ellipsoid phantom, its line integrals as targets; random-init field "
                "trained in the run)",
        "config": {
            "workload": workload,
            "rays_per_batch": args.rays, "stratified_samples_per_ray": round(s_eff, 2), "K": K,
            "samples_per_ray": round(head["samples_per_ray"], 2), "raw_ms_per_step": head["raw_ms_per_step"],
            "active_nodes": num_active, "channels": C, "grid": G, "decoder": f"{C}-{H}-{H}-1",
            "parallelism": "single GPU",
            "step": "sampling + fused fwd + MSE + fused adjoint + Adam(grids, decoder)",
            "train_iterations_per_leg": head["iterations"], "grow": grow_log,
            "loss_first_last": head["loss_first_last"],
            "step_model_bytes_per_ray": 60 + head["samples_per_ray"] * (16 * C * b + 32 * C),
            "kernel_ms": kern["kernel_ms"],
        },
        "roofline": kern["roofline"],
        "cpu_baseline": cpu,
    }
    if args.mode == "learned":
        out["config"]["proposal"] = head["proposal"]
        out["projection_mse"] = mse
    out["legs"] = summary
    print(json.dumps(out), flush=True)


def vm_leg(N, tree, batches, K, s_eff, n, args, steps):
    """Training step of the tensorial geometry (GeometryTensorQO): the fused
    path (TrainStepVM: the plane x line and explicit-grid gathers inside
    the projection kernels and their adjoint, MSE, Adam on every field)
    beside the composed operators (CreateSamplesForRays, ComputeImageVM,
    autograd, Adam) on the same fields and rays.  One JSON line of its own
    (not the headline); `value` is the fused path."""
    import torch

    from adaptivetensordf_amd import phantom

    dev = torch.device("cuda:0")
    lo, hi = [-0.6, -0.6, -0.6], [0.6, 0.6, 0.6]
    roi = tree.RoiSplit(lo, hi, 0.0)
    n_in, n_out = int(roi.in_roi_node_index.numel()), int(roi.out_roi_node_index.numel())
    line = (G - 1) * (1 << DEPTH) + 1
    f = N.VMFields()
    gen = torch.Generator(device="cpu").manual_seed(0)

    def param(*shape):
        return (torch.randn(*shape, generator=gen) * 0.01).to(dev).requires_grad_()

    f.plane_yz, f.plane_xz, f.plane_xy = (param(n_in, C, G, G) for _ in range(3))
    f.line_x, f.line_y, f.line_z = (param(1, C, line, 1) for _ in range(3))
    f.grid_out = param(n_out, C, G, G, G)
    f.tensor_op = "add"
    f.roi_min, f.roi_max = lo, hi
    mlp = phantom.init_mlp(C, H, seed=0).to(dev).requires_grad_()
    fp = field_params(N, False)
    sp = N.SamplingParams()
    sp.max_samples_per_node = K
    sp.jitter_mode = 2

    # composed operators
    params = [f.plane_yz, f.plane_xz, f.plane_xy, f.line_x, f.line_y, f.line_z, f.grid_out, mlp]
    moments = [(torch.zeros_like(p), torch.zeros_like(p)) for p in params]
    state = {"it": 0, "loss": None}

    def step_composed():
        rays, tgt = batches[state["it"] % len(batches)]
        samples = tree.CreateSamplesForRays(rays, K, True)
        img = N.ComputeImageVM(tree, roi, f, mlp, fp, samples, rays.size())
        loss = (img[0] - tgt).square().mean()
        grads = torch.autograd.grad(loss, params)
        state["it"] += 1
        for p, g, (m, v) in zip(params, grads, moments):
            N.AdamStep(p.data, g.contiguous(), m, v, state["it"], 1e-3)
        state["loss"] = loss.detach()

    # fused (state copied before the composed leg trains the fields)
    st = N.MakeVMTrainState(tree, roi, f, mlp, 1e-3, fp)
    fstate = {"it": 0}

    def step_fused():
        rays, tgt = batches[fstate["it"] % len(batches)]
        sp.jitter_seed = 1000 + fstate["it"]
        N.TrainStepVM(tree, roi, st, rays, tgt, sp, fp)
        fstate["it"] += 1

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    el_c = timed(step_composed)
    loss_c = float(state["loss"].item())
    el_f = timed(step_fused)
    loss_f = float(st.loss.item())
    R = batches[0][0].size()
    kern = vm_kernel_times(N, tree, roi, st, batches[0][0], sp, fp)
    log(f"vm composed: {el_c * 1e3 / steps:.2f} ms/step, {R * steps / el_c / 1e6:.3f} M rays/s; fused: "
        f"{el_f * 1e3 / steps:.2f} ms/step, {R * steps / el_f / 1e6:.3f} M rays/s; {n_in} VM / {n_out} explicit nodes")
    print(json.dumps({
        "metric": "fwd+bwd rays/s, tensorial (VM) field through the fused projection kernels, 512^3, 61 tilts",
        "value": R * steps / el_f, "unit": "rays/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": el_f * 1e3 / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (voxelised noisy ellipsoid phantom projections, random-init field)",
        "config": {"workload": f"tensorial geometry training step, {n}^3 phantom, 61 tilts, {n}x{n} detector, "
                               f"grown octree, ROI {lo}..{hi}: {n_in} plane/line nodes, {n_out} explicit-grid nodes, "
                               f"planes {G}^2, lines of {line}, tensor_op add; step = sampling + fused fwd + MSE + "
                               f"fused adjoint + Adam on planes, lines, grids and decoder",
                   "rays_per_batch": R, "K": K, "stratified_samples_per_ray": round(s_eff, 2), "channels": C,
                   "grid": G, "decoder": f"{C}-{H}-{H}-1", "parallelism": "single GPU",
                   "loss_last": {"fused": loss_f, "composed": loss_c},
                   "composed": {"ms_per_step": el_c * 1e3 / steps, "rays_per_s": R * steps / el_c},
                   "speedup_fused_over_composed": el_c / el_f, "kernel_ms": kern["kernel_ms"],
                   "samples_in_roi": kern["samples_in_roi"], "samples_out_roi": kern["samples_out_roi"],
                   "adjoint_form": kern["adjoint_form"]},
        "roofline": kern["roofline"], "cpu_baseline": None}), flush=True)


def vm_kernel_times(N, tree, roi, st, rays, sp, fp):
    """HIP-event durations of atdf_project_fwd / atdf_project_bwd with the
    tensorial field over one batch, and the adjoint's roofline.  Algorithmic
    bytes: an in-ROI sample gathers 3 planes x 4 corners + 3 lines x 2 corners
    of C f32 (72 C B) and the adjoint scatters as many; an explicit-grid
    sample 8 corners x C f32 each way; 28 B per ray."""
    import ctypes as Ct

    import torch

    from adaptivetensordf_amd import capi

    iv = tree.CreateIntervalsForRays(rays, sp)
    R = rays.size()
    flag = roi.in_roi_flag.to(iv.node.device)[iv.node.long()].bool()
    s_in = int(iv.count[flag].sum().item())
    s_out = int(iv.count[~flag].sum().item())
    n2p = roi.prefix_inroi.to(torch.int32).contiguous()
    n2g = roi.prefix_outroi.to(torch.int32).contiguous()
    gpl, gln = torch.zeros_like(st.planes), torch.zeros_like(st.lines)
    ggr, gml = torch.zeros_like(st.grid), torch.zeros_like(st.mlp)
    vm = capi.VM(st.planes.data_ptr(), st.lines.data_ptr(), st.planes.shape[0], n2p.data_ptr(), st.planes.shape[2],
                 st.lines.shape[1], st.roi_min_x, st.roi_max_x, gpl.data_ptr(), gln.data_ptr())
    f = capi.Field(C, H, G, capi.GRID_F32, 2.0, 0, st.grid.data_ptr() if st.grid.numel() else None, st.grid.shape[0],
                   n2g.data_ptr(), st.mlp.data_ptr(), capi.DECODER_F32, 0, Ct.pointer(vm))
    csp = capi.Sampling(sp.max_samples_per_node, 256, sp.jitter_mode, 0, sp.jitter_seed, None, None)
    civ = capi.Intervals(iv.node.data_ptr(), iv.tmin.data_ptr(), iv.tmax.data_ptr(), iv.count.data_ptr(),
                         iv.first.data_ptr(), iv.ray.data_ptr(), iv.num_intervals)
    packed = tree.PackedTree()
    s = capi.stream_ptr()
    P = torch.empty(R, device=rays.origin.device)
    dP = torch.randn(R, device=rays.origin.device) * 1e-6
    ws_bytes = capi.lib().atdf_project_bwd_workspace_bytes(R, iv.num_intervals, Ct.byref(f))
    staging = capi.lib().atdf_project_bwd_staging_bytes(iv.num_intervals, iv.num_samples, Ct.byref(f)) \
        if fp.two_pass_adjoint else 0
    ws = torch.empty(ws_bytes + staging, dtype=torch.uint8, device=rays.origin.device)
    a = (capi.ptr(packed), tree.NumNodes(), capi.ptr(rays.origin), capi.ptr(rays.direction), R, Ct.byref(csp),
         capi.ptr(iv.ray_offset), Ct.byref(civ), Ct.byref(f))

    def fwd():
        capi.call("atdf_project_fwd", *a, capi.ptr(P), s)

    def bwd():
        capi.call("atdf_project_bwd", *a, capi.ptr(dP), capi.ptr(ggr) if ggr.numel() else None, capi.ptr(gml),
                  capi.ptr(ws), ws.numel(), s)

    def timed(fn, reps=5):
        fn()
        torch.cuda.synchronize()
        cur = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(cur)
        for _ in range(reps):
            fn()
        e1.record(cur)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps / 1e3

    t_fwd = timed(fwd)
    t_bwd = timed(bwd)
    bwd_bytes = s_in * 2 * 72 * C + s_out * 2 * 8 * C * 4 + R * 28
    gbs = bwd_bytes / t_bwd / 1e9
    log(f"  vm project_fwd {t_fwd * 1e3:.2f} ms, project_bwd {t_bwd * 1e3:.2f} ms ("
        f"{'two-pass' if staging else 'fused'}, counting sort included; "
        f"{s_in} in-ROI + {s_out} explicit samples): {gbs:.0f} GB/s algorithmic")
    return {"kernel_ms": {"project_fwd": t_fwd * 1e3, "project_bwd": t_bwd * 1e3}, "samples_in_roi": s_in,
            "samples_out_roi": s_out,
            "adjoint_form": "two-pass" if staging else "fused",
            "roofline": {"bound": "hbm", "kernel": "project_bwd (tensorial)", "achieved": gbs, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS, "traffic": None,
                         "algorithmic_bytes_per_launch": bwd_bytes}}


if __name__ == "__main__":
    main()
