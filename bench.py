#!/usr/bin/env python3
"""Benchmark: fwd+bwd rays/s of one training step of the reconstruction hot
path on one MI355X.

Workload (the 512^3 configuration): 200-ellipsoid phantom (seed 1), 61 tilts
from -60 to +60 degrees, 512 x 512 detector, adaptive octree up to depth 5,
11^3 x 16-channel node grids, batches of 2^18 rays.  One step = sampling ->
fused forward -> MSE -> fused adjoint -> Adam on every grid and decoder
parameter, all on the device.  Inputs (rays, targets) are resident in HBM
before the timed region.

Modes:
  plain    FP32 grids and decoder, stratified samples (~128 per ray)
  bf16     BF16 grids (FP32 masters) and BF16-MFMA decoder, stratified samples
  learned  (default) BF16 as above with the learned proposal: 32 samples per
           ray, stratified ~128 on the KL fallback rays, the proposal trained
           online every 100 iterations inside the timed loop.  The plain FP32
           leg runs in the same process; both legs train the same number of
           iterations from the same initial field and are then judged on
           held-out rays with the stratified ~128-sample projection against
           the analytic targets ("projection MSE").

Prints one JSON line (see the keys below).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "fwd+bwd rays/s at 512³, 61 tilts; learned+bf16 vs plain at equal proj. MSE"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--mode", choices=["plain", "bf16", "learned", "vm"], default="learned",
                    help="vm: the tensorial (plane x line) field through the composed operator path")
    ap.add_argument("--train-iters", type=int, default=300,
                    help="learned mode: iterations each leg trains before the projection-MSE check")
    ap.add_argument("--mix", type=float, default=0.3, help="learned mode: uniform mixing of the proposal pdf")
    ap.add_argument("--kl-threshold", type=float, default=0.3, help="learned mode: fallback threshold")
    ap.add_argument("--proposal-rays", type=int, default=1 << 15, help="rays per proposal training pass")
    ap.add_argument("--volume", type=int, default=512)
    ap.add_argument("--rays", type=int, default=1 << 18)
    ap.add_argument("--batches", type=int, default=4)
    ap.add_argument("--budget", type=int, default=8192, help="max active octree nodes")
    ap.add_argument("--samples-per-ray", type=float, default=128.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--node-major-forward", action="store_true",
                    help="training step with the node-major forward over the adjoint's sort")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--profile-steps", type=int, default=0, help="only run this many steps (profiling)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy a 512 MiB buffer 3 times (scale factor for HBM byte counters)")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


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
    C, G, H, depth = 16, 11, 64, 5

    # ------------------------------------------------------------ workload
    t0 = time.time()
    ell = phantom.ellipsoids(200, 1)
    tree = N.HyperTreeBase(3, depth)
    tree.settree()
    active = phantom.adaptive_active_mask(tree.node_position_min.numpy(), tree.node_position_max.numpy(),
                                          tree.node_children.numpy(), tree.node_depth.numpy(), ell, args.budget, 3,
                                          depth)
    tree.node_active = torch.from_numpy(active)
    tree.UpdateActive()
    tree.to("cuda")
    num_active = int(tree.NumActiveNodes())
    angles = phantom.tilt_angles(61)

    def make_batch(nrays, seed):
        tl, px, py = phantom.sample_pixels(nrays, 61, n, n, seed)
        o, d = phantom.rays_for_pixels(tl, px, py, n, n, angles)
        o_t, d_t = torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev)
        tgt = phantom.analytic_projection_torch(o_t, d_t, ell).float()
        return N.RayList(o_t, d_t), tgt.contiguous()

    batches = [make_batch(args.rays, 100 + b) for b in range(args.batches)]
    # K so that the mean number of samples per ray is ~samples_per_ray
    sp = N.SamplingParams()
    sp.jitter_mode = 2
    sp.jitter_seed = 7
    best = None
    for K in range(1, 65):
        sp.max_samples_per_node = K
        s_eff = tree.CreateIntervalsForRays(batches[0][0], sp).num_samples / args.rays
        if best is None or abs(s_eff - args.samples_per_ray) < abs(best[1] - args.samples_per_ray):
            best = (K, s_eff)
        if s_eff > args.samples_per_ray:
            break
    K, s_eff = best
    sp.max_samples_per_node = K
    grid0 = phantom.init_grid(num_active, G, C, std=0.01, seed=0, device=dev)
    mlp0 = phantom.init_mlp(C, H, seed=0).to(dev)
    log(f"setup {time.time() - t0:.1f}s: active nodes {num_active}, K {K}, mean samples/ray {s_eff:.1f}")

    def field(bf16):
        fp = N.FieldParams()
        fp.C, fp.G, fp.H = C, G, H
        fp.bf16_decoder = bf16
        fp.node_forward = args.node_major_forward
        return fp

    steps = args.profile_steps if args.profile_steps > 0 else args.steps

    if args.mode == "vm":
        vm_leg(N, tree, batches, K, s_eff, C, G, H, depth, n, args, steps)
        return
    if args.mode != "learned":
        bf16 = args.mode == "bf16"
        leg = run_leg(N, tree, batches, sp, field(bf16), grid0, mlp0, bf16, None, args, steps)
        if args.profile_steps > 0:
            return
        out_leg = leg
        kern = kernel_times(N, tree, batches[0][0], sp, field(bf16), leg["state"], bf16, None)
        plain_summary = None
        mse = None
        workload = (f"{'plain FP32' if not bf16 else 'BF16 grids'} training step, {n}^3 phantom (200 ellipsoids), "
                    f"61 tilts, {n}x{n} detector, adaptive octree depth<=5")
        dtype = "f32" if not bf16 else "bf16 grids / f32 accumulate"
    else:
        plain = run_leg(N, tree, batches, sp, field(False), grid0, mlp0, False, None, args, steps)
        pp = N.ProposalParams()
        pp.mix = args.mix
        pp.kl_threshold = args.kl_threshold
        ps = N.MakeProposalState(phantom.init_proposal(0).to(dev), 3e-3, pp)
        learned = run_leg(N, tree, batches, sp, field(True), grid0, mlp0, True, ps, args, steps)
        if args.profile_steps > 0:
            return
        out_leg = learned
        kern = kernel_times(N, tree, batches[0][0], sp, field(True), learned["state"], True, ps)
        kern_plain = kernel_times(N, tree, batches[0][0], sp, field(False), plain["state"], False, None)
        # projection MSE on held-out rays, judged by the stratified sampler
        ev_rays, ev_tgt = make_batch(1 << 16, 999)
        sp_judge = N.SamplingParams()
        sp_judge.max_samples_per_node = K
        iv_j = tree.CreateIntervalsForRays(ev_rays, sp_judge)

        def judge(leg, fp):
            P = N.ProjectIntervals(tree, ev_rays, iv_j, leg["state"].grid, leg["state"].mlp, sp_judge, fp)
            return float(((P.double() - ev_tgt.double()) ** 2).mean().item())

        ivl, fb = N.CreateProposalIntervals(tree, ps, ev_rays, sp_judge)
        P_own = N.ProjectIntervals(tree, ev_rays, ivl, learned["state"].grid, learned["state"].mlp, sp_judge,
                                   field(True))
        mse = {
            "judge": f"stratified K={K} (~{s_eff:.0f} samples/ray, midpoints) on {ev_rays.size()} held-out rays "
                     f"vs analytic targets, after {args.train_iters} training iterations of each leg",
            "plain": judge(plain, field(False)),
            "learned_bf16": judge(learned, field(True)),
            "learned_bf16_own_sampler": float(((P_own.double() - ev_tgt.double()) ** 2).mean().item()),
            "learned_own_sampler_samples_per_ray": ivl.num_samples / ev_rays.size(),
            "learned_own_sampler_fallback_frac": float(fb.float().mean().item()),
        }
        mse["ratio_learned_over_plain"] = mse["learned_bf16"] / mse["plain"]
        # "equal projection MSE (within 1%)": the learned leg is no worse than 1% above plain
        mse["learned_within_1pct_or_better"] = mse["ratio_learned_over_plain"] <= 1.01
        log(f"projection MSE (judge): plain {mse['plain']:.5e}, learned+bf16 {mse['learned_bf16']:.5e} "
            f"(ratio {mse['ratio_learned_over_plain']:.4f}); learned own sampler {mse['learned_bf16_own_sampler']:.5e}")
        plain_summary = {
            "value": plain["rays_per_s"], "ms_per_step": plain["ms_per_step"],
            "speedup_learned_over_plain": learned["rays_per_s"] / plain["rays_per_s"],
            "kernel_ms": kern_plain["kernel_ms"], "roofline": kern_plain["roofline"],
            "loss_first_last": plain["loss_first_last"],
        }
        workload = (f"learned proposal + BF16 grids training step (proposal trained online every 100 iterations "
                    f"inside the timed loop), {n}^3 phantom (200 ellipsoids), 61 tilts, {n}x{n} detector, "
                    f"adaptive octree depth<=5; plain FP32 leg timed in the same run")
        dtype = "bf16 grids / f32 accumulate"

    rays_per_s = out_leg["rays_per_s"]
    cpu = None
    if not args.no_cpu_baseline:
        rays0, tgt0 = batches[0]
        cpu = cpu_baseline(tree, active, grid0, mlp0, rays0, tgt0, K, C, G, args.cpu_seconds)
    b = 2 if args.mode != "plain" else 4
    out = {
        "metric": METRIC,
        "value": rays_per_s,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": out_leg["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (analytic ellipsoid phantom projections, random-init field)",
        "config": {
            "workload": workload,
            "rays_per_batch": args.rays, "stratified_samples_per_ray": round(s_eff, 2), "K": K,
            "samples_per_ray": round(out_leg["samples_per_ray"], 2),
            "active_nodes": num_active, "channels": C, "grid": G, "decoder": f"{C}-{H}-{H}-1",
            "parallelism": "single GPU",
            "step": "sampling + fused fwd + MSE + fused adjoint + Adam(grids, decoder)",
            "loss_first_last": out_leg["loss_first_last"],
            "step_model_bytes_per_ray": 60 + out_leg["samples_per_ray"] * (16 * C * b + 32 * C),
            "kernel_ms": kern["kernel_ms"],
        },
        "roofline": kern["roofline"],
        "cpu_baseline": cpu,
    }
    if args.mode == "learned":
        out["config"]["proposal"] = learned["proposal"]
        out["plain_fp32"] = plain_summary
        out["projection_mse"] = mse
    print(json.dumps(out), flush=True)


def vm_leg(N, tree, batches, K, s_eff, C, G, H, depth, n, args, steps):
    """Training step of the tensorial geometry (GeometryTensorQO) through the
    composed operators: CreateSamplesForRays, ComputeImageVM (VM gathers for
    in-ROI nodes, explicit grids for the rest, DecodeFeatures, integration),
    MSE, autograd backward (slot-major gather adjoints, decoder adjoint),
    Adam on every field.  One JSON line of its own (not the headline)."""
    import torch

    dev = torch.device("cuda:0")
    lo, hi = [-0.6, -0.6, -0.6], [0.6, 0.6, 0.6]
    roi = tree.RoiSplit(lo, hi, 0.0)
    n_in, n_out = int(roi.in_roi_node_index.numel()), int(roi.out_roi_node_index.numel())
    line = (G - 1) * (1 << depth) + 1
    f = N.VMFields()
    gen = torch.Generator(device="cpu").manual_seed(0)

    def param(*shape):
        return (torch.randn(*shape, generator=gen) * 0.01).to(dev).requires_grad_()

    f.plane_yz, f.plane_xz, f.plane_xy = (param(n_in, C, G, G) for _ in range(3))
    f.line_x, f.line_y, f.line_z = (param(1, C, line, 1) for _ in range(3))
    f.grid_out = param(n_out, C, G, G, G)
    f.tensor_op = "add"
    f.roi_min, f.roi_max = lo, hi
    from adaptivetensordf_amd import phantom

    mlp = phantom.init_mlp(C, H, seed=0).to(dev).requires_grad_()
    fp = N.FieldParams()
    fp.C, fp.G, fp.H = C, G, H
    params = [f.plane_yz, f.plane_xz, f.plane_xy, f.line_x, f.line_y, f.line_z, f.grid_out, mlp]
    moments = [(torch.zeros_like(p), torch.zeros_like(p)) for p in params]
    state = {"it": 0, "loss": None}

    def step():
        rays, tgt = batches[state["it"] % len(batches)]
        samples = tree.CreateSamplesForRays(rays, K, True)
        img = N.ComputeImageVM(tree, roi, f, mlp, fp, samples, rays.size())
        loss = (img[0] - tgt).square().mean()
        grads = torch.autograd.grad(loss, params)
        state["it"] += 1
        for p, g, (m, v) in zip(params, grads, moments):
            N.AdamStep(p.data, g.contiguous(), m, v, state["it"], 1e-3)
        state["loss"] = loss.detach()

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    first = float(state["loss"].item()) if state["loss"] is not None else None
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    R = batches[0][0].size()
    log(f"vm: {steps} steps in {el:.3f}s: {el * 1e3 / steps:.2f} ms/step, {R * steps / el / 1e6:.3f} M rays/s, "
        f"{n_in} VM / {n_out} explicit nodes")
    print(json.dumps({
        "metric": "fwd+bwd rays/s, tensorial (VM) field through the composed operators, 512^3, 61 tilts",
        "value": R * steps / el, "unit": "rays/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": el * 1e3 / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (analytic ellipsoid phantom projections, random-init field)",
        "config": {"workload": f"tensorial geometry training step, {n}^3 phantom, 61 tilts, {n}x{n} detector, "
                               f"adaptive octree depth<=5, ROI {lo}..{hi}: {n_in} plane/line nodes, {n_out} "
                               f"explicit-grid nodes, lines of {line}",
                   "rays_per_batch": R, "K": K, "stratified_samples_per_ray": round(s_eff, 2), "channels": C,
                   "grid": G, "decoder": f"{C}-{H}-{H}-1", "parallelism": "single GPU",
                   "loss_first_last": [first, float(state["loss"].item())]},
        "roofline": None, "cpu_baseline": None}), flush=True)


def run_leg(N, tree, batches, sp0, fp, grid0, mlp0, bf16, ps, args, steps):
    """Train one leg: warmup, `steps` timed steps, then (learned mode) more
    untimed steps up to --train-iters.  With a proposal state, the proposal is
    trained every 100 iterations (the first time at iteration 100, before
    which the leg samples stratified)."""
    import torch

    st = N.MakeTrainState(grid0, mlp0, bf16, 1e-3, 1e-3)
    prop_rays = {}
    stats = {"updates": 0, "update_ms": [], "loss": [], "fallback_frac": None}
    sp = N.SamplingParams()
    sp.max_samples_per_node = sp0.max_samples_per_node
    sp.jitter_mode = sp0.jitter_mode

    def sub(b):
        if b not in prop_rays:
            rays, _ = batches[b]
            R = rays.size()
            stride = max(1, R // args.proposal_rays)
            idx = torch.arange(0, R, stride, device=rays.origin.device)[: args.proposal_rays]
            prop_rays[b] = N.RayList(rays.origin.index_select(0, idx).contiguous(),
                                     rays.direction.index_select(0, idx).contiguous())
        return prop_rays[b]

    def step(i):
        b = i % len(batches)
        rays, tgt = batches[b]
        sp.jitter_seed = 7 + i
        if ps is None:
            return N.TrainStep(tree, st, rays, tgt, sp, fp)
        if i > 0 and i % 100 == 0:
            t0 = time.perf_counter()
            loss = N.ProposalTrain(tree, ps, sub(b), sp, st.grid, st.mlp, fp, 200 if stats["updates"] == 0 else 50)
            stats["update_ms"].append((time.perf_counter() - t0) * 1e3)
            stats["loss"].append(loss)
            stats["updates"] += 1
        return N.TrainStepLearned(tree, st, ps, rays, tgt, sp, fp)

    # learned legs start timing once the proposal is in use (iteration 100)
    start = 100 if ps is not None else 0
    it = 0
    while it < start + args.warmup:
        step(it)
        it += 1
    torch.cuda.synchronize()
    first_loss = float(st.loss.item())
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(steps):
        step(it)
        it += 1
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    last_loss = float(st.loss.item())
    ms = elapsed * 1e3 / steps
    rps = batches[0][0].size() * steps / elapsed
    # samples per ray of the leg's own sampler on batch 0
    rays0 = batches[0][0]
    if ps is not None:
        iv, fb = N.CreateProposalIntervals(tree, ps, rays0, sp)
        stats["fallback_frac"] = float(fb.float().mean().item())
    else:
        iv = tree.CreateIntervalsForRays(rays0, sp)
    spr = iv.num_samples / rays0.size()
    log(f"{'learned+bf16' if ps is not None else ('bf16' if bf16 else 'plain')}: {steps} steps in {elapsed:.3f}s: "
        f"{ms:.2f} ms/step, {rps / 1e6:.3f} M rays/s, {spr:.1f} samples/ray, loss {first_loss:.5f} -> {last_loss:.5f}"
        + (f", fallback {stats['fallback_frac']:.3f}, proposal updates {stats['updates']} "
           f"({', '.join(f'{x:.0f}' for x in stats['update_ms'])} ms)" if ps is not None else ""))
    if args.mode == "learned":
        while it < args.train_iters:
            step(it)
            it += 1
        torch.cuda.synchronize()
    leg = {"state": st, "rays_per_s": rps, "ms_per_step": ms, "loss_first_last": [first_loss, last_loss],
           "samples_per_ray": spr}
    if ps is not None:
        leg["proposal"] = {"mix": args.mix, "kl_threshold": args.kl_threshold, "bins": 32, "samples": 32,
                           "updates_in_timed_region": sum(1 for j in range(start + args.warmup,
                                                                           start + args.warmup + steps)
                                                          if j > 0 and j % 100 == 0),
                           "update_ms": stats["update_ms"], "proposal_loss": stats["loss"],
                           "fallback_frac": stats["fallback_frac"], "train_rays": args.proposal_rays}
    return leg


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
    C, G, H = fp.C, fp.G, fp.H
    S_total = iv.num_samples
    R = rays.size()
    f = capi.Field(C, H, G, capi.GRID_BF16 if bf16 else capi.GRID_F32, 2.0, 0, st.grid.data_ptr(),
                   st.grid.shape[0], tree.node_to_grid.data_ptr(), st.mlp.data_ptr(),
                   capi.DECODER_BF16 if bf16 else capi.DECODER_F32, 0)
    csp = capi.Sampling(sp.max_samples_per_node, 256, sp.jitter_mode, 0, sp.jitter_seed, None, None)
    civ = capi.Intervals(iv.node.data_ptr(), iv.tmin.data_ptr(), iv.tmax.data_ptr(), iv.count.data_ptr(),
                         iv.first.data_ptr(), iv.ray.data_ptr(), iv.num_intervals)
    packed = torch.empty(capi.lib().atdf_tree_packed_bytes(tree.NumNodes()), dtype=torch.uint8, device=dev)
    tdesc = capi.Tree(tree.NumNodes(), tree.node_position_min.data_ptr(), tree.node_position_max.data_ptr(),
                      tree.node_children.data_ptr(), tree.node_parent.data_ptr(), tree.node_active.data_ptr())
    s = capi.stream_ptr()
    capi.call("atdf_tree_pack", Ct.byref(tdesc), capi.ptr(packed), s)
    P = torch.empty(R, device=dev)
    dP = torch.randn(R, device=dev) * 1e-6
    gg = torch.zeros(st.grid.shape, dtype=torch.float32, device=dev)
    gm = torch.zeros_like(st.mlp)
    ws = torch.empty(capi.lib().atdf_project_bwd_workspace_bytes(R, iv.num_intervals, Ct.byref(f)),
                     dtype=torch.uint8, device=dev)

    args = (capi.ptr(packed), tree.NumNodes(), capi.ptr(rays.origin), capi.ptr(rays.direction), R, Ct.byref(csp),
            capi.ptr(iv.ray_offset), Ct.byref(civ), Ct.byref(f))
    node_fwd = fp.node_forward

    # the calls the training step makes: with the node-major forward, one
    # counting sort (prepare) serves the forward and the adjoint
    def prep():
        capi.call("atdf_project_prepare", *args, capi.ptr(ws), ws.numel(), s)

    def fwd():
        if node_fwd:
            capi.call("atdf_project_fwd_prepared", *args, capi.ptr(P), capi.ptr(ws), ws.numel(), s)
        else:
            capi.call("atdf_project_fwd", *args, capi.ptr(P), s)

    def bwd():
        if node_fwd:
            capi.call("atdf_project_bwd_prepared", *args, capi.ptr(dP), capi.ptr(gg), capi.ptr(gm), None,
                      capi.ptr(ws), ws.numel(), s)
        else:
            capi.call("atdf_project_bwd", *args, capi.ptr(dP), capi.ptr(gg), capi.ptr(gm), capi.ptr(ws), ws.numel(),
                      s)

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

    t_prep = timed(prep) if node_fwd else 0.0
    if node_fwd:
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
    log(f"  {'project_prepare %.2f ms, ' % (t_prep * 1e3) if node_fwd else ''}"
        f"project_fwd{'_prepared' if node_fwd else ''} {t_fwd * 1e3:.2f} ms, "
        f"project_bwd{'_prepared' if node_fwd else ''} {t_bwd * 1e3:.2f} ms ({S_total / R:.1f} samples/ray, "
        f"{iv.num_intervals} intervals); dominant {dom}: {achieved_gbs:.0f} GB/s algorithmic, "
        f"{achieved_tflops:.1f} TFLOP/s decoder")
    mode = ("learned" if ps is not None else "bf16") if bf16 else "plain"
    traffic = measured_traffic(mode, f"project_{dom}")
    return {
        "kernel_ms": ({"project_prepare": t_prep * 1e3, "project_fwd_prepared": t_fwd * 1e3,
                       "project_bwd_prepared": t_bwd * 1e3} if node_fwd else
                      {"project_fwd": t_fwd * 1e3, "project_bwd": t_bwd * 1e3}),
        "roofline": {
            "bound": "hbm", "kernel": f"project_{dom}", "achieved": achieved_gbs, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved_gbs / HBM_PEAK_GBS, "traffic": traffic,
            "algorithmic_bytes_per_launch": bytes_dom, "samples_per_launch": S_total,
            "mfma": {"achieved": achieved_tflops, "peak": mfma_peak, "unit": "TFLOP/s",
                     "frac": achieved_tflops / mfma_peak},
        },
    }


def measured_traffic(mode, kernel):
    """HBM bytes per launch of the main kernel behind `kernel` (project_fwd or
    project_bwd) for this mode's interval lists, from the committed rocprofv3
    PMC summary of the 512^3 workload (FETCH_SIZE + WRITE_SIZE,
    copy-calibrated), or None when that combination was not measured."""
    path = os.path.join(HERE, "profiles", "traffic_512cube.json")
    template = {"plain": "<16, false, false", "learned": "<16, true, true"}.get(mode)
    if template is None or not os.path.exists(path):
        return None
    with open(path) as fh:
        data = json.load(fh)
    stem = "project_bwd_node_kernel" if kernel == "project_bwd" else "project_fwd_kernel"
    for name, rec in data.get("kernels", {}).items():
        if stem in name and template in name:
            return rec["hbm_bytes_per_launch"]
    return None


def cpu_baseline(tree, active, grid0, mlp0, rays, tgt, K, C, G, seconds):
    """The oracle's FP32 training step (OpenMP over rays, dense Adam) on the
    same workload, timed on this host for about `seconds`."""
    import torch

    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    nr = 2048
    o = rays.origin[:nr].cpu().numpy()
    d = rays.direction[:nr].cpu().numpy()
    t = tgt[:nr].cpu().numpy()
    fm = O.FusedModel(tree.node_position_min.cpu().numpy(), tree.node_position_max.cpu().numpy(), active,
                      grid0.cpu().numpy(), mlp0.cpu().numpy(), C, K=K, jitter_mode=2, seed=7)
    tb = O.TrainBaseline(fm, lr=1e-3)
    tb.train_step(o, d, t, threads, do_adam=False)  # warm-up
    n_steps, t0 = 0, time.perf_counter()
    while True:
        tb.train_step(o, d, t, threads, do_adam=False)
        n_steps += 1
        el = time.perf_counter() - t0
        if el >= seconds or n_steps >= 200:
            break
    per_ray = el / (n_steps * nr)
    t1 = time.perf_counter()
    tb.train_step(o[:1], d[:1], t[:1], threads, do_adam=True)  # one dense Adam pass (+1 ray)
    t_adam = time.perf_counter() - t1
    R = rays.origin.shape[0]
    value = R / (per_ray * R + t_adam)
    return {"value": value, "unit": "rays/s", "cores": threads, "kind": "port",
            "sample": f"{n_steps} FP32 fwd+bwd passes over {nr} rays of the same workload ({el:.1f} s; oracle "
                      f"restatement, OpenMP over rays) plus one dense Adam pass over all "
                      f"{grid0.numel() + mlp0.numel()} parameters ({t_adam:.2f} s), scaled to one "
                      f"{R}-ray step"}

if __name__ == "__main__":
    main()
