#!/usr/bin/env python3
"""Benchmark: fwd+bwd rays/s of one training step of the reconstruction hot
path on one MI355X.

Workload (default, the 512^3 configuration): 200-ellipsoid phantom (seed 1),
61 tilts from -60 to +60 degrees, 512 x 512 detector, adaptive octree up to
depth 5, 11^3 x 16-channel node grids, batches of 2^18 rays, stratified
samples with K chosen so the mean is ~128 per ray.  One step = traversal ->
fused forward -> MSE -> fused adjoint -> Adam on every grid and decoder
parameter, all on the device.  Inputs (rays, targets) are resident in HBM
before the timed region.

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
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--mode", choices=["plain", "bf16"], default="plain")
    ap.add_argument("--volume", type=int, default=512)
    ap.add_argument("--rays", type=int, default=1 << 18)
    ap.add_argument("--batches", type=int, default=4)
    ap.add_argument("--budget", type=int, default=8192, help="max active octree nodes")
    ap.add_argument("--samples-per-ray", type=float, default=128.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
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

    from adaptivetensordf_amd import capi, native, phantom

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
    batches = []
    for b in range(args.batches):
        tl, px, py = phantom.sample_pixels(args.rays, 61, n, n, 100 + b)
        o, d = phantom.rays_for_pixels(tl, px, py, n, n, angles)
        o_t, d_t = torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev)
        tgt = phantom.analytic_projection_torch(o_t, d_t, ell).float()
        batches.append((N.RayList(o_t, d_t), tgt.contiguous()))
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
    fp = N.FieldParams()
    fp.C, fp.G, fp.H = C, G, H
    fp.bf16_decoder = args.mode == "bf16"
    grid0 = phantom.init_grid(num_active, G, C, std=0.01, seed=0, device=dev)
    mlp0 = phantom.init_mlp(C, H, seed=0).to(dev)
    st = N.MakeTrainState(grid0, mlp0, args.mode == "bf16", 1e-3, 1e-3)
    log(f"setup {time.time() - t0:.1f}s: active nodes {num_active}, K {K}, mean samples/ray {s_eff:.1f}")

    # ------------------------------------------------------------ timed steps
    def step(i):
        rays, tgt = batches[i % len(batches)]
        return N.TrainStep(tree, st, rays, tgt, sp, fp)

    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()
    steps = args.profile_steps if args.profile_steps > 0 else args.steps
    first_loss = float(st.loss.item())
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(steps):
        step(args.warmup + i)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    last_loss = float(st.loss.item())
    ms_per_step = elapsed * 1e3 / steps
    rays_per_s = args.rays * steps / elapsed
    log(f"{steps} steps in {elapsed:.3f}s: {ms_per_step:.2f} ms/step, {rays_per_s / 1e6:.3f} M rays/s, "
        f"loss {first_loss:.5f} -> {last_loss:.5f}")
    if args.profile_steps > 0:
        return

    # ------------------------------------------------- per-kernel durations
    rays, tgt = batches[0]
    iv = tree.CreateIntervalsForRays(rays, sp)
    S_total = iv.num_samples
    import ctypes as Ct

    grid_dtype = capi.GRID_BF16 if args.mode == "bf16" else capi.GRID_F32
    f = capi.Field(C, H, G, grid_dtype, 2.0, 0, st.grid.data_ptr(), st.grid.shape[0], tree.node_to_grid.data_ptr(),
                   st.mlp.data_ptr(), capi.DECODER_BF16 if args.mode == "bf16" else capi.DECODER_F32, 0)
    csp = capi.Sampling(K, 256, 2, 0, 7, None)
    civ = capi.Intervals(iv.node.data_ptr(), iv.tmin.data_ptr(), iv.tmax.data_ptr(), iv.count.data_ptr(),
                         iv.first.data_ptr(), iv.ray.data_ptr(), iv.num_intervals)
    packed = torch.empty(capi.lib().atdf_tree_packed_bytes(tree.NumNodes()), dtype=torch.uint8, device=dev)
    tdesc = capi.Tree(tree.NumNodes(), tree.node_position_min.data_ptr(), tree.node_position_max.data_ptr(),
                      tree.node_children.data_ptr(), tree.node_parent.data_ptr(), tree.node_active.data_ptr())
    s = capi.stream_ptr()
    capi.call("atdf_tree_pack", Ct.byref(tdesc), capi.ptr(packed), s)
    R = rays.origin.shape[0]
    P = torch.empty(R, device=dev)
    dP = torch.randn(R, device=dev) * 1e-6
    gg = torch.zeros(st.grid.shape, dtype=torch.float32, device=dev)
    gm = torch.zeros_like(st.mlp)
    ws = torch.empty(capi.lib().atdf_project_bwd_workspace_bytes(R, iv.num_intervals, Ct.byref(f)), dtype=torch.uint8, device=dev)

    def fwd():
        capi.call("atdf_project_fwd", capi.ptr(packed), tree.NumNodes(), capi.ptr(rays.origin),
                  capi.ptr(rays.direction), R, Ct.byref(csp), capi.ptr(iv.ray_offset), Ct.byref(civ), Ct.byref(f),
                  capi.ptr(P), s)

    def bwd():
        capi.call("atdf_project_bwd", capi.ptr(packed), tree.NumNodes(), capi.ptr(rays.origin),
                  capi.ptr(rays.direction), R, Ct.byref(csp), capi.ptr(iv.ray_offset), Ct.byref(civ), Ct.byref(f),
                  capi.ptr(dP), capi.ptr(gg), capi.ptr(gm), capi.ptr(ws), ws.numel(), s)

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
    b = 2 if args.mode == "bf16" else 4
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
    mfma_peak = BF16_PEAK_TFLOPS if args.mode == "bf16" else FP32_PEAK_TFLOPS
    log(f"project_fwd {t_fwd * 1e3:.2f} ms, project_bwd {t_bwd * 1e3:.2f} ms; dominant {dom}: "
        f"{achieved_gbs:.0f} GB/s algorithmic, {achieved_tflops:.1f} TFLOP/s decoder")
    traffic = measured_traffic(args.mode, f"project_{dom}")
    # whole-step model of BASELINE.md: bytes/ray = 60 + S (16 C b + 32 C)
    step_bytes_per_ray = 60 + s_eff * (16 * C * b + 32 * C)

    # ------------------------------------------------------------ CPU baseline
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(tree, active, st, grid0, mlp0, rays, tgt, K, C, G, args.cpu_seconds)

    out = {
        "metric": METRIC,
        "value": rays_per_s,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.mode == "plain" else "bf16 grids / f32 accumulate",
        "data": "synthetic (analytic ellipsoid phantom projections, random-init field)",
        "config": {
            "workload": f"{'plain FP32' if args.mode == 'plain' else 'BF16 grids'} training step, {n}^3 phantom "
                        f"(200 ellipsoids), 61 tilts, {n}x{n} detector, adaptive octree depth<=5",
            "rays_per_batch": args.rays, "samples_per_ray": round(s_eff, 2), "K": K,
            "active_nodes": num_active, "channels": C, "grid": G, "decoder": f"{C}-{H}-{H}-1",
            "parallelism": "single GPU",
            "step": "traverse + fused fwd + MSE + fused adjoint + Adam(grids, decoder)",
            "loss_first_last": [first_loss, last_loss],
            "step_model_bytes_per_ray": step_bytes_per_ray,
            "step_model_hbm_frac": step_bytes_per_ray * rays_per_s / (HBM_PEAK_GBS * 1e9),
            "kernel_ms": {"project_fwd": t_fwd * 1e3, "project_bwd": t_bwd * 1e3},
        },
        "roofline": {
            "bound": "hbm", "kernel": f"project_{dom}", "achieved": achieved_gbs, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved_gbs / HBM_PEAK_GBS, "traffic": traffic,
            "mfma": {"achieved": achieved_tflops, "peak": mfma_peak, "unit": "TFLOP/s",
                     "frac": achieved_tflops / mfma_peak},
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def measured_traffic(mode, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary of this workload (FETCH_SIZE / WRITE_SIZE, copy-calibrated), or
    None when no summary for this mode is committed."""
    path = os.path.join(HERE, "profiles", f"traffic_{'plain_fp32' if mode == 'plain' else 'bf16'}_512cube.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        data = json.load(fh)
    for name, rec in data.get("kernels", {}).items():
        if kernel + "_node_kernel" in name or (kernel + "_kernel") in name:
            return rec["hbm_bytes_per_launch"]
    return None


def cpu_baseline(tree, active, st, grid0, mlp0, rays, tgt, K, C, G, seconds):
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
