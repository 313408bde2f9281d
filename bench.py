#!/usr/bin/env python3
"""Tracking-iteration throughput of the GS-EVT rasterizer path on one MI355X.

A step is one fine-stage tracking iteration at full resolution: render of the
last and next camera of the window, grey intensity difference, L2
normalisation, loss, adjoint through the rasterizer, Adam and the
pose/velocity update.

Workload (default): 1M anisotropic Gaussians on 64 planar patches, SH degree
3, 640x480, fx = fy = 550, one 10 ms event window sampled from the rendered
intensity change under a ground-truth twist (0.3 m/s, 20 deg/s); initial pose
perturbed by 2 cm / 1 deg.  All data synthetic and seeded.

Prints one JSON line (see --help for the knobs).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "gs-evt_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="patches_1m")
    ap.add_argument("--mode", default="plain", choices=["plain"])
    ap.add_argument("--cpu-baseline", type=int, default=1, help="time the CPU oracle on a bounded sample")
    ap.add_argument("--cpu-sample-iters", type=int, default=1)
    return ap.parse_args()


def blend_bytes(meta, nv, backward):
    """Algorithmic HBM bytes of one blend launch: 44 B per fragment (id +
    40 B of splat attributes) plus per-pixel image terms; the backward adds
    the per-visible-Gaussian gradient traffic (40 B)."""
    n, W, H = meta["n"], meta["W"], meta["H"]
    if backward:
        return 24 * W * H + 44 * n + 40 * nv
    return 44 * n + 28 * W * H


def run_plain(prob, steps, warmup):
    import torch
    from gsevt import raster
    from gsevt.tracker import plain_fine_iterations
    lr = (0.004, 0.004, 0.002, 0.002)
    cam = prob.camera()
    plain_fine_iterations(cam, prob.gaussians, prob.background, prob.sign_Ie, warmup, lr)
    torch.cuda.synchronize()
    raster.enable_kernel_timers(["blend_forward", "blend_backward"])
    losses = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain_fine_iterations(cam, prob.gaussians, prob.background, prob.sign_Ie, steps, lr, losses=losses)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    timers = raster.disable_kernel_timers()
    return dt, losses, timers, cam


def kernel_stats(timers):
    import torch
    out = {}
    for name, lst in timers.items():
        if not lst:
            continue
        ms = [s.elapsed_time(e) for s, e, _ in lst]
        nv = int((lst[0][2]["radii"] > 0).sum().item())
        nbytes = [blend_bytes(m, nv, name == "blend_backward") for _, _, m in lst]
        avg_ms = sum(ms) / len(ms)
        avg_bytes = sum(nbytes) / len(nbytes)
        out[name] = dict(avg_ms=avg_ms, launches=len(ms), avg_bytes=avg_bytes, n_fragments=lst[0][2]["n"],
                         n_visible=nv, achieved_gbs=avg_bytes / (avg_ms * 1e-3) / 1e9)
    return out


def cpu_baseline(prob, iters):
    """Time the CPU oracle's tracking iteration on the same workload."""
    import numpy as np
    import torch
    from oracle import oracle as orc
    from oracle import tracking_step
    from gsevt.camera import render_matrices
    cam = prob.camera()
    g = prob.gaussians
    scene = orc.Scene(means=g.get_xyz.cpu().numpy(), scales=g.get_scaling.cpu().numpy(),
                      rots=g.get_rotation.cpu().numpy(), opac=g.get_opacity.reshape(-1).cpu().numpy(),
                      shs=g.get_features.cpu().numpy(), sh_degree=g.active_sh_degree)

    def ocam(pose):
        v, p, pr, c = render_matrices(pose[:3, :3], pose[:3, 3], cam.FoVx, cam.FoVy)
        import math
        return orc.Cam(view=v.reshape(-1).cpu().numpy(), proj=p.reshape(-1).cpu().numpy(),
                       projraw=pr.reshape(-1).cpu().numpy(), campos=c.cpu().numpy(), W=cam.image_width,
                       H=cam.image_height, tanfovx=math.tan(cam.FoVx * 0.5), tanfovy=math.tan(cam.FoVy * 0.5),
                       bg=np.zeros(3, np.float32))
    last = (cam.last_vel_transform @ cam.curr_pose)
    nxt = (cam.next_vel_transform @ cam.curr_pose)
    cl, cn = ocam(last), ocam(nxt)
    e = prob.sign_Ie[0].cpu().numpy()
    t0 = time.perf_counter()
    for _ in range(iters):
        tracking_step.iteration(scene, cl, cn, e)
    dt = time.perf_counter() - t0
    return dict(value=iters / dt, unit="iters/s", cores=os.cpu_count(), kind="port",
                sample=f"{iters} full tracking iteration(s) of the same 1M-Gaussian 640x480 workload "
                       f"(C++ oracle, OpenMP {os.cpu_count()} threads for preprocess/blend-forward, "
                       f"single-threaded backward)")


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the tracking path has no "
              "multi-GPU component in this build", file=sys.stderr)
        sys.exit(2)
    import torch
    from gsevt import workloads
    torch.cuda.init()
    print(f"[bench] building workload {args.workload}", file=sys.stderr, flush=True)
    prob = workloads.build(args.workload)
    torch.cuda.synchronize()
    print(f"[bench] plain path: warmup {args.warmup}, steps {args.steps}", file=sys.stderr, flush=True)
    dt, losses, timers, cam = run_plain(prob, args.steps, args.warmup)
    ks = kernel_stats(timers)
    dom_name = max(ks, key=lambda k: ks[k]["avg_ms"] * ks[k]["launches"])
    dom = ks[dom_name]
    value = args.steps / dt
    cpu = None
    if args.cpu_baseline:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(prob, args.cpu_sample_iters)
    line = {
        "metric": "pose-tracking iters/sec, batched+fused vs plain, 640x480 @ 1M Gaussians",
        "value": value,
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * dt / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: 1M Gaussians SH3, 640x480, plain single-hypothesis path "
                               "(reference loop with the drop-in rasterizer), fine stage, level 0",
                   "path": args.mode, "gaussians": prob.gaussians.num_points,
                   "events": prob.events.size()},
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": dom["achieved_gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": dom["achieved_gbs"] / HBM_PEAK_GBS, "traffic": None,
                     "avg_ms": dom["avg_ms"], "algorithmic_bytes": dom["avg_bytes"],
                     "n_fragments": dom["n_fragments"], "n_visible": dom["n_visible"]},
        "kernels": {k: {"avg_ms": v["avg_ms"], "launches": v["launches"]} for k, v in ks.items()},
        "final_loss": float(losses[-1]) if losses else None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
