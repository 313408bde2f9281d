#!/usr/bin/env python3
"""Pose-tracking iterations/s of the GS-EVT rasterizer path on one MI355X,
batched + fused vs plain, 640x480 at 1M Gaussians.

A step (tracking iteration) is one Adam iteration of one (pose hypothesis,
event window) item at full resolution: render of the window's last and next
camera, grey intensity difference, L2 normalisation, event loss, adjoint
through the rasterizer, Adam and the pose/velocity update.

Workload "batched_64x4": 1M anisotropic Gaussians on 64 planar patches (SH
degree 3), 640x480, fx = fy = 550; 4 consecutive 10 ms windows of a
constant-twist motion (0.3 m/s, 20 deg/s) with Poisson(150k) events each
(clipped to 20k..500k) sampled from the rendered intensity change; 64 pose
hypotheses per window (5 cm / 3 deg).  All data synthetic and seeded.

* plain   : the reference's tracking loop (utils/tracker.py:176-240) with the
            drop-in diff_gaussian_rasterization, one item at a time
* batched : all 256 items per iteration through the fused kernels, one HIP
            graph per iteration (value)
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "gs-evt_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0
METRIC = "pose-tracking iters/sec, batched+fused vs plain, 640x480 @ 1M Gaussians"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "tr_blend_backward_traffic.md")


def measured_traffic(kernel, items, gaussians, tile):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated), when that summary was
    taken on a launch of the same shape; None otherwise."""
    import re
    try:
        text = open(TRAFFIC_FILE).read()
    except OSError:
        return None
    shape = re.search(r"^launch: kernel=(\S+) items=(\d+) gaussians=(\d+) tile=(\S+)", text, re.M)
    val = re.search(r"^traffic_bytes_per_launch: (\d+)", text, re.M)
    if not shape or not val:
        return None
    if (shape.group(1), int(shape.group(2)), int(shape.group(3)), shape.group(4)) != (kernel, items, gaussians, tile):
        return None
    return int(val.group(1))


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6, help="timed batched iterations (each = all items)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mode", default="all", choices=["all", "plain", "batched", "sequence"],
                    help="sequence: consecutive windows along a smooth trajectory (own JSON line)")
    ap.add_argument("--seq-windows", type=int, default=100, help="sequence mode: windows (2000 = full sequence)")
    ap.add_argument("--seq-iters", type=int, default=50, help="sequence mode: iterations per window")
    ap.add_argument("--seq-hyp", type=int, default=16, help="sequence mode: re-initialisation hypotheses")
    ap.add_argument("--seq-velocity", default="finite_difference",
                    choices=["reference", "optimised", "finite_difference"],
                    help="sequence mode: velocity update after each window (see gsevt/sequence.py)")
    ap.add_argument("--plain-steps", type=int, default=10)
    ap.add_argument("--hyp", type=int, default=64)
    ap.add_argument("--windows", type=int, default=4)
    ap.add_argument("--chunk", type=int, default=0, help="items per launch (0: all items in one launch)")
    ap.add_argument("--gaussians", type=int, default=1000000)
    ap.add_argument("--tile", default="16x16", help="plain path tile; batched start shape")
    ap.add_argument("--autotune", type=int, default=1, help="batched path: pick the tile shape by timing")
    ap.add_argument("--cpu-baseline", type=int, default=1)
    ap.add_argument("--cpu-sample-iters", type=int, default=1)
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="run a device-to-device copy of this size first (reference for HBM counters)")
    return ap.parse_args()


def blend_bytes(n_frag, n_vis, HW, backward):
    """Algorithmic HBM bytes of one render's blend: 44 B per fragment (id +
    40 B of splat attributes), 28 B per pixel of outputs (forward) or 24 B of
    image inputs plus 40 B per visible Gaussian of gradient traffic
    (backward)."""
    if backward:
        return 24 * HW + 44 * n_frag + 40 * n_vis
    return 44 * n_frag + 28 * HW


def run_plain(prob, steps, warmup, tile):
    import torch
    from gsevt import raster
    from gsevt.camera import Camera
    from gsevt.tracker import plain_fine_iterations
    lr = (0.004, 0.004, 0.002, 0.002)
    T0 = prob.init_poses[0]
    intr = prob.intr

    def cam():
        c = Camera(T0[:3, :3].clone(), T0[:3, 3].clone(), prob.gt_ang.clone(), prob.gt_lin.clone(), intr.fovx,
                   intr.fovy, intr.W, intr.H, delta_tau=prob.dtau[0], device=T0.device)
        c.fx, c.fy = intr.fx, intr.fy
        return c
    bg = torch.zeros(3, device=T0.device)
    e = prob.events[0:1]
    c = cam()
    plain_fine_iterations(c, prob.gaussians, bg, e, warmup, lr, tile=tile)
    torch.cuda.synchronize()
    raster.enable_kernel_timers(["blend_forward", "blend_backward"])
    losses = []
    t0 = time.perf_counter()
    plain_fine_iterations(c, prob.gaussians, bg, e, steps, lr, tile=tile, losses=losses)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    timers = raster.disable_kernel_timers()
    HW = intr.W * intr.H
    ks = {}
    for name, lst in timers.items():
        ms = [s.elapsed_time(en) for s, en, _ in lst]
        nv = int((lst[0][2]["radii"] > 0).sum().item())
        nbytes = [blend_bytes(m["n"], nv, HW, name == "blend_backward") for _, _, m in lst]
        ks[name] = dict(avg_ms=sum(ms) / len(ms), launches=len(ms), avg_bytes=sum(nbytes) / len(nbytes))
    return dict(iters_per_s=steps / dt, ms_per_iter=1000 * dt / steps, final_loss=losses[-1], kernels=ks)


def run_batched(prob, steps, warmup, chunk, tile, autotune=True):
    import torch
    from gsevt import raster
    from gsevt.fused import BatchedTracker
    intr = prob.intr
    B = len(prob.item_window)
    tr = BatchedTracker(prob.gaussians, intr.W, intr.H, intr.fx, intr.fy, prob.events, B, tile=tile,
                        chunk=chunk or None, device=prob.events.device)
    lin = prob.gt_lin.expand(B, 3)
    ang = prob.gt_ang.expand(B, 3)
    dtau = [prob.dtau[w] for w in prob.item_window]

    def reset():
        tr.set_items(prob.init_poses, lin, ang, dtau, prob.item_window)

    reset()
    nmax = tr.measure_capacity()
    tune = None
    if autotune:
        best, tune = tr.autotune()
        nmax = int(tr.nfrag.max().item())
        log("autotune (ms per iteration): " + ", ".join(f"{a}x{b}: {t:.1f}" for (a, b), t in tune.items())
            + f" -> {best[0]}x{best[1]}")
    log(f"batched: {B} items, chunk {tr.chunk}, tile {tr.tile[0]}x{tr.tile[1]}, fragment capacity {tr.cap} "
        f"(max measured {nmax})")
    # eager pass with HIP events around the blend launches (per-kernel roofline)
    reset()
    raster.enable_kernel_timers(["tr_blend_forward", "tr_blend_backward"])
    _instrument(tr)
    for _ in range(max(1, warmup)):
        tr.step()
    torch.cuda.synchronize()
    timers = raster.disable_kernel_timers()
    _uninstrument(tr)
    HW = intr.W * intr.H
    ks = {}
    for name, lst in timers.items():
        ms = [s.elapsed_time(en) for s, en, _ in lst]
        nbytes = []
        for _, _, m in lst:
            nf = m["nfrag"].long()
            nv = m["nvis"]
            nbytes.append(sum(blend_bytes(int(a), int(b), HW, name == "tr_blend_backward")
                              for a, b in zip(nf.tolist(), nv.tolist())))
        ks[name] = dict(avg_ms=sum(ms) / len(ms), launches=len(ms), avg_bytes=sum(nbytes) / len(nbytes),
                        renders_per_launch=lst[0][2]["R"])
    # graph-captured timed region
    reset()
    tr.capture()
    reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        tr.graph.replay()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    tr.check_overflow()
    # accuracy of the tracked items against each window's ground truth
    T = tr.poses()
    gt = prob.window_gt.double()[torch.tensor(prob.item_window, device=T.device)]
    terr = (T[:, :3, 3] - gt[:, :3, 3]).norm(dim=1)
    dR = T[:, :3, :3] @ gt[:, :3, :3].transpose(1, 2)
    cosang = ((dR.diagonal(dim1=1, dim2=2).sum(1) - 1) / 2).clamp(-1, 1)
    rerr = torch.acos(cosang)
    T0 = prob.init_poses.double()
    t0err = (T0[:, :3, 3] - gt[:, :3, 3]).norm(dim=1)
    loss = tr.loss.cpu()
    R = tr.R
    frags = dict(rect_slots=float(tr.nfrag[:R].double().mean()),
                 walked=float((tr.nkept if (tr.packed and tr.cull_tiles) else tr.nfrag)[:R].double().mean()))
    return dict(items=B, fragments_per_render=frags, tile_cull=bool(tr.packed and tr.cull_tiles), iters_per_s=B * steps / dt, ms_per_batch_iter=1000 * dt / steps, kernels=ks,
                mean_loss=float(loss.mean()), trans_err_mean_m=float(terr.mean()), rot_err_mean_rad=float(rerr.mean()),
                init_trans_err_mean_m=float(t0err.mean()), cap=tr.cap, chunk=tr.chunk, tracker=tr,
                tile=f"{tr.tile[0]}x{tr.tile[1]}",
                autotune=None if tune is None else {f"{a}x{b}": t for (a, b), t in tune.items()})


def _instrument(tr):
    """Wrap the batched blend launches with HIP events (eager pass only)."""
    from gsevt import fused, raster
    orig = fused.call

    def wrapped(name, *args):
        if name in ("gsevt_tr_blend_forward", "gsevt_tr_blend_backward"):
            R = args[0]
            key = "tr_blend_forward" if name.endswith("forward") else "tr_blend_backward"
            nvis = (tr.tiles[:R * tr.P].view(R, tr.P) > 0).sum(1)
            # the fragments the blend walks: the tile-culled lists when the
            # tracker culls, every rect slot otherwise
            walked = tr.nkept if (tr.packed and tr.cull_tiles) else tr.nfrag
            with raster.timed(key, R=R, nfrag=walked[:R].clone(), nvis=nvis):
                return orig(name, *args)
        return orig(name, *args)
    fused.call = wrapped
    tr._orig_call = orig


def _uninstrument(tr):
    from gsevt import fused
    fused.call = tr._orig_call


def cpu_baseline(prob, iters):
    import numpy as np
    from oracle import oracle as orc
    from oracle import tracking_step
    from gsevt.camera import Camera, render_matrices
    g = prob.gaussians
    scene = orc.Scene(means=g.get_xyz.cpu().numpy(), scales=g.get_scaling.cpu().numpy(),
                      rots=g.get_rotation.cpu().numpy(), opac=g.get_opacity.reshape(-1).cpu().numpy(),
                      shs=g.get_features.cpu().numpy(), sh_degree=g.active_sh_degree)
    intr = prob.intr
    T0 = prob.init_poses[0]
    c = Camera(T0[:3, :3].clone(), T0[:3, 3].clone(), prob.gt_ang.clone(), prob.gt_lin.clone(), intr.fovx,
               intr.fovy, intr.W, intr.H, delta_tau=prob.dtau[0], device=T0.device)

    def ocam(pose):
        v, p, pr, cp = render_matrices(pose[:3, :3], pose[:3, 3], intr.fovx, intr.fovy)
        return orc.Cam(view=v.reshape(-1).cpu().numpy(), proj=p.reshape(-1).cpu().numpy(),
                       projraw=pr.reshape(-1).cpu().numpy(), campos=cp.cpu().numpy(), W=intr.W, H=intr.H,
                       tanfovx=math.tan(intr.fovx * 0.5), tanfovy=math.tan(intr.fovy * 0.5),
                       bg=np.zeros(3, np.float32))
    cl = ocam(c.last_vel_transform @ c.curr_pose)
    cn = ocam(c.next_vel_transform @ c.curr_pose)
    e = prob.events[0].cpu().numpy()
    t0 = time.perf_counter()
    for _ in range(iters):
        tracking_step.iteration(scene, cl, cn, e)
    dt = time.perf_counter() - t0
    # OpenMP uses OMP_NUM_THREADS threads when it is set, else one per CPU
    n = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count())
    return dict(value=iters / dt, unit="iters/s", cores=n, kind="port",
                sample=f"{iters} tracking iteration(s) of item 0 of the same 1M-Gaussian 640x480 workload on the "
                       f"C++ oracle (OpenMP {n} threads in preprocess and blend forward, backward single-threaded)")


def run_sequence(args):
    """Consecutive 10 ms windows along a seeded B-spline trajectory (peak
    1.5 m/s, 90 deg/s; 10k-500k events per window in proportion to speed),
    1M Gaussians, 640x480: each window tracked for seq_iters iterations as one
    graph, 16-hypothesis re-initialisation on a loss jump.  Reports
    iterations/s beside the drift against the ground truth."""
    import torch
    from gsevt import sequence, synthetic
    from gsevt.model import GaussianModel
    intr = synthetic.PATCH_CAM
    raw = synthetic.patch_scene(P=args.gaussians, sh_degree=3)
    gm = GaussianModel.from_raw(raw.xyz, raw.features_dc, raw.features_rest, raw.scaling, raw.rotation,
                                raw.opacity, raw.sh_degree, device="cuda")
    log(f"sequence: generating {args.seq_windows} windows")
    seq = sequence.make_sequence(gm, intr, args.seq_windows, log=log)
    torch.cuda.synchronize()
    log(f"sequence: tracking ({args.seq_iters} iterations per window, {args.seq_hyp} hypotheses on re-init)")
    res = sequence.track_sequence(gm, intr, seq, iters=args.seq_iters, n_hyp=args.seq_hyp, log=log,
                                  velocity_update=args.seq_velocity)
    e, a = sequence.drift(res.est, seq.gt_mid)
    line = {
        "metric": "sequence tracking iters/sec, 640x480 @ 1M Gaussians",
        "value": res.iters_per_s(), "unit": "iters/s", "n_gpus": 1, "higher_is_better": True,
        "windows": res.windows, "iterations_per_window": args.seq_iters, "iterations": res.iterations,
        "track_seconds": res.track_seconds, "reinit_windows": len(res.reinit),
        "drift": {"ate_rmse_m": float(e.pow(2).mean().sqrt()), "final_m": float(e[-1]), "max_m": float(e.max()),
                  "final_rot_deg": math.degrees(float(a[-1])), "max_rot_deg": math.degrees(float(a.max()))},
        "events_per_window": {"min": min(seq.counts), "max": max(seq.counts),
                              "mean": sum(seq.counts) / len(seq.counts)},
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"sequence_{res.windows}: {args.gaussians} Gaussians SH3, 640x480, B-spline "
                               f"trajectory (peak 1.5 m/s, 90 deg/s), 10 ms windows",
                   "path": "batched+fused, one graph per window", "tile": "16x16",
                   "velocity_update": args.seq_velocity},
    }
    print(json.dumps(line), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU (got --gpus {args.gpus}); this build has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)
    if args.mode == "sequence":
        run_sequence(args)
        return
    import torch
    from gsevt import workloads
    tile = tuple(int(x) for x in args.tile.split("x"))
    if args.calibrate_copy_mib:
        n = args.calibrate_copy_mib * (1 << 20) // 4
        a = torch.ones(n, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.calibrate_copy_mib} MiB read + {args.calibrate_copy_mib} MiB written")
        del a, b
    log(f"building workload: {args.gaussians} Gaussians, {args.hyp} hypotheses x {args.windows} windows")
    prob = workloads.build_batched(P=args.gaussians, n_hyp=args.hyp, n_windows=args.windows)
    torch.cuda.synchronize()
    plain = batched = None
    if args.mode in ("all", "plain"):
        log(f"plain path: warmup {args.warmup}, steps {args.plain_steps}")
        plain = run_plain(prob, args.plain_steps, args.warmup, tile)
        log(f"plain: {plain['iters_per_s']:.1f} iters/s")
    if args.mode in ("all", "batched"):
        log(f"batched path: warmup {args.warmup}, steps {args.steps}")
        batched = run_batched(prob, args.steps, args.warmup, args.chunk, tile, bool(args.autotune))
        log(f"batched: {batched['iters_per_s']:.1f} iters/s")
    cpu = None
    if args.cpu_baseline:
        log("cpu baseline")
        cpu = cpu_baseline(prob, args.cpu_sample_iters)
    if batched is not None:
        ks = batched["kernels"]
        value = batched["iters_per_s"]
        ms_per_step = 1000.0 / value
        steps = args.steps * batched["items"]
    else:
        ks = plain["kernels"]
        value = plain["iters_per_s"]
        ms_per_step = plain["ms_per_iter"]
        steps = args.plain_steps
    dom = max(ks, key=lambda k: ks[k]["avg_ms"] * ks[k]["launches"])
    d = ks[dom]
    achieved = d["avg_bytes"] / (d["avg_ms"] * 1e-3) / 1e9
    used_tile = batched["tile"] if batched else args.tile
    traffic = measured_traffic(dom, batched["chunk"] if batched else 1, args.gaussians, used_tile)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"batched_{args.hyp}x{args.windows}: {args.gaussians} Gaussians SH3, 640x480, "
                               f"{args.hyp} hypotheses x {args.windows} event windows "
                               f"(events {prob.event_counts}), fine stage, level 0",
                   "path": "batched+fused, hipGraph per iteration" if batched else "plain",
                   "tile": used_tile, "items_per_launch": batched["chunk"] if batched else 1,
                   "autotune_ms_per_iteration": batched["autotune"] if batched else None,
                   "tile_cull": batched["tile_cull"] if batched else False,
                   "fragments_per_render": batched["fragments_per_render"] if batched else None},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "avg_ms": d["avg_ms"],
                     "algorithmic_bytes": d["avg_bytes"]},
        "plain": None if plain is None else {"iters_per_s": plain["iters_per_s"], "ms_per_iter": plain["ms_per_iter"],
                                             "kernels": plain["kernels"]},
        "batched_over_plain": (value / plain["iters_per_s"]) if (plain and batched) else None,
        "accuracy": None if batched is None else {
            "mean_loss": batched["mean_loss"], "trans_err_mean_m": batched["trans_err_mean_m"],
            "rot_err_mean_rad": batched["rot_err_mean_rad"], "init_trans_err_mean_m": batched["init_trans_err_mean_m"],
            "iterations": args.steps + 0},
        "kernels": {k: {"avg_ms": v["avg_ms"], "launches": v["launches"]} for k, v in ks.items()},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
