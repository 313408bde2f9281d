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
COUNTERS_FILE = os.path.join(ROOT, "profiles", "tr_blend_counters.md")
# MI355X's published FP32 vector peak, 157.3 TFLOP/s, as wave64 instructions of
# one fused multiply-add per lane: 1.23e12 per second (0.5 per SIMD and cycle
# at 2.4 GHz).  A stream of independent v_fma_f32 issues at half that rate
# (measured), and the blends issue above that half rate, so only the published
# rate bounds them.
VALU_PEAK = 157.3e12 / (2 * 64)


def _summary_values(path, kernel, items, gaussians, tile, schedule):
    """key: value lines of the block of a committed rocprofv3 summary under
    profiles/ whose 'launch:' line names the same launch shape (a block runs
    to the next 'launch:' line); None if no block does."""
    import re
    try:
        text = open(path).read()
    except OSError:
        return None
    pat = re.compile(r"^launch: kernel=(\S+) items=(\d+) gaussians=(\d+) tile=(\S+) schedule=(\S+)", re.M)
    heads = list(pat.finditer(text))
    for k, m in enumerate(heads):
        if (m.group(1), int(m.group(2)), int(m.group(3)), m.group(4), m.group(5)) != (
                kernel, items, gaussians, tile, schedule):
            continue
        block = text[m.end():heads[k + 1].start() if k + 1 < len(heads) else len(text)]
        return {v.group(1): float(v.group(2)) for v in re.finditer(r"^([a-z_]+): ([0-9.eE+-]+)$", block, re.M)}
    return None


def measured_traffic(kernel, items, gaussians, tile, schedule):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (read requests x their size + WRITE_SIZE), when that summary was
    taken on a launch of the same shape; None otherwise."""
    v = _summary_values(TRAFFIC_FILE, kernel, items, gaussians, tile, schedule)
    return None if v is None or "traffic_bytes_per_launch" not in v else int(v["traffic_bytes_per_launch"])


def measured_valu(kernel, items, gaussians, tile, schedule, avg_ms):
    """VALU issue of `kernel` from the committed counter summary of the same
    launch shape: wave64 VALU instructions per launch (SQ_INSTS_VALU) over the
    launch time measured here, against VALU_PEAK; with the summary's lane
    utilisation (SQ_THREAD_CYCLES_VALU / 64 SQ_ACTIVE_INST_VALU), LDS bank
    conflict ratio and occupancy.  None without such a summary."""
    v = _summary_values(COUNTERS_FILE, kernel, items, gaussians, tile, schedule)
    if v is None or "valu_insts_per_launch" not in v:
        return None
    achieved = v["valu_insts_per_launch"] / (avg_ms * 1e-3)
    return {"bound": "valu", "achieved": achieved, "peak": VALU_PEAK, "unit": "wave64 instructions/s",
            "frac": achieved / VALU_PEAK, "lane_utilisation": v.get("valu_lane_utilisation"),
            "lds_bank_conflict_ratio": v.get("lds_bank_conflict_ratio"),
            "waves_per_simd": v.get("waves_per_simd")}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6, help="timed batched iterations (each = all items)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mode", default="all", choices=["all", "plain", "batched", "sequence", "c2"],
                    help="sequence: consecutive windows along a smooth trajectory (own JSON line)")
    ap.add_argument("--seq-windows", type=int, default=100, help="sequence mode: windows (2000 = full sequence)")
    ap.add_argument("--seq-iters", type=int, default=50, help="sequence mode: iterations per window")
    ap.add_argument("--seq-hyp", type=int, default=16, help="sequence mode: re-initialisation hypotheses")
    ap.add_argument("--seq-velocity", default="finite_difference",
                    choices=["reference", "optimised", "finite_difference"],
                    help="sequence mode: velocity update after each window (see gsevt/sequence.py)")
    ap.add_argument("--seq-diag", default="", help="sequence mode: write per-window diagnostics (JSON) here")
    ap.add_argument("--seq-first", type=int, default=0, help="sequence mode: first window tracked")
    ap.add_argument("--seq-count", type=int, default=0, help="sequence mode: windows tracked (0: to the end)")
    ap.add_argument("--seq-hyp-vel", default="0.3,0.4",
                    help="sequence mode: velocity sigma (m/s,rad/s) of half the re-initialisation hypotheses")
    ap.add_argument("--seq-max-accel", default="30,30",
                    help="sequence mode: bound on the velocity change per window, 'm/s^2,rad/s^2' (empty: none)")
    ap.add_argument("--seq-accept-margin", type=float, default=0.0,
                    help="sequence mode: relative loss margin a re-initialisation hypothesis must win by")
    ap.add_argument("--plain-steps", type=int, default=10)
    ap.add_argument("--c2-iters", type=int, default=100, help="c2 mode: timed plain iterations")
    ap.add_argument("--hyp", type=int, default=64)
    ap.add_argument("--windows", type=int, default=4)
    ap.add_argument("--chunk", type=int, default=0, help="items per launch (0: all items in one launch)")
    ap.add_argument("--gaussians", type=int, default=1000000)
    ap.add_argument("--tile", default="16x16", help="plain path tile; batched start shape")
    ap.add_argument("--autotune", type=int, default=1,
                    help="batched path: pick the tile shape and blend schedule by timing")
    ap.add_argument("--schedules", default="grid,queue,queue_xcd",
                    help="batched path: blend schedules to try (first one without --autotune)")
    ap.add_argument("--queue-workgroups", type=int, default=2048, help="persistent blends: workgroups per launch")
    ap.add_argument("--reorder-map", type=int, default=2, choices=[0, 1, 2],
                    help="batched path, frozen-map layout before timing: 0 model order, 1 front to back for "
                         "the first item, 2 the same with the first item's contributing Gaussians first")
    ap.add_argument("--cpu-baseline", type=int, default=1)
    ap.add_argument("--cpu-sample-iters", type=int, default=1)
    ap.add_argument("--cpu-reps", type=int, default=5,
                    help="CPU baseline: repetitions of the 50-iteration C1 run (median reported)")
    ap.add_argument("--cpu-single-iters", type=int, default=5, help="CPU baseline: timed 1-thread C1 iterations")
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


def run_batched(prob, steps, warmup, chunk, tile, autotune=True, schedules=("grid",), queue_workgroups=2048,
                reorder=2):
    import torch
    from gsevt import raster
    from gsevt.fused import BatchedTracker
    intr = prob.intr
    B = len(prob.item_window)
    tr = BatchedTracker(prob.gaussians, intr.W, intr.H, intr.fx, intr.fy, prob.events, B, tile=tile,
                        chunk=chunk or None, device=prob.events.device, schedule=schedules[0],
                        queue_workgroups=queue_workgroups)
    lin = prob.gt_lin.expand(B, 3)
    ang = prob.gt_ang.expand(B, 3)
    dtau = [prob.dtau[w] for w in prob.item_window]

    def reset():
        tr.set_items(prob.init_poses, lin, ang, dtau, prob.item_window)

    reset()
    nmax = tr.measure_capacity()
    if reorder:
        # frozen-map layout front to back as the first hypothesis of the first
        # window sees it, its contributing Gaussians first (once, before
        # tuning and timing)
        tr.reorder_map_by_depth(0, contributors_first=reorder == 2)
        reset()
        nmax = tr.measure_capacity()
    tune = None
    if autotune:
        best, tune = tr.autotune(schedules=schedules)
        nmax = int(tr.nfrag.max().item())
        log("autotune (ms per iteration): " + ", ".join(f"{a}x{b} {sch}: {t:.1f}" for ((a, b), sch), t in tune.items())
            + f" -> {best[0][0]}x{best[0][1]} {best[1]}")
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
                tracker_bytes=tr.device_bytes(),
                tile=f"{tr.tile[0]}x{tr.tile[1]}", schedule=tr.schedule, reorder_map=reorder,
                autotune=None if tune is None else {f"{a}x{b} {sch}": t for ((a, b), sch), t in tune.items()})


def _instrument(tr):
    """Wrap the batched blend launches with HIP events (eager pass only)."""
    from gsevt import fused, raster
    orig = fused.call

    def wrapped(name, *args):
        if name.startswith("gsevt_tr_blend_forward") or name.startswith("gsevt_tr_blend_backward"):
            R = args[0]
            key = "tr_blend_forward" if "forward" in name else "tr_blend_backward"
            if tr.tiles is None:
                # packed: each depth-sorted value carries its Gaussian's tile
                # count above the index bits; a nonzero count is a visible one
                v = tr.dval[:R * tr.P].view(R, tr.P)
                nvis = ((v < 0) | (v >= (1 << tr.gid_bits))).sum(1)
            else:
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


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def _oracle_scene(g):
    from oracle import oracle as orc
    return orc.Scene(means=g.get_xyz.cpu().numpy(), scales=g.get_scaling.cpu().numpy(),
                     rots=g.get_rotation.cpu().numpy(), opac=g.get_opacity.reshape(-1).cpu().numpy(),
                     shs=g.get_features.cpu().numpy(), sh_degree=g.active_sh_degree)


def cpu_baseline(prob, iters, c1_iters=50, c1_warmup=5, c1_reps=1, c1_single_iters=5):
    """The oracle (C++ OpenMP restatement of the reference's tracking
    iteration) on this machine's host cores:
    * value: `iters` fine-stage iterations of item 0 of the benchmark's own
      1M-Gaussian 640x480 workload (the sample of the metric's workload);
    * c1: the CPU configuration (20k Gaussians SH0, 346x260, 8k events, 50
      Adam iterations of the fine stage from a 2 cm / 1 deg perturbation):
      c1_warmup untimed iterations, then c1_iters timed, median of c1_reps
      repetitions, at the OpenMP default thread count and at 1 thread
      (c1_single_iters timed iterations after 1 untimed)."""
    import statistics
    import numpy as np
    from oracle import oracle as orc
    from oracle import tracking_step as ts
    from gsevt import workloads
    n = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count())
    prev = orc.set_num_threads(n)
    intr = prob.intr
    scene = _oracle_scene(prob.gaussians)
    T0 = prob.init_poses[0].double().cpu().numpy()
    lin, ang = prob.gt_lin.double().cpu().numpy(), prob.gt_ang.double().cpu().numpy()
    e = prob.events[0].cpu().numpy()
    t0 = time.perf_counter()
    ts.track_fine(scene, T0, lin, ang, prob.dtau[0], e, iters, intr.W, intr.H, intr.fx, intr.fy)
    dt = time.perf_counter() - t0
    del scene
    # C1 on the CPU, as its configuration states it
    c1 = workloads.build("cpu_small")
    ci = c1.intr
    s1 = _oracle_scene(c1.gaussians)
    T1 = c1.init_pose.double().cpu().numpy()
    l1, a1 = c1.init_lin.double().cpu().numpy(), c1.init_ang.double().cpu().numpy()
    e1 = c1.sign_Ie[0].cpu().numpy()

    def run_c1(k):
        return ts.track_fine(s1, T1, l1, a1, c1.delta_tau, e1, k, ci.W, ci.H, ci.fx, ci.fy)

    run_c1(c1_warmup)
    times = []
    for _ in range(c1_reps):
        t1 = time.perf_counter()
        _, _, _, losses = run_c1(c1_iters)
        times.append(time.perf_counter() - t1)
    med = statistics.median(times)
    orc.set_num_threads(1)
    run_c1(1)
    t1 = time.perf_counter()
    run_c1(c1_single_iters)
    single = (time.perf_counter() - t1) / c1_single_iters
    orc.set_num_threads(prev)
    return dict(value=iters / dt, unit="iters/s", cores=n, kind="port", cpu_model=_cpu_model(),
                sample=f"{iters} fine-stage iteration(s) of item 0 of the same 1M-Gaussian SH3 640x480 workload "
                       f"(renders, loss, gradient, Adam, pose update) on the C++ oracle, OpenMP {n} threads",
                c1={"iters_per_s": c1_iters / med, "threads": n, "iterations": c1_iters, "warmup": c1_warmup,
                    "repetitions": c1_reps, "rep_seconds": times, "final_loss": float(losses[-1]),
                    "single_thread_iters_per_s": 1.0 / single, "single_thread_iterations": c1_single_iters,
                    "workload": "C1: 20k Gaussians SH0, 346x260, 8k events, fine stage from 2 cm / 1 deg",
                    "inputs_sha256": c1.raw_sha256})


def run_c2(args):
    """The drop-in baseline configuration: plain single-window tracking of the
    300k-Gaussian patch scene (SH3, 640x480, fx = fy = 550, 200k events drawn
    from the rendered intensity change under a 0.3 m/s, 20 deg/s twist), 100
    fine-stage Adam iterations through diff_gaussian_rasterization, one item
    at a time (utils/tracker.py:176-240 without the early exit).  Own JSON
    line."""
    import torch
    from gsevt import raster, workloads
    from gsevt.tracker import plain_fine_iterations
    prob = workloads.build("patches_300k")
    intr = prob.intr
    lr = (0.004, 0.004, 0.002, 0.002)
    tile = tuple(int(x) for x in args.tile.split("x"))
    plain_fine_iterations(prob.camera(), prob.gaussians, prob.background, prob.sign_Ie, args.warmup, lr, tile=tile)
    torch.cuda.synchronize()
    cam = prob.camera()
    e0 = float((cam.T - prob.gt_pose[:3, 3]).norm())
    raster.enable_kernel_timers(["blend_forward", "blend_backward"])
    losses = []
    t0 = time.perf_counter()
    plain_fine_iterations(cam, prob.gaussians, prob.background, prob.sign_Ie, args.c2_iters, lr, tile=tile,
                          losses=losses)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    timers = raster.disable_kernel_timers()
    HW = intr.W * intr.H
    ks = {}
    for name, lst in timers.items():
        ms = [st.elapsed_time(en) for st, en, _ in lst]
        nv = int((lst[0][2]["radii"] > 0).sum().item())
        nbytes = [blend_bytes(m["n"], nv, HW, name == "blend_backward") for _, _, m in lst]
        avg_ms = sum(ms) / len(ms)
        avg_b = sum(nbytes) / len(nbytes)
        ks[name] = dict(avg_ms=avg_ms, launches=len(ms), avg_bytes=avg_b,
                        hbm_frac=avg_b / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                        fragments=sum(m["n"] for _, _, m in lst) / len(lst), visible=nv)
    e1 = float((cam.T - prob.gt_pose[:3, 3]).norm())
    line = {
        "metric": "plain single-window tracking iters/sec (drop-in rasterizer), 640x480 @ 300k Gaussians",
        "value": args.c2_iters / dt, "unit": "iters/s", "n_gpus": 1, "higher_is_better": True,
        "iterations": args.c2_iters, "ms_per_iter": 1000 * dt / args.c2_iters, "renders_per_s": 2 * args.c2_iters / dt,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "patches_300k: 300000 Gaussians SH3 (64 planar patches), 640x480, 200k events, "
                               f"{args.c2_iters} fine-stage iterations", "path": "plain drop-in", "tile": args.tile},
        "kernels": ks, "loss": {"first": losses[0], "last": losses[-1]},
        "translation_error_m": {"initial": e0, "final": e1}, "inputs_sha256": prob.raw_sha256,
    }
    print(json.dumps(line), flush=True)


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
    seq = sequence.make_sequence(gm, intr, args.seq_windows, log=log, first=args.seq_first,
                                 count=args.seq_count or None)
    torch.cuda.synchronize()
    log(f"sequence: tracking ({args.seq_iters} iterations per window, {args.seq_hyp} hypotheses on re-init)")
    res = sequence.track_sequence(gm, intr, seq, iters=args.seq_iters, n_hyp=args.seq_hyp, log=log,
                                  velocity_update=args.seq_velocity, diag=bool(args.seq_diag),
                                  hyp_vel_sigma=tuple(float(x) for x in args.seq_hyp_vel.split(",")),
                                  accept_margin=args.seq_accept_margin,
                                  max_accel=tuple(float(x) for x in args.seq_max_accel.split(","))
                                  if args.seq_max_accel else None)
    e, a = sequence.drift(res.est, seq.gt_mid)
    if args.seq_diag:
        with open(args.seq_diag, "w") as f:
            json.dump(res.diag, f)
        log(f"sequence: per-window diagnostics in {args.seq_diag}")
    line = {
        "metric": "sequence tracking iters/sec, 640x480 @ 1M Gaussians",
        "value": res.iters_per_s(), "unit": "iters/s", "n_gpus": 1, "higher_is_better": True,
        "value_definition": "windows x iterations per window / wall time of the whole sequence "
                            "(re-initialisation time included, its hypothesis iterations not counted)",
        "windows": res.windows, "iterations_per_window": args.seq_iters, "iterations": res.iterations,
        "windows_per_s": res.windows_per_s(),
        "single_track_iters_per_s": res.iterations / max(res.single_seconds, 1e-9),
        "reinit": {"windows": len(res.reinit), "hypotheses": args.seq_hyp, "iterations": res.reinit_iterations,
                   "seconds": res.reinit_seconds,
                   "iters_per_s": res.reinit_iterations / res.reinit_seconds if res.reinit_seconds else None},
        "track_seconds": res.track_seconds, "reinit_windows": len(res.reinit),
        "drift": {"ate_rmse_m": float(e.pow(2).mean().sqrt()), "final_m": float(e[-1]), "max_m": float(e.max()),
                  "final_rot_deg": math.degrees(float(a[-1])), "max_rot_deg": math.degrees(float(a.max()))},
        "events_per_window": {"min": min(seq.counts), "max": max(seq.counts),
                              "mean": sum(seq.counts) / len(seq.counts)},
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"sequence_{res.windows}: {args.gaussians} Gaussians SH3, 640x480, B-spline "
                               f"trajectory (peak 1.5 m/s, 90 deg/s), 10 ms windows",
                   "path": "batched+fused, one graph per window", "tile": "16x16",
                   "velocity_update": args.seq_velocity, "trajectory_windows": args.seq_windows,
                   "first_window": args.seq_first, "reinit_hypotheses": args.seq_hyp,
                   "reinit_velocity_sigma": args.seq_hyp_vel, "reinit_accept_margin": args.seq_accept_margin,
                   "max_accel": args.seq_max_accel or None},
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
    if args.mode == "c2":
        run_c2(args)
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
        batched = run_batched(prob, args.steps, args.warmup, args.chunk, tile, bool(args.autotune),
                              tuple(args.schedules.split(",")), args.queue_workgroups, args.reorder_map)
        log(f"batched: {batched['iters_per_s']:.1f} iters/s")
    cpu = None
    if args.cpu_baseline:
        log("cpu baseline")
        cpu = cpu_baseline(prob, args.cpu_sample_iters, c1_reps=args.cpu_reps,
                           c1_single_iters=args.cpu_single_iters)
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
    used_sched = batched["schedule"] if batched else "grid"
    traffic = measured_traffic(dom, batched["chunk"] if batched else 1, args.gaussians, used_tile, used_sched)
    valu = measured_valu(dom, batched["chunk"] if batched else 1, args.gaussians, used_tile, used_sched, d["avg_ms"])
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
                   "blend_schedule": batched["schedule"] if batched else "grid",
                   "map_order": {0: "model order", 1: "front to back for item 0",
                                 2: "contributors to item 0 first, each part front to back"}[batched["reorder_map"]]
                   if batched else "model order",
                   "autotune_ms_per_iteration": batched["autotune"] if batched else None,
                   "tile_cull": batched["tile_cull"] if batched else False,
                   "fragments_per_render": batched["fragments_per_render"] if batched else None,
                   "tracker_hbm_gb": batched["tracker_bytes"] / 1e9 if batched else None,
                   "peak_hbm_gb": torch.cuda.max_memory_allocated() / 1e9},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "avg_ms": d["avg_ms"],
                     "algorithmic_bytes": d["avg_bytes"]},
        "valu_roofline": valu,
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
