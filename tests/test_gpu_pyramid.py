"""The reference's three-level window optimisation (utils/tracker.py:149-240:
coarse stage at the top level switching to the fine stage on convergence,
fine stages at levels 1 and 0, caps, one Adam state across levels) on the
batched fused path (gsevt.pyramid.PyramidTracker; the stage switches happen
on the device), one window of the 20k-Gaussian cube scene (346x260):

* the device's stage decisions are the reference's: replaying
  tracker.py:176-236 (check_convergence of tracker.py:65-76 on the losses,
  the switch to the fine stage, the caps) on each item's own per-level loss
  sequence gives the same fine-stage start and the same iteration count, at
  a threshold where stages converge early and at one where they hit the caps;
* against the plain Tracker with the drop-in rasterizer on the same items:
  the same per-level iteration counts and stage switches where the stages run
  to their caps, and final poses within 5e-4 m / 2e-3 rad after 63
  iterations over the three levels.  The two paths round differently (grey
  splats against RGB renders made grey in torch; float atomics in both), and
  Adam normalises each entry's step: an entry whose gradient is at the FP32
  floor (as rotation about the optical axis is on the 1M workload) steps by
  up to its learning rate,
  0.004 rad, in a direction rounding decides.  Over two runs (four items)
  the rotation difference was 3e-4 to 1.1e-3 rad and the translation
  difference 2e-5 to 1.3e-4 m; the rotation bound is half an Adam step (over
  30 iterations at one level the bound is 1e-4 m, test_gpu_fused.py)."""
import copy
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _config(W, H, fx, fy, max_iter, thr):
    return {
        "Event": {"img_width": W, "img_height": H, "gaussian_kernel_size": 9,
                  "intrinsic": {"data": [fx, 0, W / 2, 0, fy, H / 2, 0, 0, 1]},
                  "distortion_factors": [0, 0, 0, 0, 0]},
        "Optimizer": {"cam_rot_delta": 0.004, "cam_trans_delta": 0.004, "cam_w_delta": 0.002,
                      "cam_v_delta": 0.002, "converged_threshold": thr, "max_optim_iter": max_iter},
    }


def _reference_flow(losses, coarse, max_iter, thr):
    """tracker.py:176-236 on a given loss sequence: (iterations run, fine
    stage start), or None if the sequence ends before the loop would."""
    from gsevt.tracker import check_convergence
    opt_vel, start, hist = not coarse, 0, []
    for it, L in enumerate(losses):
        hist.append(L)
        if check_convergence(hist, thr):
            if not opt_vel:
                opt_vel, start = True, it
            else:
                return it + 1, start
        if not opt_vel:
            if it >= max_iter:
                return it + 1, start
        elif it >= start + max_iter:
            return it + 1, start
    return None


def _setup(cuda, B, seed):
    from gsevt import synthetic, workloads
    from gsevt.lie import se3_exp
    prob = workloads.build("cpu_small", device=cuda)
    rng = np.random.default_rng(seed)
    starts = [prob.init_pose] + [
        se3_exp(torch.tensor(synthetic.fixed_norm_twist(rng, 0.01, math.radians(0.5)), dtype=torch.float32,
                             device=cuda)) @ prob.gt_pose for _ in range(B - 1)]
    return prob, starts


def _predicted(prob, starts, dtau):
    out = []
    for T0 in starts:
        pred = prob.camera(pose=T0)
        pred.const_vel_model(dtau / 2)      # track_window's prediction (last_delta_tau = 0)
        T = torch.eye(4, dtype=torch.float64)
        T[:3, :3], T[:3, 3] = pred.R.detach().double().cpu(), pred.T.detach().double().cpu()
        out.append(T)
    return torch.stack(out)


def _fused(prob, cuda, B, preds, max_iter, thr):
    from gsevt.frame import EventFrame
    from gsevt.pyramid import PyramidTracker
    intr = prob.intr
    K = np.array([[intr.fx, 0, intr.W / 2], [0, intr.fy, intr.H / 2], [0, 0, 1.0]])
    ef = EventFrame(intr.W, intr.H, K, np.zeros(5), 9, prob.events, device=cuda)
    pt = PyramidTracker(prob.gaussians, intr.W, intr.H, intr.fx, intr.fy, B, max_optim_iter=max_iter,
                        converged_threshold=thr, device=cuda)
    info = pt.track(preds, prob.init_lin.expand(B, 3), prob.init_ang.expand(B, 3), prob.events.duration(),
                    ef.sign_delta_Ie[0])
    return pt, info


@pytest.mark.parametrize("max_iter,thr", [(40, 0.02), (40, 0.05), (20, 1e-4)])
def test_pyramid_stage_decisions_follow_reference_flow(cuda, max_iter, thr):
    B = 4
    prob, starts = _setup(cuda, B, 9)
    preds = _predicted(prob, starts, prob.events.duration())
    pt, info = _fused(prob, cuda, B, preds, max_iter, thr)
    switched = converged = 0
    for lv in info:
        coarse = lv["level"] == 2
        for b in range(B):
            ref = _reference_flow(lv["losses"][b], coarse, max_iter, thr)
            print(f"[pyramid] thr {thr} item {b} level {lv['level']}: device {lv['iterations'][b]} iterations, "
                  f"fine from {lv['start_vel'][b]}; reference flow on these losses {ref}", flush=True)
            assert ref == (lv["iterations"][b], lv["start_vel"][b])
            assert lv["stage"][b] == 2
            switched += coarse and lv["start_vel"][b] > 0
            converged += lv["iterations"][b] < max_iter + 1
    if thr >= 0.02:
        assert switched > 0 and converged > 0     # the convergence branches ran


def test_pyramid_fused_vs_plain_tracker(cuda):
    from gsevt.tracker import Tracker
    B, max_iter, thr = 2, 20, 1e-4
    prob, starts = _setup(cuda, B, 8)
    intr = prob.intr
    dtau = prob.events.duration()
    cfg = _config(intr.W, intr.H, intr.fx, intr.fy, max_iter, thr)
    plain_poses, plain_levels = [], []
    for b in range(B):
        cam = prob.camera(pose=starts[b])
        tr = Tracker(cfg, [prob.events], cam, prob.gaussians, None, prob.background, device=cuda)
        tr.track_window(0, 0.0)
        Tp = np.eye(4)
        Tp[:3, :3], Tp[:3, 3] = cam.R.detach().cpu().numpy(), cam.T.detach().cpu().numpy()
        plain_poses.append(Tp)
        plain_levels.append({l["level"]: l for l in tr.window_log[0]["levels"]})
    pt, info = _fused(prob, cuda, B, _predicted(prob, starts, dtau), max_iter, thr)
    fused = pt.trackers[0].poses().cpu().numpy()
    for b in range(B):
        for lv in info:
            p = plain_levels[b][lv["level"]]
            print(f"[pyramid] item {b} level {lv['level']}: fused {lv['iterations'][b]} iterations (fine from "
                  f"{lv['start_vel'][b]}), plain optim_iter {p['iters']} (fine from {p['start_vel']})", flush=True)
            assert lv["iterations"][b] == p["iters"] + 1 and lv["start_vel"][b] == p["start_vel"]
        dt = np.abs(fused[b, :3, 3] - plain_poses[b][:3, 3]).max()
        dR = fused[b, :3, :3] @ plain_poses[b][:3, :3].T
        ang_err = math.acos(max(-1.0, min(1.0, (np.trace(dR) - 1) / 2)))
        moved = np.abs(plain_poses[b][:3, 3] - starts[b][:3, 3].cpu().numpy()).max()
        print(f"[pyramid] item {b}: |dt| {dt:.2e} m, angle {ang_err:.2e} rad (the plain path moved {moved:.2e} m)",
              flush=True)
        assert dt < 5e-4 and ang_err < 2e-3
