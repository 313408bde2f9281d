"""Event-based pose tracker (plain path).

Mirrors utils/tracker.py:33-286 of the reference with the drop-in rasterizer:
per window a fresh Adam over (cam_rot_delta, cam_trans_delta, cam_w_delta,
cam_v_delta), constant-velocity prediction, event pyramid, coarse stage at the
top level (pose only, unsigned images) switching to the fine stage on
convergence, fine stages with the learning-rate ramp
fraction = (it - start) / (max_iter / 2) (pose LR x fraction, velocity LR x
(1 - fraction), fraction = 1 outside 1..max_iter/2), iteration caps, the
velocity blend after window 5 and a TUM trajectory line per window.  With
config["Tracking"]["save_path"] set, a run writes what tracker.py:105-288
writes there: tracking_log.log, tracking_pose_tum.txt, one overlay frame per
window in tracking_frames/, tracking_video.gif and (with OpenCV installed)
tracking_video.mp4.

Deviations, each because the reference code cannot run as written:
* the SIGINT/SIGTERM handler takes (signum, frame) (tracker.py:24 takes none
  and raises TypeError when a signal arrives);
* if the first window's coarse stage hits its cap without converging,
  start_vel_opt_iter is 0 (tracker.py:245 would raise UnboundLocalError).
"""
from __future__ import annotations

import copy
import logging
import os
import signal
import time
from typing import List

import numpy as np
import torch

from .frame import EventArray, EventFrame, RenderFrame, image_pyramid

log = logging.getLogger("gsevt.tracker")

_stop = {"flag": False}


def _handle_stop(signum=None, frame=None):
    _stop["flag"] = True


def install_signal_handlers():
    signal.signal(signal.SIGINT, _handle_stop)
    signal.signal(signal.SIGTERM, _handle_stop)


def check_convergence(losses, threshold=1e-4):
    """tracker.py:65-76."""
    if len(losses) <= 10:
        return False
    slopes = np.diff(losses[-11:])
    return bool(np.mean(np.abs(slopes)) < threshold)


def tracking_loss(delta_Ir, delta_Ie):
    """tracker.py:93-103 (huber=False, no mask): un-squared Frobenius norm."""
    return torch.norm(delta_Ir - delta_Ie)


def lr_fraction(optim_iter, start_vel_opt_iter, max_optim_iter):
    """tracker.py:188-191."""
    fraction_num = max_optim_iter / 2
    if 1 <= (optim_iter - start_vel_opt_iter) <= fraction_num:
        return (optim_iter - start_vel_opt_iter) / fraction_num
    return 1


def rotation_to_quat_xyzw(R: np.ndarray) -> np.ndarray:
    """scipy Rotation.from_matrix(R).as_quat() for a proper rotation (x, y, z, w)."""
    from scipy.spatial.transform import Rotation
    return Rotation.from_matrix(R).as_quat()


class Tracker:
    PYRAMID_LEVELS = 3

    def __init__(self, config, event_arrays: List[EventArray], viewpoint, gaussians, pipeline, background,
                 device="cuda", tile=(16, 16)):
        self.config = config
        self.event_arrays = event_arrays
        self.viewpoint = viewpoint
        self.gaussians = gaussians
        self.pipeline = pipeline
        self.background = background
        self.device = device
        self.tile = tile
        ev = config["Event"]
        self.img_width = ev["img_width"]
        self.img_height = ev["img_height"]
        self.gaussian_kernel_size = ev["gaussian_kernel_size"]
        self.intrinsic = np.array(ev["intrinsic"]["data"]).reshape(3, 3)
        self.distortion_factors = np.array(ev["distortion_factors"])
        opt = config["Optimizer"]
        self.lr = dict(rot=opt["cam_rot_delta"], trans=opt["cam_trans_delta"], w=opt["cam_w_delta"],
                       v=opt["cam_v_delta"])
        self.converged_threshold = opt["converged_threshold"]
        self.max_optim_iter = opt["max_optim_iter"]
        self.trajectory = []     # TUM rows (t, tx, ty, tz, qx, qy, qz, qw)
        self.window_log = []
        self.iterations = 0
        self.save_path = (config.get("Tracking") or {}).get("save_path")
        self.run_log = None     # the run's file + console log while tracking() runs

    def _log(self, level, msg):
        getattr(log, level)(msg)
        if self.run_log is not None:
            getattr(self.run_log, level)(msg)

    def _optimizer(self):
        vp = self.viewpoint
        return torch.optim.Adam([
            {"params": [vp.cam_rot_delta], "lr": self.lr["rot"]},
            {"params": [vp.cam_trans_delta], "lr": self.lr["trans"]},
            {"params": [vp.cam_w_delta], "lr": self.lr["w"]},
            {"params": [vp.cam_v_delta], "lr": self.lr["v"]},
        ])

    def _set_lr(self, optimizer, rot_trans_scale, vel_scale):
        scales = [rot_trans_scale, rot_trans_scale, vel_scale, vel_scale]
        bases = [self.lr["rot"], self.lr["trans"], self.lr["w"], self.lr["v"]]
        for g, b, s in zip(optimizer.param_groups, bases, scales):
            g["lr"] = b * s

    def track_window(self, frame_idx, last_delta_tau):
        vp = self.viewpoint
        optimizer = self._optimizer()
        ea = self.event_arrays[frame_idx]
        delta_tau = ea.duration()
        vp.delta_tau = delta_tau
        vp.const_vel_model((delta_tau + last_delta_tau) / 2)
        initial = copy.deepcopy([vp.T.detach(), vp.R.detach(), vp.angular_vel.detach(), vp.linear_vel.detach()])
        ef = EventFrame(self.img_width, self.img_height, self.intrinsic, self.distortion_factors,
                        self.gaussian_kernel_size, ea, device=self.device)
        sign_pyr = image_pyramid(ef.sign_delta_Ie, self.PYRAMID_LEVELS)
        unsign_pyr = image_pyramid(ef.unsign_delta_Ie, self.PYRAMID_LEVELS)
        start_vel_opt_iter = 0
        info = dict(frame_idx=frame_idx, delta_tau=delta_tau, levels=[])
        self._log("info", f"frame_idx:\t{frame_idx} / {len(self.event_arrays)}")
        self._log("info", f"delta_tau:\t{delta_tau:.4f}")
        rf = None
        for lvl in range(self.PYRAMID_LEVELS - 1, -1, -1):
            sign_Ie, unsign_Ie = sign_pyr[lvl], unsign_pyr[lvl]
            opt_vel = lvl != self.PYRAMID_LEVELS - 1
            if opt_vel:
                start_vel_opt_iter = 0
            self._set_lr(optimizer, 1.0, 1.0)
            losses = []
            optim_iter = 0
            t0 = time.time()
            while True:
                if not opt_vel:
                    vp.cam_w_delta.requires_grad_(False)
                    vp.cam_v_delta.requires_grad_(False)
                    vp.cam_rot_delta.requires_grad_(True)
                    vp.cam_trans_delta.requires_grad_(True)
                else:
                    for p in (vp.cam_w_delta, vp.cam_v_delta, vp.cam_rot_delta, vp.cam_trans_delta):
                        p.requires_grad_(True)
                    f = lr_fraction(optim_iter, start_vel_opt_iter, self.max_optim_iter)
                    self._set_lr(optimizer, f, 1 - f)
                rf = RenderFrame(vp, self.gaussians, self.pipeline, self.background, lvl, tile=self.tile)
                if not opt_vel:
                    loss = tracking_loss(rf.unsign_delta_Ir, unsign_Ie)
                else:
                    loss = tracking_loss(rf.sign_delta_Ir, sign_Ie)
                loss.backward()
                losses.append(loss.item())
                self.iterations += 1
                with torch.no_grad():
                    optimizer.step()
                    converged = check_convergence(losses, self.converged_threshold)
                    if not opt_vel:
                        vp.update_pose()
                    else:
                        vp.update_vwRT()
                    optimizer.zero_grad()
                if converged:
                    if not opt_vel:
                        opt_vel = True
                        start_vel_opt_iter = optim_iter
                    else:
                        break
                if not opt_vel:
                    if optim_iter >= self.max_optim_iter:
                        self._log("error", "coarse stage optimization iter exceeded the max_optim_iter!")
                        break
                else:
                    if optim_iter >= start_vel_opt_iter + self.max_optim_iter:
                        self._log("error", "fine stage optimization iter exceeded the max_optim_iter!")
                        break
                optim_iter += 1
            opt_time = time.time() - t0
            info["levels"].append(dict(level=lvl, iters=optim_iter, start_vel=start_vel_opt_iter,
                                       time=opt_time, final_loss=losses[-1]))
            self._log("info", f"level:\t{lvl}")
            self._log("info", f"optim_iter:\t{optim_iter} ({start_vel_opt_iter}+{optim_iter - start_vel_opt_iter})")
            self._log("info", f"opt_time:\t{opt_time:.4f}")
        if frame_idx >= 5:
            vp.cal_weighted_velocity(initial[:2], (delta_tau + last_delta_tau) / 2, 0.5)
        t = ea.time()
        tr = vp.T.detach().cpu().numpy()
        q = rotation_to_quat_xyzw(vp.R.detach().cpu().numpy().astype(np.float64))
        self.trajectory.append((t, *tr.tolist(), *q.tolist()))
        self.window_log.append(info)
        self._log("debug", f"angular_vel:\t{vp.angular_vel}")
        self._log("debug", f"linear_vel:\t{vp.linear_vel}")
        self._log("info", "=" * 20)
        if self.save_path:
            from .visualize import delta_Ie_image, delta_Ir_image, overlay, write_png
            with open(f"{self.save_path}/tracking_pose_tum.txt", "a") as f:
                f.write(" ".join(str(v) for v in self.trajectory[-1]) + "\n")
            img = overlay(delta_Ir_image(rf.sign_delta_Ir), delta_Ie_image(ef.sign_delta_Ie))
            os.makedirs(f"{self.save_path}/tracking_frames", exist_ok=True)
            write_png(os.path.join(f"{self.save_path}/tracking_frames", f"frame_{frame_idx}.png"), img)
        return delta_tau

    def tracking(self):
        if self.save_path:
            from .visualize import TrackingLogger
            os.makedirs(f"{self.save_path}/tracking_frames", exist_ok=True)
            open(f"{self.save_path}/tracking_pose_tum.txt", "w").close()
            self.run_log = TrackingLogger("TrackingLogger", f"{self.save_path}/tracking_log.log", logging.INFO)
        last_delta_tau = 0
        frame_idx = 0
        t0 = time.time()
        while not _stop["flag"] and frame_idx < len(self.event_arrays):
            last_delta_tau = self.track_window(frame_idx, last_delta_tau)
            frame_idx += 1
        self._log("info", f"totoal tracking time cost: {time.time() - t0:.4f}s")
        if self.save_path and frame_idx > 0:
            from .visualize import save_gif, save_video
            img_dir = f"{self.save_path}/tracking_frames"
            video_save_path = f"{self.save_path}/tracking_video.mp4"
            gif_save_path = f"{self.save_path}/tracking_video.gif"
            try:
                save_video(img_dir, video_save_path, fps=120)
                self._log("info", f"save video in: {video_save_path}")
            except RuntimeError as e:
                self._log("warning", f"no video: {e}")
            save_gif(img_dir, gif_save_path, duration=2)
            self._log("info", f"save gif in  : {gif_save_path}")
        if self.run_log is not None:
            self.run_log.close()
            self.run_log = None
        return self.trajectory

    def write_tum(self, path):
        with open(path, "w") as f:
            for row in self.trajectory:
                f.write(" ".join(str(v) for v in row) + "\n")


def plain_fine_iterations(viewpoint, gaussians, background, sign_Ie, n_iters, lr, max_optim_iter=200,
                          tile=(16, 16), losses=None, sync_each_iter=True):
    """Bench mode of the plain path: level 0, fine stage, fixed iteration
    count (tracker.py:176-240 without the early exit).  Each iteration is one
    render of the last and next camera, the loss, autograd through the
    drop-in rasterizer, Adam and the pose/velocity update."""
    vp = viewpoint
    opt = torch.optim.Adam([
        {"params": [vp.cam_rot_delta], "lr": lr[0]},
        {"params": [vp.cam_trans_delta], "lr": lr[1]},
        {"params": [vp.cam_w_delta], "lr": lr[2]},
        {"params": [vp.cam_v_delta], "lr": lr[3]},
    ])
    for p in (vp.cam_w_delta, vp.cam_v_delta, vp.cam_rot_delta, vp.cam_trans_delta):
        p.requires_grad_(True)
    for it in range(n_iters):
        f = lr_fraction(it, 0, max_optim_iter)
        for g, b, s in zip(opt.param_groups, lr, (f, f, 1 - f, 1 - f)):
            g["lr"] = b * s
        rf = RenderFrame(vp, gaussians, None, background, 0, tile=tile)
        loss = tracking_loss(rf.sign_delta_Ir, sign_Ie)
        loss.backward()
        lv = loss.item() if sync_each_iter else loss.detach()
        if losses is not None:
            losses.append(lv)
        with torch.no_grad():
            opt.step()
            vp.update_vwRT()
            opt.zero_grad()
    return vp
