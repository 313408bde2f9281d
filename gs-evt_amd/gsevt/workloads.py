"""Tracking problems built from the synthetic generators: scene on the device,
ground-truth pose and twist, an event window sampled from the rendered
intensity change under the ground-truth motion, and a perturbed initial
camera.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from . import synthetic
from .camera import Camera
from .frame import EventArray, event_frame
from .lie import se3_exp
from .model import GaussianModel


@dataclass
class Problem:
    name: str
    gaussians: GaussianModel
    intr: synthetic.Intrinsics
    gt_pose: torch.Tensor          # [4,4] world->camera
    gt_lin: torch.Tensor
    gt_ang: torch.Tensor
    delta_tau: float
    events: EventArray
    K: np.ndarray
    dist: np.ndarray
    sign_Ie: torch.Tensor          # [1,H,W]
    unsign_Ie: torch.Tensor
    init_pose: torch.Tensor
    init_lin: torch.Tensor
    init_ang: torch.Tensor
    background: torch.Tensor
    raw_sha256: str

    def camera(self, pose=None, lin=None, ang=None):
        pose = self.init_pose if pose is None else pose
        cam = Camera(pose[:3, :3].clone(), pose[:3, 3].clone(), (self.init_ang if ang is None else ang).clone(),
                     (self.init_lin if lin is None else lin).clone(), self.intr.fovx, self.intr.fovy, self.intr.W,
                     self.intr.H, delta_tau=self.delta_tau, device=pose.device)
        cam.fx = self.intr.fx
        cam.fy = self.intr.fy
        return cam


def _gt_twist(rng, lin_speed, ang_speed_deg):
    a = rng.standard_normal(3)
    b = rng.standard_normal(3)
    lin = a / np.linalg.norm(a) * lin_speed
    ang = b / np.linalg.norm(b) * math.radians(ang_speed_deg)
    return lin, ang


def build(name="patches_1m", P=None, n_events=None, device="cuda", seed_offset=0, init_trans=0.02,
          init_rot_deg=1.0, delta_tau=0.01):
    """Named problems: cpu_small (20k cube, SH0, 346x260), patches_300k,
    patches_1m (SH3, 640x480, fx = fy = 550)."""
    dev = torch.device(device)
    if name == "cpu_small":
        raw = synthetic.cube_scene(P=P or 20000, sh_degree=0)
        intr = synthetic.CPU_SMALL_CAM
        n_events = n_events or 8000
    elif name.startswith("patches"):
        P = P or (1000000 if name == "patches_1m" else 300000)
        raw = synthetic.patch_scene(P=P, sh_degree=3)
        intr = synthetic.PATCH_CAM
        n_events = n_events or (150000 if name == "patches_1m" else 200000)
    else:
        raise ValueError(name)
    act = raw.activated()
    gm = GaussianModel.from_raw(raw.xyz, raw.features_dc, raw.features_rest, raw.scaling, raw.rotation,
                                raw.opacity, raw.sh_degree, device=dev)
    rng = np.random.Generator(np.random.PCG64(synthetic.SEED_EVENTS + seed_offset))
    lin, ang = _gt_twist(rng, 0.3, 20.0)
    f32 = dict(dtype=torch.float32, device=dev)
    gt_pose = torch.eye(4, **f32)
    gt_lin = torch.tensor(lin, **f32)
    gt_ang = torch.tensor(ang, **f32)
    bg = torch.zeros(3, **f32)
    # events from the rendered intensity change under the ground-truth motion
    from .frame import RenderFrame
    gt_cam = Camera(gt_pose[:3, :3].clone(), gt_pose[:3, 3].clone(), gt_ang.clone(), gt_lin.clone(), intr.fovx,
                    intr.fovy, intr.W, intr.H, delta_tau=delta_tau, device=dev)
    gt_cam.fx, gt_cam.fy = intr.fx, intr.fy
    with torch.no_grad():
        rf = RenderFrame(gt_cam, gm, None, bg, 0)
        dI = rf.sign_delta_Ir[0].float().cpu().numpy()
    ev = synthetic.events_from_delta(dI, n_events, noise_frac=0.1, duration_us=int(delta_tau * 1e6),
                                     seed=synthetic.SEED_EVENTS + 10 + seed_offset)
    ea = EventArray(ev["ts"], ev["x"], ev["y"], ev["p"])
    K = np.array([[intr.fx, 0, intr.W / 2], [0, intr.fy, intr.H / 2], [0, 0, 1.0]])
    dist = np.zeros(5)
    sign, absf, _ = event_frame(ea.x, ea.y, ea.p, intr.W, intr.H, K, dist, 9, dev)
    rng2 = np.random.Generator(np.random.PCG64(synthetic.SEED_INIT + seed_offset))
    xi = synthetic.fixed_norm_twist(rng2, init_trans, math.radians(init_rot_deg))
    init_pose = se3_exp(torch.tensor(xi, **f32)) @ gt_pose
    return Problem(name=name, gaussians=gm, intr=intr, gt_pose=gt_pose, gt_lin=gt_lin, gt_ang=gt_ang,
                   delta_tau=ea.duration(), events=ea, K=K, dist=dist, sign_Ie=sign, unsign_Ie=absf,
                   init_pose=init_pose, init_lin=gt_lin.clone(), init_ang=gt_ang.clone(), background=bg,
                   raw_sha256=raw.sha256())


@dataclass
class BatchedProblem:
    name: str
    gaussians: GaussianModel
    intr: synthetic.Intrinsics
    events: torch.Tensor          # [n_windows, H, W] signed event frames
    event_counts: list
    window_gt: torch.Tensor       # [n_windows, 4, 4]
    gt_lin: torch.Tensor
    gt_ang: torch.Tensor
    dtau: list                    # per window (s)
    init_poses: torch.Tensor      # [B, 4, 4] item order: window-major
    item_window: list
    n_hyp: int
    raw_sha256: str


def build_batched(P=1000000, n_hyp=64, n_windows=4, mean_events=150000, device="cuda", hyp_trans=0.05,
                  hyp_rot_deg=3.0, delta_tau=0.01):
    """n_hyp pose hypotheses (Exp(xi) T_gt, xi ~ N(0, diag(5 cm^2, 3 deg^2)))
    for each of n_windows consecutive 10 ms windows of a constant-twist motion
    (0.3 m/s, 20 deg/s); events per window Poisson(mean_events) clipped to
    [20k, 500k], sampled from the rendered intensity change of the window."""
    from .frame import RenderFrame
    dev = torch.device(device)
    raw = synthetic.patch_scene(P=P, sh_degree=3)
    intr = synthetic.PATCH_CAM
    gm = GaussianModel.from_raw(raw.xyz, raw.features_dc, raw.features_rest, raw.scaling, raw.rotation,
                                raw.opacity, raw.sh_degree, device=dev)
    rng = np.random.Generator(np.random.PCG64(synthetic.SEED_EVENTS))
    lin, ang = _gt_twist(rng, 0.3, 20.0)
    f32 = dict(dtype=torch.float32, device=dev)
    gt_lin = torch.tensor(lin, **f32)
    gt_ang = torch.tensor(ang, **f32)
    bg = torch.zeros(3, **f32)
    counts = np.clip(rng.poisson(mean_events, n_windows), 20000, 500000)
    K = np.array([[intr.fx, 0, intr.W / 2], [0, intr.fy, intr.H / 2], [0, 0, 1.0]])
    frames, gts, dts = [], [], []
    xi_step = torch.cat([gt_lin, gt_ang]) * delta_tau
    for w in range(n_windows):
        T = se3_exp(xi_step * w)
        gts.append(T)
        cam = Camera(T[:3, :3].clone(), T[:3, 3].clone(), gt_ang.clone(), gt_lin.clone(), intr.fovx, intr.fovy,
                     intr.W, intr.H, delta_tau=delta_tau, device=dev)
        cam.fx, cam.fy = intr.fx, intr.fy
        with torch.no_grad():
            dI = RenderFrame(cam, gm, None, bg, 0).sign_delta_Ir[0].float().cpu().numpy()
        ev = synthetic.events_from_delta(dI, int(counts[w]), noise_frac=0.1, duration_us=int(delta_tau * 1e6),
                                         seed=synthetic.SEED_EVENTS + 100 + w, t0_us=int(w * delta_tau * 1e6))
        ea = EventArray(ev["ts"], ev["x"], ev["y"], ev["p"])
        sign, _, _ = event_frame(ea.x, ea.y, ea.p, intr.W, intr.H, K, np.zeros(5), 9, dev)
        frames.append(sign[0])
        dts.append(ea.duration())
    rh = np.random.Generator(np.random.PCG64(synthetic.SEED_HYP))
    poses, item_window = [], []
    for w in range(n_windows):
        for h in range(n_hyp):
            xi = np.concatenate([rh.normal(0, hyp_trans, 3), rh.normal(0, math.radians(hyp_rot_deg), 3)])
            poses.append(se3_exp(torch.tensor(xi, **f32)) @ gts[w])
            item_window.append(w)
    return BatchedProblem(name=f"batched_{n_hyp}x{n_windows}", gaussians=gm, intr=intr, events=torch.stack(frames),
                          event_counts=[int(c) for c in counts], window_gt=torch.stack(gts), gt_lin=gt_lin,
                          gt_ang=gt_ang, dtau=dts, init_poses=torch.stack(poses), item_window=item_window,
                          n_hyp=n_hyp, raw_sha256=raw.sha256())
