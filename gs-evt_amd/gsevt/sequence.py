"""Full-sequence tracking on the batched fused path.

The per-window loop of the reference tracker (utils/tracker.py:105-286) is
kept: constant-velocity prediction of each window's pose from the previous
estimate (camera.py:183-201, tau = (dtau + dtau_prev) / 2), a fixed number of
fine-stage iterations (pose and velocity, signed event frame, pyramid level
0), and the 0.5 velocity blend after window 5 (camera.py:157-181).  The
iterations of a window run as one captured HIP graph.  When a window's final
loss exceeds `reinit_factor` times the median of the accepted losses so far,
the window is tracked again from `n_hyp` pose hypotheses around the
prediction in one batched graph, and the lowest-loss hypothesis is kept.

The trajectory is a seeded uniform cubic B-spline (positions and rotation
vectors), time-scaled so that its peak linear and angular speeds are the
requested ones.  Each window's events (count proportional to the
instantaneous speed, clipped to [min_events, max_events]) are sampled from the
rendered intensity change under the ground-truth motion.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import synthetic
from .camera import Camera
from .frame import EventArray, RenderFrame, event_frame
from .lie import se3_exp, so3_exp, so3_log


def _bspline_basis(u):
    """Uniform cubic B-spline weights and their derivatives (per unit knot
    spacing) at local parameter u in [0, 1)."""
    b = np.stack([(1 - u) ** 3, 3 * u ** 3 - 6 * u ** 2 + 4, -3 * u ** 3 + 3 * u ** 2 + 3 * u + 1, u ** 3], -1) / 6
    d = np.stack([-3 * (1 - u) ** 2, 9 * u ** 2 - 12 * u, -9 * u ** 2 + 6 * u + 3, 3 * u ** 2], -1) / 6
    return b, d


def _spline_eval(ctrl, t, knot):
    """Value and time derivative of the uniform cubic B-spline with control
    points ctrl [n, 3] and knot spacing `knot` at times t."""
    s = t / knot
    k = np.clip(np.floor(s).astype(np.int64), 0, len(ctrl) - 4)
    u = s - k
    b, d = _bspline_basis(u)
    idx = k[:, None] + np.arange(4)[None, :]
    val = np.einsum("tj,tjc->tc", b, ctrl[idx])
    der = np.einsum("tj,tjc->tc", d, ctrl[idx]) / knot
    return val, der


def _rot(rv):
    """Rotation matrices [n, 3, 3] of rotation vectors [n, 3] (float64)."""
    return so3_exp(torch.as_tensor(np.asarray(rv), dtype=torch.float64)).numpy()


@dataclass
class Trajectory:
    knot: float
    pos_ctrl: np.ndarray     # [n, 3] m
    rot_ctrl: np.ndarray     # [n, 3] rad (rotation vectors)
    duration: float

    def pose(self, t):
        """World->camera poses [len(t), 4, 4] (float64)."""
        t = np.atleast_1d(np.asarray(t, np.float64))
        p, _ = _spline_eval(self.pos_ctrl, t, self.knot)
        r, _ = _spline_eval(self.rot_ctrl, t, self.knot)
        T = np.zeros((len(t), 4, 4))
        T[:, :3, :3] = _rot(r)
        T[:, :3, 3] = p
        T[:, 3, 3] = 1.0
        return T

    def speeds(self, t, h=1e-4):
        """Linear (m/s) and angular (rad/s) speeds at times t."""
        t = np.atleast_1d(np.asarray(t, np.float64))
        _, dp = _spline_eval(self.pos_ctrl, t, self.knot)
        r0, _ = _spline_eval(self.rot_ctrl, t, self.knot)
        r1, _ = _spline_eval(self.rot_ctrl, t + h, self.knot)
        D = torch.as_tensor(_rot(r1) @ np.swapaxes(_rot(r0), 1, 2))
        ang = torch.linalg.norm(so3_log(D), dim=-1).numpy() / h
        return np.linalg.norm(dp, axis=1), ang


def bspline_trajectory(duration, peak_lin=1.5, peak_ang_deg=90.0, knot=0.25, pos_amp=0.3, rot_amp=0.25,
                       seed=synthetic.SEED_TRAJ, grid=4000):
    """Seeded smooth trajectory whose peak speeds over [0, duration] are
    peak_lin (m/s) and peak_ang_deg (deg/s): random control points (uniform
    in +-pos_amp m, +-rot_amp rad), then scaled to hit the peaks."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = int(math.ceil(duration / knot)) + 4
    pos = rng.uniform(-pos_amp, pos_amp, (n, 3))
    rot = rng.uniform(-rot_amp, rot_amp, (n, 3))
    tg = np.linspace(0.0, duration, grid, endpoint=False)
    tr = Trajectory(knot, pos, rot, duration)
    lin, _ = tr.speeds(tg)
    tr.pos_ctrl = pos * (peak_lin / lin.max())
    target = math.radians(peak_ang_deg)
    for _ in range(4):   # angular speed is nearly linear in the rotation amplitude
        _, ang = tr.speeds(tg)
        tr.rot_ctrl = tr.rot_ctrl * (target / ang.max())
    return tr


@dataclass
class Sequence:
    traj: Trajectory
    dtau: float
    gt_mid: torch.Tensor        # [N, 4, 4] world->camera at each window's midpoint
    gt_lin: torch.Tensor        # [N, 3] twist of the window (m/s)
    gt_ang: torch.Tensor        # [N, 3] (rad/s)
    frames: torch.Tensor        # [N, H, W] signed event frames
    counts: list
    first: int = 0              # index of the first window in the trajectory's window list


def _window_twist(T0, T1, dtau):
    """Left twist [v; w] with Exp(dtau [v; w]) T0 = T1 (T0, T1 float64 4x4)."""
    D = torch.as_tensor(T1 @ np.linalg.inv(T0))
    w = so3_log(D[:3, :3])
    th = float(torch.linalg.norm(w))
    Wm = torch.zeros(3, 3, dtype=torch.float64)
    Wm[0, 1], Wm[0, 2], Wm[1, 0], Wm[1, 2], Wm[2, 0], Wm[2, 1] = -w[2], w[1], w[2], -w[0], -w[1], w[0]
    if th < 1e-9:
        Vinv = torch.eye(3, dtype=torch.float64) - 0.5 * Wm
    else:
        Vinv = (torch.eye(3, dtype=torch.float64) - 0.5 * Wm
                + (1 / th ** 2) * (1 - th * math.sin(th) / (2 * (1 - math.cos(th)))) * (Wm @ Wm))
    rho = Vinv @ D[:3, 3]
    return (rho / dtau).numpy(), (w / dtau).numpy()


def make_sequence(gaussians, intr, n_windows, dtau=0.01, peak_lin=1.5, peak_ang_deg=90.0, min_events=10000,
                  max_events=500000, seed=synthetic.SEED_TRAJ, device="cuda", log=None, first=0, count=None):
    """Ground truth and signed event frames of consecutive windows of the
    trajectory of n_windows windows: windows first .. first + count - 1 (all
    of them by default; a part shares the whole sequence's trajectory,
    events and seeds window by window)."""
    dev = torch.device(device)
    traj = bspline_trajectory(n_windows * dtau, peak_lin, peak_ang_deg, seed=seed)
    t0 = np.arange(n_windows) * dtau
    lin_s, ang_s = traj.speeds(t0 + dtau / 2)
    s = lin_s / peak_lin + ang_s / math.radians(peak_ang_deg)
    counts = np.clip(np.round(max_events * s / s.max()), min_events, max_events).astype(np.int64)
    count = n_windows - first if count is None else count
    if not (0 <= first and count > 0 and first + count <= n_windows):
        raise ValueError(f"windows {first}..{first + count - 1} are not in 0..{n_windows - 1}")
    K = np.array([[intr.fx, 0, intr.W / 2], [0, intr.fy, intr.H / 2], [0, 0, 1.0]])
    bg = torch.zeros(3, dtype=torch.float32, device=dev)
    mids, lins, angs, frames = [], [], [], []
    for w in range(first, first + count):
        Ts = traj.pose([t0[w], t0[w] + dtau / 2, t0[w] + dtau])
        v, om = _window_twist(Ts[0], Ts[2], dtau)
        Tm = torch.as_tensor(Ts[1], dtype=torch.float32, device=dev)
        lin = torch.as_tensor(v, dtype=torch.float32, device=dev)
        ang = torch.as_tensor(om, dtype=torch.float32, device=dev)
        cam = Camera(Tm[:3, :3].clone(), Tm[:3, 3].clone(), ang.clone(), lin.clone(), intr.fovx, intr.fovy, intr.W,
                     intr.H, delta_tau=dtau, device=dev)
        cam.fx, cam.fy = intr.fx, intr.fy
        with torch.no_grad():
            dI = RenderFrame(cam, gaussians, None, bg, 0).sign_delta_Ir[0].float().cpu().numpy()
        ev = synthetic.events_from_delta(dI, int(counts[w]), noise_frac=0.1, duration_us=int(dtau * 1e6),
                                         seed=seed + 10 + w, t0_us=int(t0[w] * 1e6))
        ea = EventArray(ev["ts"], ev["x"], ev["y"], ev["p"])
        sign, _, _ = event_frame(ea.x, ea.y, ea.p, intr.W, intr.H, K, np.zeros(5), 9, dev)
        mids.append(torch.as_tensor(Ts[1], dtype=torch.float64))
        lins.append(torch.as_tensor(v))
        angs.append(torch.as_tensor(om))
        frames.append(sign[0])
        if log is not None and (w + 1 - first) % 100 == 0:
            log(f"sequence: {w + 1 - first}/{count} windows generated")
    return Sequence(traj, dtau, torch.stack(mids), torch.stack(lins), torch.stack(angs), torch.stack(frames),
                    [int(c) for c in counts[first:first + count]], first)


@dataclass
class SequenceResult:
    est: torch.Tensor                    # [N, 4, 4] estimated world->camera poses (float64, host)
    losses: list
    reinit: list                         # window indices re-initialised
    iterations: int                      # single-track iterations (windows x iterations per window)
    track_seconds: float                 # wall time of the whole sequence
    windows: int
    reinit_iterations: int = 0           # hypothesis iterations of the re-initialisations
    single_seconds: float = 0.0          # wall time inside the single-track graphs
    reinit_seconds: float = 0.0          # wall time inside the re-initialisation graphs
    diag: list = field(default_factory=list)   # per-window diagnostics (track_sequence(diag=True))

    def iters_per_s(self):
        """Single-track iterations per second of the whole sequence's wall
        time: re-initialisations cost time and add no iterations."""
        return self.iterations / self.track_seconds

    def windows_per_s(self):
        return self.windows / self.track_seconds

    def tum_rows(self, dtau):
        """Trajectory rows as the tracker writes them (utils/tracker.py:257-265):
        window mid time (s), the world->camera translation, the rotation as a
        quaternion (x, y, z, w)."""
        from .tracker import rotation_to_quat_xyzw
        rows = []
        for w in range(self.windows):
            T = self.est[w].numpy()
            q = rotation_to_quat_xyzw(T[:3, :3])
            rows.append((w * dtau + dtau / 2, *T[:3, 3].tolist(), *q.tolist()))
        return rows

    def write_tum(self, path, dtau):
        with open(path, "w") as f:
            for row in self.tum_rows(dtau):
                f.write(" ".join(str(v) for v in row) + "\n")


def _cam_center(T):
    return -T[..., :3, :3].transpose(-1, -2) @ T[..., :3, 3:4]


def drift(est, gt):
    """Per-window camera-centre error (m) and rotation error (rad)."""
    e = (_cam_center(est) - _cam_center(gt)).squeeze(-1).norm(dim=-1)
    dR = est[:, :3, :3] @ gt[:, :3, :3].transpose(1, 2)
    ang = torch.acos(((dR.diagonal(dim1=1, dim2=2).sum(1) - 1) / 2).clamp(-1, 1))
    return e, ang


VELOCITY_UPDATES = ("reference", "optimised", "finite_difference")


def _pose_errors(T, G):
    """Camera-centre error (m) and rotation error (rad) of pose T against G."""
    e, a = drift(T.unsqueeze(0), G.unsqueeze(0))
    return float(e[0]), float(a[0])


def track_sequence(gaussians, intr, seq: Sequence, iters=50, n_hyp=16, reinit_factor=1.25, hyp_trans=0.05,
                   hyp_rot_deg=3.0, tile=(16, 16), seed=synthetic.SEED_HYP, device="cuda", log=None,
                   velocity_update="reference", window_log=None, diag=False, hyp_vel_sigma=(0.0, 0.0),
                   accept_margin=0.0, max_accel=None):
    """Track every window in order (see the module docstring).

    velocity_update, applied after each window from window 5 on:
    * "reference": v <- 0.5 * log(T_est T_pred^-1) / tau + 0.5 * v_opt, the
      blend of camera.py:157-181, whose first term is the window's correction
      of the prediction;
    * "optimised": v <- v_opt (the velocity the fine stage optimised);
    * "finite_difference": v <- 0.5 * log(T_est T_est_prev^-1) / tau +
      0.5 * v_opt, with the previous window's estimate in place of the
      prediction.
    window_log: optional list receiving (window, loss, reinit) per window.
    diag: record per window the prediction's and the estimate's errors
    against the ground truth, the velocities and the re-initialisation's
    losses (SequenceResult.diag).
    Re-initialisation hypotheses: hypothesis 0 is the prediction; with
    hyp_vel_sigma = (m/s, rad/s) non-zero, the odd ones keep the predicted
    pose and draw the linear and angular velocity around the predicted ones
    (N(0, sigma^2) per axis), the even ones perturb the pose (hyp_trans m,
    hyp_rot_deg deg); otherwise all perturb the pose.  A hypothesis replaces
    the single track only if its loss is below (1 - accept_margin) times the
    single track's.
    max_accel = (m/s^2, rad/s^2): the velocities carried to the next window
    change by at most max_accel * tau from those the window was predicted
    with (a window whose estimate jumps cannot hand a velocity far outside
    the motion's acceleration range to the next prediction); None: no bound."""
    if velocity_update not in VELOCITY_UPDATES:
        raise ValueError(f"velocity_update must be one of {VELOCITY_UPDATES}")
    from .fused import BatchedTracker
    dev = torch.device(device)
    N = seq.frames.shape[0]
    ev = seq.frames[0:1].clone()
    one = BatchedTracker(gaussians, intr.W, intr.H, intr.fx, intr.fy, ev, 1, tile=tile, device=dev)
    many = BatchedTracker(gaussians, intr.W, intr.H, intr.fx, intr.fy, ev, n_hyp, tile=tile, device=dev,
                          chunk=n_hyp)
    rng = np.random.Generator(np.random.PCG64(seed))
    # the estimate starts at the ground-truth pose of t = 0 with the first
    # window's velocities (the reference reads both from its config); the
    # first prediction then moves it by dtau / 2, to the window midpoint
    T = torch.as_tensor(seq.traj.pose([seq.first * seq.dtau])[0])
    lin = seq.gt_lin[0].clone().float()
    ang = seq.gt_ang[0].clone().float()
    last_dtau = 0.0
    est, losses, reinit, records = [], [], [], []
    accepted = []
    n_it = n_reinit_it = 0
    t_single = t_reinit = 0.0
    graphs = {}

    def run(tr, poses, l, a):
        B = tr.B
        l, a = l.reshape(-1, 3).expand(B, 3), a.reshape(-1, 3).expand(B, 3)
        tr.events[0].copy_(seq.frames[w])
        tr.set_items(poses, l.expand(B, 3), a.expand(B, 3), [seq.dtau] * B, [0] * B)
        if tr.cap == 0:
            tr.measure_capacity(margin=1.5)
        if graphs.get(id(tr)) is None:
            tr.capture(steps=iters)
            graphs[id(tr)] = tr.graph
            tr.set_items(poses, l.expand(B, 3), a.expand(B, 3), [seq.dtau] * B, [0] * B)
        tr.graph.replay()
        torch.cuda.synchronize(dev)
        if int(tr.overflow.item()) != 0:   # capacity too small for this window: size again, replay
            tr.overflow.zero_()
            tr.set_items(poses, l.expand(B, 3), a.expand(B, 3), [seq.dtau] * B, [0] * B)
            tr.measure_capacity(margin=1.5)
            tr.capture(steps=iters)
            graphs[id(tr)] = tr.graph
            tr.set_items(poses, l.expand(B, 3), a.expand(B, 3), [seq.dtau] * B, [0] * B)
            tr.graph.replay()
            torch.cuda.synchronize(dev)
            tr.check_overflow()
        return tr.loss.clone(), tr.poses().cpu(), tr.lin.clone(), tr.ang.clone()

    t_start = time.perf_counter()
    last_log = t_start
    for w in range(N):
        tau = (seq.dtau + last_dtau) / 2
        xi = torch.cat([lin, ang]).double() * tau
        T_pred = se3_exp(xi.cpu()) @ T
        t0 = time.perf_counter()
        loss, P, l1, a1 = run(one, T_pred.unsqueeze(0), lin.to(dev), ang.to(dev))
        t_single += time.perf_counter() - t0
        n_it += iters
        L = float(loss[0])
        L_single = L
        T_new, lin_new, ang_new = P[0], l1[0].cpu().double(), a1[0].cpu().double()
        rec = None
        if diag:
            rec = dict(w=w, window=seq.first + w, events=int(seq.counts[w]), loss=L_single)
            rec["pred_err_m"], rec["pred_err_rad"] = _pose_errors(T_pred, seq.gt_mid[w])
            rec["single_err_m"], rec["single_err_rad"] = _pose_errors(T_new, seq.gt_mid[w])
        if len(accepted) >= 3 and L > reinit_factor * float(np.median(accepted)):
            hyps, hl, ha = [], [], []
            with_vel = hyp_vel_sigma[0] > 0 or hyp_vel_sigma[1] > 0
            for h in range(n_hyp):
                x = np.concatenate([rng.normal(0, hyp_trans, 3), rng.normal(0, math.radians(hyp_rot_deg), 3)])
                dv = np.concatenate([rng.normal(0, hyp_vel_sigma[0], 3), rng.normal(0, hyp_vel_sigma[1], 3)])
                if h == 0:
                    hyps.append(T_pred)
                    hl.append(lin)
                    ha.append(ang)
                elif with_vel and h % 2 == 1:
                    hyps.append(T_pred)
                    hl.append(lin + torch.as_tensor(dv[:3], dtype=lin.dtype))
                    ha.append(ang + torch.as_tensor(dv[3:], dtype=ang.dtype))
                else:
                    hyps.append(se3_exp(torch.as_tensor(x)) @ T_pred)
                    hl.append(lin)
                    ha.append(ang)
            t0 = time.perf_counter()
            lh, Ph, lh_lin, lh_ang = run(many, torch.stack(hyps), torch.stack(hl).to(dev), torch.stack(ha).to(dev))
            t_reinit += time.perf_counter() - t0
            n_reinit_it += iters * n_hyp
            k = int(torch.argmin(lh))
            reinit.append(w)
            if diag:
                errs = [_pose_errors(Ph[h], seq.gt_mid[w]) for h in range(n_hyp)]
                rec["hyp_loss"] = [float(x) for x in lh]
                rec["hyp_err_m"] = [e for e, _ in errs]
                rec["hyp_best"] = k
            if float(lh[k]) < L * (1.0 - accept_margin):
                L = float(lh[k])
                T_new, lin_new, ang_new = Ph[k], lh_lin[k].cpu().double(), lh_ang[k].cpu().double()
        accepted.append(L)
        losses.append(L)
        # velocity blend of camera.py:157-181 after window 5
        if w >= 5 and velocity_update != "optimised":
            ref = T_pred if velocity_update == "reference" else T
            D = T_new @ torch.linalg.inv(ref)
            fd_lin = D[:3, 3] / tau
            fd_ang = so3_log(D[:3, :3]) / tau
            lin_new = 0.5 * fd_lin + 0.5 * lin_new
            ang_new = 0.5 * fd_ang + 0.5 * ang_new
        if max_accel is not None:
            for new_v, old_v, a_max in ((lin_new, lin, max_accel[0]), (ang_new, ang, max_accel[1])):
                d = new_v - old_v.double()
                n = float(torch.linalg.norm(d))
                if n > a_max * tau:
                    new_v.copy_(old_v.double() + d * (a_max * tau / n))
        if window_log is not None:
            window_log.append((w, L, bool(reinit and reinit[-1] == w)))
        if diag:
            rec["final_loss"] = L
            rec["err_m"], rec["err_rad"] = _pose_errors(T_new, seq.gt_mid[w])
            rec["lin_err"] = float((lin_new - seq.gt_lin[w]).norm())
            rec["ang_err"] = float((ang_new - seq.gt_ang[w]).norm())
            records.append(rec)
        T, lin, ang = T_new, lin_new.float(), ang_new.float()
        est.append(T.clone())
        last_dtau = seq.dtau
        # every 100 windows, and at least every 30 s (windows that re-initialise
        # run 16 items each and can make 100 windows take minutes)
        if log is not None and ((w + 1) % 100 == 0 or time.perf_counter() - last_log > 30.0):
            log(f"sequence: {w + 1}/{N} windows tracked, {len(reinit)} re-initialised")
            last_log = time.perf_counter()
    dt = time.perf_counter() - t_start
    return SequenceResult(torch.stack(est), losses, reinit, n_it, dt, N, reinit_iterations=n_reinit_it,
                          single_seconds=t_single, reinit_seconds=t_reinit, diag=records)
