"""Window-sequence loop host logic: the B-spline trajectory hits the requested
peak speeds and is smooth, the window twist reproduces the window's end pose
exactly, and drift is measured on camera centres."""
import math
import sys
import os

import numpy as np
import torch

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gs-evt_amd"))

from gsevt import sequence as sq   # noqa: E402
from gsevt.lie import se3_exp      # noqa: E402


def test_trajectory_peak_speeds_and_smoothness():
    tr = sq.bspline_trajectory(3.0, peak_lin=1.5, peak_ang_deg=90.0, seed=5001)
    t = np.linspace(0.0, 3.0, 3000, endpoint=False)
    lin, ang = tr.speeds(t)
    assert abs(lin.max() - 1.5) < 1e-6
    assert abs(math.degrees(ang.max()) - 90.0) < 0.5     # grid of the scaling vs this grid
    # C1 continuity: finite-difference velocity matches the analytic one
    T = tr.pose(t)
    c = -np.einsum("nji,nj->ni", T[:, :3, :3], T[:, :3, 3])
    assert np.all(np.isfinite(c)) and np.abs(c).max() < 1.0
    _, dp = sq._spline_eval(tr.pos_ctrl, t, tr.knot)
    h = 1e-6
    p1, _ = sq._spline_eval(tr.pos_ctrl, t + h, tr.knot)
    p0, _ = sq._spline_eval(tr.pos_ctrl, t, tr.knot)
    np.testing.assert_allclose((p1 - p0) / h, dp, atol=1e-4)


def test_window_twist_reproduces_end_pose():
    tr = sq.bspline_trajectory(1.0, seed=5002)
    for t0 in (0.0, 0.37, 0.81):
        T0, T1 = tr.pose([t0, t0 + 0.01])
        v, w = sq._window_twist(T0, T1, 0.01)
        T1b = se3_exp(torch.as_tensor(np.concatenate([v, w]) * 0.01)).numpy() @ T0
        np.testing.assert_allclose(T1b, T1, atol=1e-12)


def test_drift_is_camera_centre_error():
    gt = torch.eye(4, dtype=torch.float64).repeat(3, 1, 1)
    est = gt.clone()
    est[1, :3, 3] = torch.tensor([0.0, 0.0, 0.02], dtype=torch.float64)
    R = se3_exp(torch.tensor([0, 0, 0, 0, 0, 0.1], dtype=torch.float64))
    est[2] = R
    e, a = sq.drift(est, gt)
    assert abs(float(e[1]) - 0.02) < 1e-12 and float(e[0]) == 0.0
    assert abs(float(a[2]) - 0.1) < 1e-9
