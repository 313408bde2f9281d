"""CPU restatement of one fine-stage tracking iteration (oracle side).

TEST INFRASTRUCTURE ONLY (used by tests/ and as bench.py's cpu_baseline).
Follows utils/render_camera/frame.py:61-94 (grey intensity, difference, L2
normalisation) and utils/tracker.py:93-103,212-222 (un-squared norm loss,
backward, Adam, left-multiplicative pose update), with the rasterizer
replaced by the C++ oracle.
"""
from __future__ import annotations

import numpy as np

import math

from . import oracle as orc

GREY = np.array([0.2989, 0.5870, 0.1140])
ZNEAR, ZFAR = 0.01, 100.0


def skew(v):
    return np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]], dtype=np.float64)


def se3_exp(xi):
    """SE(3) exponential of [rho; theta] in float64 (utils/pose.py:78-91,
    the small-angle branch below 1e-12 rad)."""
    rho, th = np.asarray(xi[:3], np.float64), np.asarray(xi[3:], np.float64)
    a = float(np.linalg.norm(th))
    Wm = skew(th)
    if a < 1e-12:
        R = np.eye(3) + Wm + 0.5 * Wm @ Wm
        V = np.eye(3) + 0.5 * Wm + Wm @ Wm / 6.0
    else:
        R = np.eye(3) + math.sin(a) / a * Wm + (1 - math.cos(a)) / a ** 2 * Wm @ Wm
        V = np.eye(3) + (1 - math.cos(a)) / a ** 2 * Wm + (a - math.sin(a)) / a ** 3 * Wm @ Wm
    T = np.eye(4)
    T[:3, :3] = R
    T[:3, 3] = V @ rho
    return T


def camera(T, W, H, fx, fy, vel=None, delta_time=0.0, lin=None, ang=None):
    """Oracle camera of world->camera pose T (utils/render_camera/camera.py:
    68-91, graphics_utils.py:33-69): matrices in the settings' transposed
    storage, rounded to float32 as the rasterizer receives them; vel is the
    4x4 velocity transform of the render (stored transposed)."""
    fovx, fovy = 2 * math.atan(W / (2 * fx)), 2 * math.atan(H / (2 * fy))
    tx, ty = math.tan(fovx / 2), math.tan(fovy / 2)
    Pm = np.zeros((4, 4))
    Pm[0, 0] = 2 * ZNEAR / (2 * tx * ZNEAR)
    Pm[1, 1] = 2 * ZNEAR / (2 * ty * ZNEAR)
    Pm[3, 2] = 1.0
    Pm[2, 2] = ZFAR / (ZFAR - ZNEAR)
    Pm[2, 3] = -(ZFAR * ZNEAR) / (ZFAR - ZNEAR)
    T = np.asarray(T, np.float64)
    f = lambda a: np.asarray(a, np.float32)
    return orc.Cam(view=f(T.T.reshape(-1)), proj=f((Pm @ T).T.reshape(-1)), projraw=f(Pm.T.reshape(-1)),
                   campos=f(-T[:3, :3].T @ T[:3, 3]), W=W, H=H, tanfovx=float(np.float32(tx)),
                   tanfovy=float(np.float32(ty)), bg=np.zeros(3, np.float32),
                   vel=None if vel is None else np.asarray(vel, np.float64).T.reshape(-1), delta_time=delta_time,
                   lin=lin, ang=ang)


def delta_ir_and_grad(color_last, color_next, sign_Ie, want_grad=True):
    """Returns (loss, dL/dcolor_last [3,H,W], dL/dcolor_next [3,H,W], r)."""
    I_l = np.tensordot(GREY, color_last.astype(np.float64), axes=(0, 0))
    I_n = np.tensordot(GREY, color_next.astype(np.float64), axes=(0, 0))
    d = I_n - I_l
    nrm = np.linalg.norm(d)
    r = d / nrm
    res = r - sign_Ie.astype(np.float64)
    loss = np.linalg.norm(res)
    if not want_grad:
        return loss, None, None, r
    g = res / loss
    gd = (g - r * np.sum(r * g)) / nrm
    gn = GREY[:, None, None] * gd[None]
    return loss, -gn, gn, r


def iteration(scene, cam_last, cam_next, sign_Ie, dtype=np.float32):
    """One iteration's rendering, loss and 12-float gradient
    (theta, rho, w, v), summed over the two renders."""
    st_l = orc.render(scene, cam_last, dtype=dtype, n_touched=False)
    st_n = orc.render(scene, cam_next, dtype=dtype, n_touched=False)
    loss, g_l, g_n, _ = delta_ir_and_grad(st_l["fwd"]["color"], st_n["fwd"]["color"], sign_Ie)
    b_l = orc.render_backward(scene, cam_last, st_l, g_l.astype(dtype), dtype=dtype)
    b_n = orc.render_backward(scene, cam_next, st_n, g_n.astype(dtype), dtype=dtype)
    grads = {k: b_l["pose"][k] + b_n["pose"][k] for k in ("theta", "rho", "w", "v")}
    return loss, grads


class Adam:
    """torch.optim.Adam defaults (betas 0.9/0.999, eps 1e-8), one state per
    3-vector parameter; parameters whose gradient is None are skipped, as
    torch does."""

    def __init__(self, lrs):
        self.lrs = dict(lrs)
        self.m = {k: np.zeros(3) for k in lrs}
        self.v = {k: np.zeros(3) for k in lrs}
        self.t = {k: 0 for k in lrs}

    def step(self, params, grads, lr_scale=None):
        out = {}
        for k, p in params.items():
            g = grads.get(k)
            if g is None:
                out[k] = p
                continue
            self.t[k] += 1
            t = self.t[k]
            self.m[k] = 0.9 * self.m[k] + 0.1 * g
            self.v[k] = 0.999 * self.v[k] + 0.001 * g * g
            lr = self.lrs[k] * (1.0 if lr_scale is None else lr_scale[k])
            bc1 = 1 - 0.9 ** t
            bc2 = 1 - 0.999 ** t
            denom = np.sqrt(self.v[k]) / np.sqrt(bc2) + 1e-8
            out[k] = p - (lr / bc1) * self.m[k] / denom
        return out


def lr_fraction(optim_iter, start, max_optim_iter):
    """utils/tracker.py:188-191."""
    n = max_optim_iter / 2
    return (optim_iter - start) / n if 1 <= optim_iter - start <= n else 1


def track_fine(scene, T0, lin, ang, dtau, sign_Ie, iters, W, H, fx, fy, lr=(0.004, 0.004, 0.002, 0.002),
               max_optim_iter=200):
    """The fine stage at pyramid level 0 for a fixed number of iterations
    (utils/tracker.py:176-240 without the early exit) on the oracle: per
    iteration the last and next renders at Exp(-/+ dtau/2 [v; w]) T, the loss
    and 12-float gradient, Adam with the learning-rate ramp (pose x fraction,
    velocity x (1 - fraction)), then update_vwRT (velocity, then the
    left-multiplicative pose update).  Returns (T, lin, ang, losses)."""
    adam = Adam({"theta": lr[0], "rho": lr[1], "w": lr[2], "v": lr[3]})
    T = np.asarray(T0, np.float64).copy()
    lin, ang = np.asarray(lin, np.float64).copy(), np.asarray(ang, np.float64).copy()
    zero = {k: np.zeros(3) for k in ("theta", "rho", "w", "v")}
    losses = []
    for it in range(iters):
        fr = lr_fraction(it, 0, max_optim_iter)
        vel = np.concatenate([lin, ang])
        El, En = se3_exp(-0.5 * dtau * vel), se3_exp(0.5 * dtau * vel)
        cl = camera(El @ T, W, H, fx, fy, vel=El, delta_time=-0.5 * dtau, lin=lin, ang=ang)
        cn = camera(En @ T, W, H, fx, fy, vel=En, delta_time=0.5 * dtau, lin=lin, ang=ang)
        loss, grads = iteration(scene, cl, cn, sign_Ie)
        losses.append(loss)
        d = adam.step(zero, grads, lr_scale={"theta": fr, "rho": fr, "w": 1 - fr, "v": 1 - fr})
        ang, lin = ang + d["w"], lin + d["v"]
        T = se3_exp(np.concatenate([d["rho"], d["theta"]])) @ T
    return T, lin, ang, losses
