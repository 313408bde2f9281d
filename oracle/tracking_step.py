"""CPU restatement of one fine-stage tracking iteration (oracle side).

TEST INFRASTRUCTURE ONLY (used by tests/ and as bench.py's cpu_baseline).
Follows utils/render_camera/frame.py:61-94 (grey intensity, difference, L2
normalisation) and utils/tracker.py:93-103,212-222 (un-squared norm loss,
backward, Adam, left-multiplicative pose update), with the rasterizer
replaced by the C++ oracle.
"""
from __future__ import annotations

import numpy as np

from . import oracle as orc

GREY = np.array([0.2989, 0.5870, 0.1140])


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
