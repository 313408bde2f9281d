"""Front-end of the GPU FP64 finite-difference engine (oracle/fd64.hip).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the pose and
velocity adjoint at full size, never by the product.

`FD64` holds one Gaussian map in double precision on the device and computes
the tracking loss of utils/tracker.py:93-103 (un-squared norm of the
normalised intensity change of utils/render_camera/frame.py:61-94 minus the
event frame) for the two renders of a window, and its central differences in
the 12 parameters (theta, rho, w, v) that gaussian_renderer/__init__.py:
301-308 routes through the rasterizer:

    T_last = Exp(-dtau/2 [v + dv; w + dw]) Exp([drho; dtheta]) T_c
    T_next = Exp(+dtau/2 [v + dv; w + dw]) Exp([drho; dtheta]) T_c

Every discrete decision of the two renders (culls, clamps, which fragment is
blended into which pixel, alpha clamps, binning) is taken at the base point
and frozen for the perturbed evaluations, so the difference quotient
converges to the derivative of the branches taken (the quantity the adjoint
computes) instead of straddling thresholds.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np
import torch

from .tracking_step import se3_exp

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
GREY = (0.2989, 0.5870, 0.1140)
ZNEAR, ZFAR = 0.01, 100.0


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "libfd64.so")
    if not os.path.exists(path):
        subprocess.run(["make", "-s", "-C", _HERE, "libfd64.so"], check=True)
    L = ctypes.CDLL(path)
    vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    L.fd64_last_error.restype = ctypes.c_char_p
    L.fd64_preprocess.argtypes = [i, i, i, vp, vp, vp, vp, vp, vp, vp, i, i, d, d, i, i, i, vp, vp, vp, vp, vp, vp,
                                  vp]
    L.fd64_blend.argtypes = [i, i, i, i, vp, vp, vp, vp, vp, d, i, vp, vp, vp]
    L.fd64_record32.argtypes = [i, i, i, i, vp, vp, ctypes.c_uint32, vp, i, vp, vp, vp]
    _LIB = L
    return L


def _call(fn, *args):
    rc = fn(*args)
    if rc != 0:
        raise RuntimeError(f"{fn.__name__} failed ({rc}): {lib().fd64_last_error().decode()}")


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _h(a):
    a = np.ascontiguousarray(a, np.float64)
    return a, a.ctypes.data_as(ctypes.c_void_p)


def cov6_from_scale_rot(scales, rots, modifier=1.0):
    """3D covariances [P, 6] (00, 01, 02, 11, 12, 22) = R S S^T R^T with R of
    the (w, x, y, z) quaternion as given (gaussian_model.py:52-60)."""
    s = np.asarray(scales, np.float64) * modifier
    w, x, y, z = np.asarray(rots, np.float64).T
    R = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
                  2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
                  2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], 1).reshape(-1, 3, 3)
    M = R * s[:, None, :]
    S = M @ M.transpose(0, 2, 1)
    return np.stack([S[:, 0, 0], S[:, 0, 1], S[:, 0, 2], S[:, 1, 1], S[:, 1, 2], S[:, 2, 2]], 1)


class FD64:
    """means [P,3], scales [P,3], rots [P,4] (w, x, y, z, normalised),
    opac [P], shs [P,M,3] (activated values, any float type); pinhole
    W x H with focal lengths fx, fy and the principal point at the centre."""

    def __init__(self, means, scales, rots, opac, shs, sh_degree, W, H, fx, fy, tile=(16, 16), device="cuda"):
        dev = torch.device(device)
        self.dev = dev
        d64 = lambda a: torch.as_tensor(np.ascontiguousarray(np.asarray(a, np.float64))).to(dev)
        self.P = int(np.asarray(means).shape[0])
        self.D = int(sh_degree)
        shs = np.asarray(shs, np.float64)
        self.M = int(shs.shape[1])
        self.means = d64(np.asarray(means).reshape(-1, 3))
        self.cov6 = d64(cov6_from_scale_rot(np.asarray(scales).reshape(-1, 3), np.asarray(rots).reshape(-1, 4)))
        self.opac = d64(np.asarray(opac).reshape(-1))
        self.shs = d64(shs.reshape(self.P, self.M, 3))
        self.W, self.H = int(W), int(H)
        self.tile = (int(tile[0]), int(tile[1]))
        self.tanfovx = math.tan(0.5 * (2 * math.atan(W / (2 * fx))))
        self.tanfovy = math.tan(0.5 * (2 * math.atan(H / (2 * fy))))
        P = self.P
        f64 = dict(dtype=torch.float64, device=dev)
        self.xy = torch.empty(P, 2, **f64)
        self.conic = torch.empty(P, 4, **f64)
        self.inten = torch.empty(P, **f64)
        self.depth = torch.empty(P, **f64)
        self.rect = torch.empty(P, 4, dtype=torch.int32, device=dev)
        self.gx = (self.W + self.tile[0] - 1) // self.tile[0]
        self.gy = (self.H + self.tile[1] - 1) // self.tile[1]

    def _proj(self):
        tx, ty = self.tanfovx, self.tanfovy
        top, right = ty * ZNEAR, tx * ZNEAR
        Pm = np.zeros((4, 4))
        Pm[0, 0] = 2 * ZNEAR / (2 * right)
        Pm[1, 1] = 2 * ZNEAR / (2 * top)
        Pm[3, 2] = 1.0
        Pm[2, 2] = ZFAR / (ZFAR - ZNEAR)
        Pm[2, 3] = -(ZFAR * ZNEAR) / (ZFAR - ZNEAR)
        return Pm

    def _preprocess(self, T, record, flags):
        """FP64 preprocess at world->camera pose T (4x4)."""
        T = np.asarray(T, np.float64)
        view, vp = _h(T[:3, :4].reshape(-1))
        full, fp = _h((self._proj() @ T).reshape(-1))
        campos, cp = _h(-T[:3, :3].T @ T[:3, 3])
        s = ctypes.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)
        _call(lib().fd64_preprocess, self.P, self.D, self.M, _p(self.means), _p(self.cov6), _p(self.opac),
              _p(self.shs), vp, fp, cp, self.W, self.H, self.tanfovx, self.tanfovy, self.tile[0], self.tile[1],
              int(record), _p(self.xy), _p(self.conic), _p(self.inten), _p(self.depth), _p(self.rect), _p(flags), s)

    def _bin(self, flags):
        """Fragments of every visible Gaussian's tile rect, tile-major and
        front to back (ties in Gaussian order): (ranges [n_tiles, 2], vals)."""
        dev = self.dev
        vis = torch.nonzero((flags & 1) != 0).reshape(-1)
        r = self.rect[vis].long()
        w, h = r[:, 2] - r[:, 0], r[:, 3] - r[:, 1]
        cnt = w * h
        n = int(cnt.sum())
        rep = torch.repeat_interleave(torch.arange(vis.numel(), device=dev), cnt)
        start = torch.cumsum(cnt, 0) - cnt
        loc = torch.arange(n, device=dev) - start[rep]
        tx = r[rep, 0] + loc % w[rep]
        ty = r[rep, 1] + loc // w[rep]
        tile = ty * self.gx + tx
        gid = vis[rep]
        # the rasterizer's key holds the depth as a float32: fragments whose
        # depths round to the same float32 keep Gaussian order
        o1 = torch.sort(self.depth[gid].float(), stable=True).indices
        o2 = torch.sort(tile[o1], stable=True).indices
        order = o1[o2]
        tile_s = tile[order].contiguous()
        vals = gid[order].to(torch.int32).contiguous()
        t = torch.arange(self.gx * self.gy, device=dev)
        lo = torch.searchsorted(tile_s, t, right=False)
        hi = torch.searchsorted(tile_s, t, right=True)
        ranges = torch.stack([lo, hi], 1).to(torch.int32).contiguous()
        return ranges, vals

    def _blend(self, ranges, vals, bits, record):
        img = torch.empty(self.H, self.W, dtype=torch.float64, device=self.dev)
        s = ctypes.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)
        _call(lib().fd64_blend, self.W, self.H, self.tile[0], self.tile[1], _p(ranges), _p(vals), _p(self.xy),
              _p(self.conic), _p(self.inten), 0.0, int(record), _p(bits), _p(img), s)
        return img

    def record(self, T):
        """Grey render at pose T with every decision taken there; returns
        (image [H, W] float64, frozen state for `frozen`)."""
        flags = torch.zeros(self.P, dtype=torch.int32, device=self.dev)
        self._preprocess(T, True, flags)
        ranges, vals = self._bin(flags)
        bits = torch.zeros(max(1, vals.numel()) * 8, dtype=torch.int64, device=self.dev)
        img = self._blend(ranges, vals, bits, True)
        return img, dict(flags=flags, ranges=ranges, vals=vals, bits=bits)

    def record_from_forward(self, T, ranges, vals, n, splats, S, n_contrib, gid_mask=0xFFFFFFFF):
        """Frozen decisions taken from an FP32 forward render at pose T: its
        binning (ranges [n_tiles*2], vals [>= n]), splat records (S float4
        per Gaussian) and n_contrib [H, W] give which fragment is blended into
        which pixel and which alpha is clamped, as that forward decided them
        (oracle/fd64.hip k_fd_record32; the Gaussian of a fragment is
        vals & gid_mask); the per-Gaussian decisions are taken in FP64 at T.
        The frozen FP64 function is then the one the FP32 forward's adjoint
        differentiates."""
        flags = torch.zeros(self.P, dtype=torch.int32, device=self.dev)
        self._preprocess(T, True, flags)
        ranges = ranges.reshape(-1).to(torch.int32).contiguous()
        vals = vals.reshape(-1)[:max(int(n), 1)].to(torch.int32).contiguous()
        bits = torch.zeros(max(1, int(n)) * 8, dtype=torch.int64, device=self.dev)
        s = ctypes.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)
        _call(lib().fd64_record32, self.W, self.H, self.tile[0], self.tile[1], _p(ranges), _p(vals),
              int(gid_mask), _p(splats.contiguous()), int(S), _p(n_contrib.reshape(-1).to(torch.int32).contiguous()),
              _p(bits), s)
        if gid_mask != 0xFFFFFFFF:   # the frozen blend reads plain Gaussian indices
            vals = (vals.long() & gid_mask).to(torch.int32).contiguous()
        return dict(flags=flags, ranges=ranges, vals=vals, bits=bits)

    def frozen(self, T, st):
        """Grey render at pose T with the decisions of `st`."""
        self._preprocess(T, False, st["flags"])
        return self._blend(st["ranges"], st["vals"], st["bits"], False)

    @staticmethod
    def poses(T_c, lin, ang, dtau, delta12=None):
        """(T_last, T_next) for the 12-vector (dtheta, drho, dw, dv)."""
        d = np.zeros(12) if delta12 is None else np.asarray(delta12, np.float64)
        D = se3_exp(np.concatenate([d[3:6], d[0:3]]))
        vel = np.concatenate([np.asarray(lin, np.float64) + d[9:12], np.asarray(ang, np.float64) + d[6:9]])
        Tc = np.asarray(T_c, np.float64)
        return se3_exp(-0.5 * dtau * vel) @ D @ Tc, se3_exp(0.5 * dtau * vel) @ D @ Tc

    @staticmethod
    def loss(I_last, I_next, events):
        d = I_next - I_last
        return float(torch.linalg.norm(d / torch.linalg.norm(d) - events))

    def loss_and_fd(self, T_c, lin, ang, dtau, events, h=1e-6, decisions=None):
        """Loss at the base point and its central differences (step h) in
        (theta, rho, w, v), decisions frozen at the base point: taken here in
        FP64 (record), or given as the (last, next) states of
        record_from_forward."""
        e = torch.as_tensor(np.asarray(events, np.float64) if not torch.is_tensor(events) else events,
                            dtype=torch.float64).to(self.dev).reshape(self.H, self.W)
        Tl, Tn = self.poses(T_c, lin, ang, dtau)
        if decisions is None:
            Il, sl = self.record(Tl)
            In, sn = self.record(Tn)
        else:
            sl, sn = decisions
            Il, In = self.frozen(Tl, sl), self.frozen(Tn, sn)
        L0 = self.loss(Il, In, e)
        fd = np.zeros(12)
        for k in range(12):
            vals = []
            for sgn in (1.0, -1.0):
                dk = np.zeros(12)
                dk[k] = sgn * h
                Tl, Tn = self.poses(T_c, lin, ang, dtau, dk)
                vals.append(self.loss(self.frozen(Tl, sl), self.frozen(Tn, sn), e))
            fd[k] = (vals[0] - vals[1]) / (2 * h)
        return L0, fd
