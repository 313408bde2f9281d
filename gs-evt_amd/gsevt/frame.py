"""Rendered and event frames of one tracking window.

RenderFrame mirrors utils/render_camera/frame.py:10-94: two renders at the
last / next poses of the window, grey intensity with weights
(0.2989, 0.5870, 0.1140), difference, L2 normalisation and absolute value.

EventFrame mirrors utils/event_camera/event.py:103-128 but builds the frame on
the GPU: int32 atomic scatter of the polarities, undistortion remap,
9x9 Gaussian blur, unit L2 normalisation, |.| - see csrc/events.hip.
image_pyramid mirrors utils/tracker.py:78-91 (nearest resize, levels 0..2).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from ._lib import call, ptr
from .camera import Camera, focal2fov
from .renderer import render2

GREY = (0.2989, 0.5870, 0.1140)


class RenderFrame:
    def __init__(self, viewpoint, gaussians, pipeline, background, pyramid_level, tile=(16, 16)):
        self.viewpoint = viewpoint
        self.gaussians = gaussians
        self.pipeline = pipeline
        self.background = background
        self.tile = tile
        self.sign_delta_Ir, self.unsign_delta_Ir = self.get_delta_Ir(pyramid_level)

    @staticmethod
    def get_intensity_frame(color_frame):
        w = torch.tensor(GREY).view(1, 3, 1, 1).to(color_frame.device)
        return (color_frame * w).sum(dim=1)

    def get_delta_Ir(self, pyramid_level=0):
        vp = self.viewpoint
        assert vp.delta_tau != 0, "delta_tau should not be zero"
        scale = 0.5 ** pyramid_level
        new_w = int(vp.image_width * scale)
        new_h = int(vp.image_height * scale)
        last_pose = vp.last_vel_transform @ vp.curr_pose
        next_pose = vp.next_vel_transform @ vp.curr_pose
        fovx = focal2fov(vp.fx * scale, new_w)
        fovy = focal2fov(vp.fy * scale, new_h)
        last = Camera(last_pose[:3, :3], last_pose[:3, 3], vp.angular_vel, vp.linear_vel, fovx, fovy, new_w, new_h,
                      delta_tau=vp.delta_tau, device=vp.device)
        nxt = Camera(next_pose[:3, :3], next_pose[:3, 3], vp.angular_vel, vp.linear_vel, fovx, fovy, new_w, new_h,
                     delta_tau=vp.delta_tau, device=vp.device)
        last_pkg, next_pkg = render2(last, vp, nxt, self.gaussians, self.background, tile=self.tile)
        last_I = self.get_intensity_frame(last_pkg["render"])
        next_I = self.get_intensity_frame(next_pkg["render"])
        sign = next_I - last_I
        l2 = torch.norm(sign, p=2)
        normalized = sign / l2
        return normalized, torch.abs(normalized)


class EventArray:
    """utils/event_camera/event.py:82-100 over numpy columns (ts in us)."""

    def __init__(self, ts, x, y, p):
        self.ts = np.asarray(ts, np.int64)
        self.x = np.asarray(x, np.int32)
        self.y = np.asarray(y, np.int32)
        self.p = np.asarray(p, np.int32)

    def size(self):
        return int(self.ts.shape[0])

    def duration(self):
        if self.size() > 0:
            return (int(self.ts[-1]) - int(self.ts[0])) / 1e6
        return 0

    def time(self):
        return (int(self.ts[0]) + (int(self.ts[-1]) - int(self.ts[0])) / 2) / 1e6


_MAP_CACHE = {}


def undistort_map(W, H, K, dist, device):
    key = (W, H, tuple(np.asarray(K, np.float64).reshape(-1)), tuple(np.asarray(dist, np.float64).reshape(-1)),
           str(device))
    m = _MAP_CACHE.get(key)
    if m is None:
        Kd = torch.tensor(np.asarray(K, np.float64).reshape(-1), device=device)
        dd = torch.zeros(5, dtype=torch.float64, device=device)
        d = np.asarray(dist, np.float64).reshape(-1)
        dd[:d.shape[0]] = torch.from_numpy(d[:5])
        m = torch.empty(H * W * 2, dtype=torch.int32, device=device)
        call("gsevt_undistort_map", W, H, ptr(Kd), ptr(dd), ptr(m), _lib.stream_of(device))
        _MAP_CACHE[key] = (m, Kd, dd)
    return m[0] if isinstance(m, tuple) else m


def event_frame(x, y, p, W, H, K, dist, ksize=9, device="cuda", outside="drop"):
    """Signed and absolute event frames [1,H,W] on the device.

    outside: what happens to events whose (x, y) lies outside the W x H frame.
    "drop" skips them (the scatter kernel's rule); "index" follows the numpy
    indexing of event.py:119-120: coordinates in [-W, 0) and [-H, 0) wrap
    around, anything else outside raises IndexError (one device-to-host read)."""
    if outside not in ("drop", "index"):
        raise ValueError(f"outside must be 'drop' or 'index', not {outside!r}")
    dev = torch.device(device)
    st = _lib.stream_of(dev)
    xt = torch.as_tensor(x, dtype=torch.int32).to(dev)
    yt = torch.as_tensor(y, dtype=torch.int32).to(dev)
    pt = torch.as_tensor(p, dtype=torch.int32).to(dev)
    if outside == "index" and xt.numel() > 0:
        bad = (xt >= W) | (xt < -W) | (yt >= H) | (yt < -H)
        if bool(bad.any()):
            i = int(torch.nonzero(bad)[0, 0])
            raise IndexError(f"event {i} at (x={int(xt[i])}, y={int(yt[i])}) is outside the {W}x{H} frame")
        xt = torch.where(xt < 0, xt + W, xt)
        yt = torch.where(yt < 0, yt + H, yt)
    count = torch.zeros(H, W, dtype=torch.int32, device=dev)
    call("gsevt_event_scatter", int(xt.shape[0]), ptr(xt), ptr(yt), ptr(pt), W, H, ptr(count), None, st)
    m = undistort_map(W, H, K, dist, dev)
    scratch = torch.empty(int(_lib.load().gsevt_event_frame_scratch_bytes(W, H)), dtype=torch.uint8, device=dev)
    sign = torch.empty(1, H, W, dtype=torch.float32, device=dev)
    absf = torch.empty(1, H, W, dtype=torch.float32, device=dev)
    call("gsevt_event_frame", W, H, ptr(count), ptr(m), ksize, 0.0, ptr(scratch), ptr(sign), ptr(absf), st)
    return sign, absf, count


class EventFrame:
    def __init__(self, img_width, img_height, intrinsic, distortion_factors, gaussian_kernel_size,
                 event_array: EventArray, device="cuda"):
        self.sign_delta_Ie, self.unsign_delta_Ie, self.count = event_frame(
            event_array.x, event_array.y, event_array.p, img_width, img_height, intrinsic, distortion_factors,
            gaussian_kernel_size, device, outside="index")


def image_pyramid(image, levels=3):
    """utils/tracker.py:78-91: nearest resize to (int(w 2^-l), int(h 2^-l))."""
    _, h, w = image.shape
    dev = image.device
    out = [image.contiguous()]
    for lvl in range(1, levels):
        s = 0.5 ** lvl
        dw, dh = int(w * s), int(h * s)
        dst = torch.empty(1, dh, dw, dtype=torch.float32, device=dev)
        call("gsevt_resize_nearest", w, h, ptr(out[0]), dw, dh, ptr(dst), _lib.stream_of(dev))
        out.append(dst)
    return out
