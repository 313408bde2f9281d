"""CPU oracle of the event-frame construction (numpy).

TEST INFRASTRUCTURE ONLY.  Restates utils/event_camera/event.py:116-128 and
utils/tracker.py:78-91 of the reference.  OpenCV (opencv_python 4.8.1.78,
requirements.txt:5) is not installed here, so its functions are restated from
their published algorithms:
  cv2.undistort     initUndistortRectifyMap (new camera matrix = K) with the
                    map quantised to 1/32 pixel (INTER_BITS = 5) and bilinear
                    remap with the 32x32 float weight table, BORDER_CONSTANT 0
  cv2.GaussianBlur  separable kernel of getGaussianKernel (sigma 0 ->
                    0.3((k-1)/2 - 1) + 0.8), BORDER_REPLICATE
  cv2.normalize     NORM_L2 to unit norm
  cv2.resize        INTER_NEAREST, src = floor(dst * (src / dst))
Bit-exact agreement with OpenCV itself is unpinned (no OpenCV here); the
integer event counts are exact (checked against a per-event loop).
"""
from __future__ import annotations

import numpy as np

INTER_BITS = 5
INTER_TAB = 1 << INTER_BITS


def scatter(x, y, p, W, H):
    """count[y, x] += +1 / -1 (event.py:119-120)."""
    c = np.zeros(H * W, np.int64)
    x = np.asarray(x, np.int64)
    y = np.asarray(y, np.int64)
    ok = (x >= 0) & (x < W) & (y >= 0) & (y < H)
    np.add.at(c, (y[ok] * W + x[ok]), np.where(np.asarray(p)[ok] != 0, 1, -1))
    return c.reshape(H, W)


def undistort_map(W, H, K, dist):
    """1/32-pixel fixed-point source coordinates of every destination pixel."""
    K = np.asarray(K, np.float64)
    d = np.zeros(5)
    d[:len(dist)] = np.asarray(dist, np.float64)[:5]
    k1, k2, p1, p2, k3 = d
    fx, fy, cx, cy = K[0, 0], K[1, 1], K[0, 2], K[1, 2]
    j, i = np.meshgrid(np.arange(W, dtype=np.float64), np.arange(H, dtype=np.float64))
    ifx, ify = 1.0 / fx, 1.0 / fy
    x = j * ifx - cx * ifx
    y = i * ify - cy * ify
    x2, y2 = x * x, y * y
    r2 = x2 + y2
    _2xy = 2 * x * y
    kr = 1 + ((k3 * r2 + k2) * r2 + k1) * r2
    xd = x * kr + p1 * _2xy + p2 * (r2 + 2 * x2)
    yd = y * kr + p1 * (r2 + 2 * y2) + p2 * _2xy
    u = fx * xd + cx
    v = fy * yd + cy
    iu = np.rint(u * INTER_TAB).astype(np.int64)
    iv = np.rint(v * INTER_TAB).astype(np.int64)
    return iu, iv


def remap_bilinear(src, iu, iv):
    H, W = src.shape
    sx, sy = iu >> INTER_BITS, iv >> INTER_BITS
    ax = (iu & (INTER_TAB - 1)).astype(np.float32) * np.float32(1.0 / INTER_TAB)
    ay = (iv & (INTER_TAB - 1)).astype(np.float32) * np.float32(1.0 / INTER_TAB)
    one = np.float32(1)
    w00, w01 = (one - ay) * (one - ax), (one - ay) * ax
    w10, w11 = ay * (one - ax), ay * ax

    def at(yy, xx):
        ok = (xx >= 0) & (xx < W) & (yy >= 0) & (yy < H)
        out = np.zeros(xx.shape, np.float32)
        out[ok] = src[yy[ok], xx[ok]]
        return out
    return (at(sy, sx) * w00 + at(sy, sx + 1) * w01 + at(sy + 1, sx) * w10 + at(sy + 1, sx + 1) * w11).astype(
        np.float32)


def undistort(img, K, dist):
    H, W = img.shape
    iu, iv = undistort_map(W, H, K, dist)
    return remap_bilinear(np.asarray(img, np.float32), iu, iv)


def gaussian_kernel(ksize, sigma):
    if sigma <= 0:
        sigma = ((ksize - 1) * 0.5 - 1) * 0.3 + 0.8
    x = np.arange(ksize) - (ksize - 1) * 0.5
    t = np.exp(-0.5 / (sigma * sigma) * x * x).astype(np.float32)
    s = float(np.sum(t.astype(np.float64)))
    return (t.astype(np.float64) / s).astype(np.float32)


def blur(img, ksize=9, sigma=0.0):
    k = gaussian_kernel(ksize, sigma)
    r = ksize // 2
    H, W = img.shape
    xs = np.clip(np.arange(W)[:, None] + np.arange(-r, r + 1)[None], 0, W - 1)
    rows = np.zeros_like(img, dtype=np.float32)
    for t in range(ksize):
        rows = rows + k[t] * img[:, xs[:, t]]
    ys = np.clip(np.arange(H)[:, None] + np.arange(-r, r + 1)[None], 0, H - 1)
    out = k[r] * rows
    for t in range(1, r + 1):
        out = out + k[r + t] * (rows[ys[:, r + t], :] + rows[ys[:, r - t], :])
    return out.astype(np.float32)


def normalize_l2(img):
    n = np.sqrt(np.sum(img.astype(np.float64) ** 2))
    s = np.float32(1.0 / n) if n > 0 else np.float32(0)
    return (img * s).astype(np.float32)


def event_frame(x, y, p, W, H, K, dist, ksize=9):
    c = scatter(x, y, p, W, H).astype(np.float32)
    s = normalize_l2(blur(undistort(c, K, dist), ksize))
    return s, np.abs(s)


def resize_nearest(img, dw, dh):
    sh, sw = img.shape
    ifx = 1.0 / (dw / sw)
    ify = 1.0 / (dh / sh)
    xs = np.minimum(np.floor(np.arange(dw) * ifx).astype(np.int64), sw - 1)
    ys = np.minimum(np.floor(np.arange(dh) * ify).astype(np.int64), sh - 1)
    return img[ys][:, xs]
