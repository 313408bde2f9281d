"""Independent FP64 brute-force renderer (no binning, no sort machinery).

TEST INFRASTRUCTURE ONLY.  Used to pin the C++ oracle: every pixel takes every
visible Gaussian whose tile rectangle covers the pixel's tile, orders them by
(depth, index) and blends front to back with the published 3DGS rules
(alpha = min(0.99, o exp(power)), skip alpha < 1/255 or power > 0, stop before
T drops below 1e-4).  Written with numpy matrix algebra rather than the
oracle's scalar expressions, and evaluates SH with the polynomial form of
gaussian_splatting/utils/sh_utils.py:55-118, so a dropped term, a transposed
matrix or a wrong sign in the oracle shows up as a mismatch.
"""
from __future__ import annotations

import numpy as np

C0 = 0.28209479177387814
C1 = 0.4886025119029199
C2 = [1.0925484305920792, -1.0925484305920792, 0.31539156525252005, -1.0925484305920792, 0.5462742152960396]
C3 = [-0.5900435899266435, 2.890611442640554, -0.4570457994644658, 0.3731763325901154, -0.4570457994644658,
      1.445305721320277, -0.5900435899266435]


def eval_sh(deg, sh, dirs):
    """sh: [P, K, 3] (coefficient-major), dirs: [P, 3] unit -> [P, 3]."""
    x, y, z = dirs[:, 0:1], dirs[:, 1:2], dirs[:, 2:3]
    r = C0 * sh[:, 0]
    if deg > 0:
        r = r - C1 * y * sh[:, 1] + C1 * z * sh[:, 2] - C1 * x * sh[:, 3]
    if deg > 1:
        xx, yy, zz, xy, yz, xz = x * x, y * y, z * z, x * y, y * z, x * z
        r = (r + C2[0] * xy * sh[:, 4] + C2[1] * yz * sh[:, 5] + C2[2] * (2.0 * zz - xx - yy) * sh[:, 6]
             + C2[3] * xz * sh[:, 7] + C2[4] * (xx - yy) * sh[:, 8])
    if deg > 2:
        r = (r + C3[0] * y * (3 * xx - yy) * sh[:, 9] + C3[1] * xy * z * sh[:, 10]
             + C3[2] * y * (4 * zz - xx - yy) * sh[:, 11] + C3[3] * z * (2 * zz - 3 * xx - 3 * yy) * sh[:, 12]
             + C3[4] * x * (4 * zz - xx - yy) * sh[:, 13] + C3[5] * z * (xx - yy) * sh[:, 14]
             + C3[6] * x * (xx - 3 * yy) * sh[:, 15])
    return r


def quat_to_rot(q):
    # the rasterizer uses the quaternion as given (the caller normalises it,
    # gaussian_model.py:59); no renormalisation here either
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    R = np.empty((q.shape[0], 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z); R[:, 0, 1] = 2 * (x * y - w * z); R[:, 0, 2] = 2 * (x * z + w * y)
    R[:, 1, 0] = 2 * (x * y + w * z); R[:, 1, 1] = 1 - 2 * (x * x + z * z); R[:, 1, 2] = 2 * (y * z - w * x)
    R[:, 2, 0] = 2 * (x * z - w * y); R[:, 2, 1] = 2 * (y * z + w * x); R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def project(scene, cam):
    """Vectorised FP64 preprocess. Returns per-Gaussian arrays."""
    means = np.asarray(scene.means, np.float64)
    P = means.shape[0]
    Vw = np.asarray(cam.view, np.float64).reshape(4, 4).T      # world->camera, column-vector form
    Pf = np.asarray(cam.proj, np.float64).reshape(4, 4).T
    hom = np.concatenate([means, np.ones((P, 1))], 1)
    pv = hom @ Vw.T
    ph = hom @ Pf.T
    ndc = ph[:, :2] / (ph[:, 3:4] + 1e-7)
    Wr = Vw[:3, :3]
    if scene.cov_pre is not None:
        c = np.asarray(scene.cov_pre, np.float64)
        Sig = np.stack([np.stack([c[:, 0], c[:, 1], c[:, 2]], -1), np.stack([c[:, 1], c[:, 3], c[:, 4]], -1),
                        np.stack([c[:, 2], c[:, 4], c[:, 5]], -1)], -2)
    else:
        R = quat_to_rot(np.asarray(scene.rots, np.float64))
        S = np.asarray(scene.scales, np.float64) * scene.scale_modifier
        Mm = R * S[:, None, :]
        Sig = Mm @ np.transpose(Mm, (0, 2, 1))
    fx = cam.W / (2 * cam.tanfovx)
    fy = cam.H / (2 * cam.tanfovy)
    z = pv[:, 2]
    u = np.clip(pv[:, 0] / z, -1.3 * cam.tanfovx, 1.3 * cam.tanfovx)
    v = np.clip(pv[:, 1] / z, -1.3 * cam.tanfovy, 1.3 * cam.tanfovy)
    J = np.zeros((P, 2, 3))
    J[:, 0, 0] = fx / z
    J[:, 0, 2] = -fx * u / z
    J[:, 1, 1] = fy / z
    J[:, 1, 2] = -fy * v / z
    T = J @ Wr
    cov = T @ Sig @ np.transpose(T, (0, 2, 1)) + 0.3 * np.eye(2)
    det = cov[:, 0, 0] * cov[:, 1, 1] - cov[:, 0, 1] ** 2
    conic = np.stack([cov[:, 1, 1], -cov[:, 0, 1], cov[:, 0, 0]], -1) / det[:, None]
    mid = 0.5 * (cov[:, 0, 0] + cov[:, 1, 1])
    lam = mid + np.sqrt(np.maximum(0.1, mid * mid - det))
    radius = np.ceil(3 * np.sqrt(lam))
    px = ((ndc[:, 0] + 1) * cam.W - 1) * 0.5
    py = ((ndc[:, 1] + 1) * cam.H - 1) * 0.5
    gx = (cam.W + cam.tile_w - 1) // cam.tile_w
    gy = (cam.H + cam.tile_h - 1) // cam.tile_h
    x0 = np.clip((px - radius) / cam.tile_w, 0, gx).astype(np.int64)
    y0 = np.clip((py - radius) / cam.tile_h, 0, gy).astype(np.int64)
    x1 = np.clip((px + radius + cam.tile_w - 1) / cam.tile_w, 0, gx).astype(np.int64)
    y1 = np.clip((py + radius + cam.tile_h - 1) / cam.tile_h, 0, gy).astype(np.int64)
    visible = (z > 0.2) & (det != 0) & ((x1 - x0) * (y1 - y0) > 0)
    if scene.colors_pre is not None:
        rgb = np.asarray(scene.colors_pre, np.float64)
    else:
        d = means - np.asarray(cam.campos, np.float64)[None]
        d = d / np.linalg.norm(d, axis=1, keepdims=True)
        rgb = np.maximum(eval_sh(scene.sh_degree, np.asarray(scene.shs, np.float64), d) + 0.5, 0.0)
    return dict(visible=visible, depth=z, xy=np.stack([px, py], -1), conic=conic,
                opac=np.asarray(scene.opac, np.float64).reshape(-1), rgb=rgb, radius=radius.astype(np.int64),
                rect=np.stack([x0, y0, x1, y1], -1), tiles=np.where(visible, (x1 - x0) * (y1 - y0), 0))


def render(scene, cam, bg=None):
    g = project(scene, cam)
    W, H = cam.W, cam.H
    bg = np.zeros(3) if bg is None else np.asarray(bg, np.float64)
    color = np.zeros((3, H, W))
    depth = np.zeros((H, W))
    opac = np.zeros((H, W))
    finalT = np.ones((H, W))
    vis = np.nonzero(g["visible"])[0]
    order = vis[np.lexsort((vis, g["depth"][vis]))]
    rect = g["rect"][order]
    for py in range(H):
        ty = py // cam.tile_h
        for px in range(W):
            tx = px // cam.tile_w
            sel = order[(rect[:, 0] <= tx) & (tx < rect[:, 2]) & (rect[:, 1] <= ty) & (ty < rect[:, 3])]
            if sel.size == 0:
                color[:, py, px] = bg
                continue
            d = g["xy"][sel] - np.array([px, py], np.float64)
            c = g["conic"][sel]
            power = -0.5 * (c[:, 0] * d[:, 0] ** 2 + c[:, 2] * d[:, 1] ** 2) - c[:, 1] * d[:, 0] * d[:, 1]
            alpha = np.minimum(0.99, g["opac"][sel] * np.exp(power))
            ok = (power <= 0) & (alpha >= 1.0 / 255.0)
            a = np.where(ok, alpha, 0.0)
            Tafter = np.cumprod(1 - a)
            stop = np.nonzero(ok & (Tafter < 1e-4))[0]
            n = stop[0] if stop.size else sel.size
            a = a[:n]
            Tb = np.concatenate([[1.0], Tafter[:n - 1]]) if n > 0 else np.zeros(0)
            w = a * Tb
            color[:, py, px] = (g["rgb"][sel[:n]] * w[:, None]).sum(0)
            depth[py, px] = (g["depth"][sel[:n]] * w).sum()
            opac[py, px] = w.sum()
            T = Tafter[n - 1] if n > 0 else 1.0
            finalT[py, px] = T
            color[:, py, px] += T * bg
    return dict(color=color, depth=depth, opacity=opac, final_T=finalT, proj=g)
