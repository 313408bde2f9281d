"""Shared test helpers: numpy cameras in the rasterizer's storage convention,
scene conversion for the oracle, and the SE(3) exponential in float64."""
from __future__ import annotations

import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "gs-evt_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

from oracle import oracle as orc  # noqa: E402
from gsevt import synthetic  # noqa: E402


def skew(v):
    return np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]], dtype=np.float64)


def se3_exp(xi):
    """float64 SE(3) exponential, twist [rho; theta] (utils/pose.py:78-91)."""
    rho, th = np.asarray(xi[:3], np.float64), np.asarray(xi[3:], np.float64)
    a = np.linalg.norm(th)
    W = skew(th)
    if a < 1e-12:
        R = np.eye(3) + W + 0.5 * W @ W
        V = np.eye(3) + 0.5 * W + W @ W / 6.0
    else:
        R = np.eye(3) + math.sin(a) / a * W + (1 - math.cos(a)) / a ** 2 * W @ W
        V = np.eye(3) + (1 - math.cos(a)) / a ** 2 * W + (a - math.sin(a)) / a ** 3 * W @ W
    T = np.eye(4)
    T[:3, :3] = R
    T[:3, 3] = V @ rho
    return T


def proj_matrix(fovx, fovy, znear=0.01, zfar=100.0):
    ty, tx = math.tan(fovy / 2), math.tan(fovx / 2)
    top, right = ty * znear, tx * znear
    P = np.zeros((4, 4))
    P[0, 0] = 2 * znear / (2 * right)
    P[1, 1] = 2 * znear / (2 * top)
    P[3, 2] = 1.0
    P[2, 2] = zfar / (zfar - znear)
    P[2, 3] = -(zfar * znear) / (zfar - znear)
    return P


def make_cam(Tw2c, W, H, fx, fy, dtype=np.float32, tile=(16, 16), bg=None, vel=None, delta_time=0.0,
             lin=None, ang=None):
    """Cam in settings storage (transposed 4x4 matrices) from a world->camera
    pose.  The matrices are rounded to `dtype` first, as the rasterizer
    receives them."""
    fovx = 2 * math.atan(W / (2 * fx))
    fovy = 2 * math.atan(H / (2 * fy))
    Rt = np.asarray(Tw2c, np.float64)
    P = proj_matrix(fovx, fovy)
    view = Rt.T.reshape(-1)
    praw = P.T.reshape(-1)
    full = (P @ Rt).T.reshape(-1)
    campos = -Rt[:3, :3].T @ Rt[:3, 3]
    c = lambda a: np.asarray(a, dtype)
    return orc.Cam(view=c(view), proj=c(full), projraw=c(praw), campos=c(campos), W=W, H=H,
                   tanfovx=float(dtype(math.tan(fovx * 0.5))), tanfovy=float(dtype(math.tan(fovy * 0.5))),
                   bg=c(np.zeros(3) if bg is None else bg), tile_w=tile[0], tile_h=tile[1],
                   vel=None if vel is None else np.asarray(vel, np.float64).T.reshape(-1),
                   delta_time=delta_time, lin=lin, ang=ang)


def oracle_scene(raw: synthetic.RawScene, dtype=np.float32):
    a = raw.activated()
    return orc.Scene(means=a["means"].astype(dtype), scales=a["scales"].astype(dtype),
                     rots=a["rots"].astype(dtype), opac=a["opac"].astype(dtype), shs=a["shs"].astype(dtype),
                     sh_degree=raw.sh_degree)


def scene_astype(s: "orc.Scene", dtype):
    cv = lambda a: None if a is None else np.asarray(a).astype(dtype)
    return orc.Scene(means=cv(s.means), scales=cv(s.scales), rots=cv(s.rots), opac=cv(s.opac), shs=cv(s.shs),
                     sh_degree=s.sh_degree, scale_modifier=s.scale_modifier, colors_pre=cv(s.colors_pre),
                     cov_pre=cv(s.cov_pre))


def cam_astype(c: "orc.Cam", dtype):
    import dataclasses
    cv = lambda a: None if a is None else np.asarray(a).astype(dtype)
    return dataclasses.replace(c, view=cv(c.view), proj=cv(c.proj), projraw=cv(c.projraw), campos=cv(c.campos),
                               bg=cv(c.bg))


def gpu_cam(c: "orc.Cam", device):
    """raster.Camera on the device from an oracle Cam (same float32 values)."""
    import torch
    from gsevt import raster
    t = lambda a: torch.as_tensor(np.asarray(a, np.float32)).to(device)
    return raster.Camera(viewmatrix=t(c.view).reshape(4, 4), projmatrix=t(c.proj).reshape(4, 4),
                         projmatrix_raw=t(c.projraw).reshape(4, 4), campos=t(c.campos), tanfovx=c.tanfovx,
                         tanfovy=c.tanfovy, width=c.W, height=c.H, bg=t(c.bg), tile_w=c.tile_w, tile_h=c.tile_h)


def gpu_map(s: "orc.Scene", device):
    import torch
    from gsevt import raster
    t = lambda a: None if a is None else torch.as_tensor(np.ascontiguousarray(a, np.float32)).to(device)
    return raster.GaussianMap(means3D=t(s.means), opacities=t(s.opac).reshape(-1, 1), shs=t(s.shs),
                              colors_precomp=t(s.colors_pre), scales=t(s.scales), rotations=t(s.rots),
                              cov3D_precomp=t(s.cov_pre), sh_degree=s.sh_degree,
                              scale_modifier=s.scale_modifier)


def splat_fields(splats, P, channels=3):
    """Unpack the device splat records into oracle-comparable arrays."""
    S = 3 if channels == 3 else 2
    r = np.asarray(splats, np.float32).reshape(-1, S, 4)[:P]
    out = dict(xy=r[:, 0, :2].reshape(-1), conic=r[:, 0, 2:4], conic_c=r[:, 1, 0], opac=r[:, 1, 1],
               depth=r[:, 1, 2])
    if channels == 3:
        out["rgb"] = np.stack([r[:, 1, 3], r[:, 2, 0], r[:, 2, 1]], -1).reshape(-1)
    else:
        out["intensity"] = r[:, 1, 3]
    return out


def workload_cases():
    """(name, scene, cam) cases for parity tests, small enough for the oracle."""
    cases = []
    raw = synthetic.small_test_scene(P=3000, seed=11, sh_degree=3, spread=0.8, log_scale_mean=-3.5)
    T = np.eye(4)
    cases.append(("sh3_small_160x120", oracle_scene(raw), make_cam(T, 160, 120, 140.0, 140.0)))
    raw = synthetic.cube_scene(P=20000)
    cases.append(("cube20k_sh0_346x260", oracle_scene(raw),
                  make_cam(se3_exp([0.01, -0.02, 0.0, 0.005, -0.01, 0.003]), 346, 260, 300.0, 300.0)))
    raw = synthetic.patch_scene(P=30000, sh_degree=3)
    cases.append(("patch30k_sh3_320x240_tile8x32", oracle_scene(raw),
                  make_cam(np.eye(4), 320, 240, 275.0, 275.0, tile=(8, 32))))
    return cases


def device_cam_records(tr):
    """Oracle Cams of a BatchedTracker's camera records (the exact float32
    matrices the device computed for each render)."""
    raw = tr.cams.cpu().numpy().view(np.uint8).reshape(tr.R, -1)
    out = []
    for r in range(tr.R):
        fl = raw[r, :53 * 4].view(np.float32)
        out.append(orc.Cam(view=fl[0:16].copy(), proj=fl[16:32].copy(), projraw=fl[32:48].copy(),
                           campos=fl[48:51].copy(), W=tr.W, H=tr.H, tanfovx=float(fl[51]), tanfovy=float(fl[52]),
                           bg=np.zeros(3, np.float32), tile_w=tr.tile[0], tile_h=tr.tile[1]))
    return out


def oracle_scene_from_model(gm):
    """Oracle Scene of a GaussianModel: the activated tensors the device used."""
    c = lambda t: np.ascontiguousarray(t.detach().float().cpu().numpy())
    return orc.Scene(means=c(gm.get_xyz), scales=c(gm.get_scaling), rots=c(gm.get_rotation),
                     opac=c(gm.get_opacity).reshape(-1), shs=c(gm.get_features), sh_degree=gm.active_sh_degree)


def check_culled_binning(kept_gids, kept_ranges, b, pre, cam):
    """A tile-culled fragment list against the oracle's full binning `b`:
    per tile, the kept Gaussians are the oracle's in the oracle's order with
    some removed; every fragment whose tile the Gaussian may reach in FP64
    (min q over the tile <= 2 ln(255 o), with 1e-5 relative slack) is kept;
    every fragment far outside the cull's margins (min q above
    2 L' + 1e-4 mag + 1e-3, the accept bound of the device's rectangle test,
    by 1 % and 0.01, or opacity below 1/255) is dropped.  The emission's band
    test widens the x-interval by 1e-3 + 1e-3 sqrt(2 L' / a), which moves q by
    at most 0.04, inside that slack.  Returns (kept, slots)."""
    gx, gy = orc.grid(cam)
    n_tiles = gx * gy
    tile = (b["keys"] >> np.uint64(32)).astype(np.int64)
    P = pre["radii"].shape[0]
    full = tile * P + b["vals"].astype(np.int64)
    kr = np.asarray(kept_ranges, np.int64).reshape(n_tiles, 2)
    ne = kr[:, 1] > kr[:, 0]
    assert np.all(kr[~ne] == 0), "an empty tile's range is not (0, 0)"
    if ne.any():   # the non-empty ranges tile the kept list in tile order
        assert kr[ne][0, 0] == 0 and kr[ne][-1, 1] == kept_gids.shape[0]
        assert np.all(kr[ne][1:, 0] == kr[ne][:-1, 1])
    else:
        assert kept_gids.shape[0] == 0
    kt = np.repeat(np.arange(n_tiles), kr[:, 1] - kr[:, 0])
    kept = kt * P + kept_gids.astype(np.int64)
    keep_mask = np.isin(full, kept)
    np.testing.assert_array_equal(full[keep_mask], kept)   # ordered subsequence, nothing extra
    qmin, L, mag = orc.tile_cull_bounds(pre, b, cam)
    must = qmin <= 2.0 * L * (1 + 1e-5) + 1e-5
    Lp = np.maximum(L, 0.0) * 1.0001 + 1e-4
    op = pre["conic_opac"].reshape(-1, 4)[b["vals"].astype(np.int64), 3].astype(np.float64)
    far = (qmin > (2.0 * Lp + 1e-4 * mag + 1e-3) * 1.01 + 0.01) | (op * 255.0 < 0.999 * (1 - 1e-6))
    assert np.all(keep_mask[must]), f"{int((must & ~keep_mask).sum())} reachable fragments dropped"
    assert not np.any(keep_mask[far]), f"{int((far & keep_mask).sum())} unreachable fragments kept"
    return int(keep_mask.sum()), int(b["n"])
