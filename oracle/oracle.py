"""ctypes front-end of the CPU oracle (gsevt_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the product.

The oracle restates the published 3D Gaussian Splatting rasterizer that the
reference calls through diff_gaussian_rasterization
(gaussian_splatting/gaussian_renderer/__init__.py:313-384); the rasterizer
itself is absent from the reference tree (.gitmodules:1-3), so parity against
it is unpinned and the oracle is pinned by tests against closed forms, an FP64
brute-force renderer (oracle/brute.py) and central finite differences.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_vp = ctypes.c_void_p
_i = ctypes.c_int


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        build()
    L = ctypes.CDLL(path)
    for suf, fp, ft in (("f32", _f32p, ctypes.c_float), ("f64", _f64p, ctypes.c_double)):
        getattr(L, f"orc_preprocess_{suf}").argtypes = [
            _i, _i, _i, fp, _vp, ft, _vp, fp, _vp, _vp, _vp, fp, fp, fp, _i, _i, ft, ft, _i, _i,
            _i32p, _i32p, fp, fp, fp, fp, _u8p, _i32p]
        getattr(L, f"orc_bin_{suf}").argtypes = [
            _i, _i, _i, _i, _i, _i32p, _i32p, fp, ctypes.c_int64, _u64p, _u32p, _u32p]
        getattr(L, f"orc_bin_{suf}").restype = ctypes.c_int64
        getattr(L, f"orc_blend_fwd_{suf}").argtypes = [
            _i, _i, _i, _i, _u32p, _u32p, fp, fp, fp, _i, fp, fp, fp, fp, fp, fp, _i32p, _vp, _i]
        getattr(L, f"orc_blend_bwd_{suf}").argtypes = [
            _i, _i, _i, _i, _u32p, _u32p, fp, fp, fp, _i, fp, fp, fp, _i32p, fp, _vp, _vp, fp, _i]
        getattr(L, f"orc_preprocess_bwd_{suf}").argtypes = [
            _i, _i, _i, fp, _vp, ft, _vp, _vp, _vp, _vp, fp, fp, fp, fp, _i, _i, ft, ft, _i, _i,
            fp, _i, _f64p, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    L.orc_pose_finalize.argtypes = [_f64p, _f64p, _f64p, ctypes.c_double, _f64p, _f64p, _f64p]
    L.orc_pose_finalize.restype = None
    L.orc_det_expf.argtypes = [ctypes.c_float]
    L.orc_det_expf.restype = ctypes.c_float
    L.orc_set_num_threads.argtypes = [_i]
    L.orc_set_num_threads.restype = _i
    _LIB = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


@dataclass
class Cam:
    """One render's camera, in the settings' storage convention (4x4 matrices
    transposed, i.e. flat arrays read column-major)."""
    view: np.ndarray       # [16]
    proj: np.ndarray       # [16] full projection
    projraw: np.ndarray    # [16]
    campos: np.ndarray     # [3]
    W: int
    H: int
    tanfovx: float
    tanfovy: float
    bg: np.ndarray = None  # [3]
    tile_w: int = 16
    tile_h: int = 16
    vel: np.ndarray = None       # [16] vel_transofrm (transposed storage)
    delta_time: float = 0.0
    lin: np.ndarray = None
    ang: np.ndarray = None


@dataclass
class Scene:
    means: np.ndarray      # [P,3]
    scales: np.ndarray     # [P,3] activated
    rots: np.ndarray       # [P,4] (w,x,y,z) normalised
    opac: np.ndarray       # [P]
    shs: np.ndarray        # [P,M,3]
    sh_degree: int
    scale_modifier: float = 1.0
    colors_pre: np.ndarray = None
    cov_pre: np.ndarray = None

    @property
    def P(self):
        return self.means.shape[0]


def _c(a, dt):
    return None if a is None else np.ascontiguousarray(a, dtype=dt)


def preprocess(scene: Scene, cam: Cam, dtype=np.float32):
    L = lib()
    suf = "f32" if dtype == np.float32 else "f64"
    P = scene.P
    M = scene.shs.shape[1] if scene.shs is not None else 1
    means = _c(scene.means, dtype).reshape(-1)
    scales = _c(scene.scales, dtype)
    rots = _c(scene.rots, dtype)
    opac = _c(scene.opac, dtype).reshape(-1)
    shs = _c(scene.shs, dtype)
    cpre = _c(scene.colors_pre, dtype)
    cov = _c(scene.cov_pre, dtype)
    out = dict(
        radii=np.zeros(P, np.int32), tiles=np.zeros(P, np.int32), xy=np.zeros(2 * P, dtype),
        depth=np.zeros(P, dtype), conic_opac=np.zeros(4 * P, dtype), rgb=np.zeros(3 * P, dtype),
        clamped=np.zeros(3 * P, np.uint8), rect=np.zeros(4 * P, np.int32))
    getattr(L, f"orc_preprocess_{suf}")(
        P, scene.sh_degree, M, means, _ptr(scales), dtype(scene.scale_modifier), _ptr(rots), opac,
        _ptr(shs), _ptr(cpre), _ptr(cov), _c(cam.view, dtype), _c(cam.proj, dtype), _c(cam.campos, dtype),
        cam.W, cam.H, dtype(cam.tanfovx), dtype(cam.tanfovy), cam.tile_w, cam.tile_h,
        out["radii"], out["tiles"], out["xy"], out["depth"], out["conic_opac"], out["rgb"], out["clamped"],
        out["rect"])
    return out


def grid(cam: Cam):
    return (cam.W + cam.tile_w - 1) // cam.tile_w, (cam.H + cam.tile_h - 1) // cam.tile_h


def bin(pre, cam: Cam, dtype=np.float32):
    L = lib()
    suf = "f32" if dtype == np.float32 else "f64"
    P = pre["radii"].shape[0]
    n = int(pre["tiles"].astype(np.int64).sum())
    gx, gy = grid(cam)
    keys = np.zeros(max(n, 1), np.uint64)
    vals = np.zeros(max(n, 1), np.uint32)
    ranges = np.zeros(2 * gx * gy, np.uint32)
    got = getattr(L, f"orc_bin_{suf}")(P, cam.W, cam.H, cam.tile_w, cam.tile_h, pre["tiles"], pre["rect"],
                                      pre["depth"], n, keys, vals, ranges)
    assert got == n
    return dict(keys=keys[:n], vals=vals[:n], ranges=ranges, n=n)


def blend_forward(pre, binning, cam: Cam, colors=None, C=3, dtype=np.float32, n_touched=True):
    L = lib()
    suf = "f32" if dtype == np.float32 else "f64"
    P = pre["radii"].shape[0]
    HW = cam.W * cam.H
    colors = pre["rgb"] if colors is None else np.ascontiguousarray(colors, dtype).reshape(-1)
    bg = np.zeros(C, dtype) if cam.bg is None else _c(cam.bg, dtype)[:C].copy()
    vals = binning["vals"] if binning["n"] > 0 else np.zeros(1, np.uint32)
    out = dict(color=np.zeros(C * HW, dtype), depth=np.zeros(HW, dtype), opacity=np.zeros(HW, dtype),
               final_T=np.zeros(HW, dtype), n_contrib=np.zeros(HW, np.int32),
               n_touched=np.zeros(P, np.int32))
    getattr(L, f"orc_blend_fwd_{suf}")(
        cam.W, cam.H, cam.tile_w, cam.tile_h, binning["ranges"], vals, pre["xy"], pre["conic_opac"],
        colors, C, pre["depth"], bg, out["color"], out["depth"], out["opacity"], out["final_T"],
        out["n_contrib"], _ptr(out["n_touched"]) if n_touched else None, P)
    out["color"] = out["color"].reshape(C, cam.H, cam.W)
    for k in ("depth", "opacity", "final_T", "n_contrib"):
        out[k] = out[k].reshape(cam.H, cam.W)
    return out


def blend_backward(pre, binning, cam: Cam, fwd, g_color, g_depth=None, g_opac=None, colors=None, C=3,
                   dtype=np.float32):
    L = lib()
    suf = "f32" if dtype == np.float32 else "f64"
    P = pre["radii"].shape[0]
    colors = pre["rgb"] if colors is None else np.ascontiguousarray(colors, dtype).reshape(-1)
    bg = np.zeros(C, dtype) if cam.bg is None else _c(cam.bg, dtype)[:C].copy()
    vals = binning["vals"] if binning["n"] > 0 else np.zeros(1, np.uint32)
    grad10 = np.zeros(10 * P, dtype)
    gD = _c(g_depth, dtype)
    gA = _c(g_opac, dtype)
    getattr(L, f"orc_blend_bwd_{suf}")(
        cam.W, cam.H, cam.tile_w, cam.tile_h, binning["ranges"], vals, pre["xy"], pre["conic_opac"], colors,
        C, pre["depth"], bg, _c(fwd["final_T"], dtype).reshape(-1),
        np.ascontiguousarray(fwd["n_contrib"], np.int32).reshape(-1),
        _c(g_color, dtype).reshape(-1), _ptr(gD), _ptr(gA), grad10, P)
    return grad10.reshape(P, 10)


def preprocess_backward(scene: Scene, cam: Cam, grad10, C=3, dtype=np.float32, map_grads=False):
    L = lib()
    suf = "f32" if dtype == np.float32 else "f64"
    P = scene.P
    M = scene.shs.shape[1] if scene.shs is not None else 1
    pose18 = np.zeros(18, np.float64)
    g = {}
    if map_grads:
        g = dict(means3D=np.zeros(3 * P, dtype), scales=np.zeros(3 * P, dtype), rotations=np.zeros(4 * P, dtype),
                 shs=np.zeros(P * M * 3, dtype), opacities=np.zeros(P, dtype), means2D=np.zeros(3 * P, dtype),
                 colors=np.zeros(3 * P, dtype), cov3D=np.zeros(6 * P, dtype))
    getattr(L, f"orc_preprocess_bwd_{suf}")(
        P, scene.sh_degree, M, _c(scene.means, dtype).reshape(-1), _ptr(_c(scene.scales, dtype)),
        dtype(scene.scale_modifier), _ptr(_c(scene.rots, dtype)), _ptr(_c(scene.shs, dtype)),
        _ptr(_c(scene.colors_pre, dtype)), _ptr(_c(scene.cov_pre, dtype)), _c(cam.view, dtype),
        _c(cam.proj, dtype), _c(cam.projraw, dtype), _c(cam.campos, dtype), cam.W, cam.H, dtype(cam.tanfovx),
        dtype(cam.tanfovy), cam.tile_w, cam.tile_h, np.ascontiguousarray(grad10, dtype).reshape(-1), C, pose18,
        *[_ptr(g.get(k)) for k in ("means3D", "scales", "rotations", "shs", "opacities", "means2D", "colors",
                                    "cov3D")])
    return pose18, g


def pose_finalize(pose18, cam: Cam):
    L = lib()
    out = np.zeros(12, np.float64)
    vel = np.eye(4).reshape(-1) if cam.vel is None else np.asarray(cam.vel, np.float64).reshape(-1)
    lin = np.zeros(3) if cam.lin is None else np.asarray(cam.lin, np.float64)
    ang = np.zeros(3) if cam.ang is None else np.asarray(cam.ang, np.float64)
    L.orc_pose_finalize(np.ascontiguousarray(pose18, np.float64), np.ascontiguousarray(cam.view, np.float64),
                        np.ascontiguousarray(vel), float(cam.delta_time), np.ascontiguousarray(lin),
                        np.ascontiguousarray(ang), out)
    return dict(theta=out[0:3], rho=out[3:6], w=out[6:9], v=out[9:12])


def render(scene: Scene, cam: Cam, dtype=np.float32, n_touched=True):
    """Full oracle forward: preprocess -> bin -> blend.  Returns a dict with
    the preprocess state, binning and the blend outputs."""
    pre = preprocess(scene, cam, dtype)
    b = bin(pre, cam, dtype)
    f = blend_forward(pre, b, cam, dtype=dtype, n_touched=n_touched)
    return dict(pre=pre, bin=b, fwd=f)


def render_backward(scene: Scene, cam: Cam, state, g_color, g_depth=None, g_opac=None, dtype=np.float32,
                    map_grads=False):
    g10 = blend_backward(state["pre"], state["bin"], cam, state["fwd"], g_color, g_depth, g_opac, dtype=dtype)
    pose18, g = preprocess_backward(scene, cam, g10, dtype=dtype, map_grads=map_grads)
    return dict(grad10=g10, pose18=pose18, pose=pose_finalize(pose18, cam), map=g)


def set_num_threads(n: int) -> int:
    """OpenMP threads of the oracle's parallel loops on the calling thread;
    returns the previous maximum."""
    return int(lib().orc_set_num_threads(int(n)))


def det_expf(x: float) -> float:
    return lib().orc_det_expf(float(x))


def mark_visible(means: np.ndarray, view: np.ndarray) -> np.ndarray:
    """Near-plane mask of the rasterizer's mark_visible: False where the
    view-space z (float32, products summed left to right, no fused multiply-add)
    is <= 0.2, True otherwise (NaN included)."""
    m = np.ascontiguousarray(means, dtype=np.float32).reshape(-1, 3)
    V = np.ascontiguousarray(view, dtype=np.float32).reshape(16)
    z = ((V[2] * m[:, 0] + V[6] * m[:, 1]) + V[10] * m[:, 2]) + V[14]
    return ~(z <= np.float32(0.2))


def tile_cull_bounds(pre, binning, cam: Cam):
    """FP64 reach of every fragment of a binning over its tile: qmin, the
    minimum of q = a dx^2 + 2 b dx dy + c dy^2 (the conic of the fragment's
    Gaussian, d = pixel - mean) over the tile's rectangle of pixel centres
    [tx*tw, tx*tw + tw - 1] x [ty*th, ty*th + th - 1]; mag = a dx^2 +
    |2 b dx dy| + c dy^2 at that minimiser (the scale of the rounding in a
    float32 evaluation of q there); and L = ln(255 opacity) (-inf for
    opacity <= 0).  A pixel of the tile has alpha >= 1/255 only if
    qmin <= 2 L.  The minimum of a positive-definite quadratic over a rectangle
    is 0 with the mean inside, else on a side facing the mean, at the
    stationary point of q along that side clamped to the side."""
    gx, _ = grid(cam)
    keys = binning["keys"]
    tile = (keys >> np.uint64(32)).astype(np.int64)
    g = binning["vals"].astype(np.int64)
    xy = pre["xy"].reshape(-1, 2).astype(np.float64)[g]
    co = pre["conic_opac"].reshape(-1, 4).astype(np.float64)[g]
    mx, my = xy[:, 0], xy[:, 1]
    a, b, c, op = co[:, 0], co[:, 1], co[:, 2], co[:, 3]
    x0 = (tile % gx).astype(np.float64) * cam.tile_w
    y0 = (tile // gx).astype(np.float64) * cam.tile_h
    x1, y1 = x0 + (cam.tile_w - 1), y0 + (cam.tile_h - 1)
    cand = []
    for sx in (x0, x1):   # vertical sides
        dx = sx - mx
        cand.append((dx, np.clip(-b * dx / c, y0 - my, y1 - my)))
    for sy in (y0, y1):   # horizontal sides
        dy = sy - my
        cand.append((np.clip(-b * dy / a, x0 - mx, x1 - mx), dy))
    qs = np.stack([a * dx * dx + 2.0 * b * dx * dy + c * dy * dy for dx, dy in cand])
    ms = np.stack([a * dx * dx + np.abs(2.0 * b * dx * dy) + c * dy * dy for dx, dy in cand])
    k = np.argmin(qs, axis=0)
    idx = np.arange(g.shape[0])
    inside = (mx >= x0) & (mx <= x1) & (my >= y0) & (my <= y1)
    qmin = np.where(inside, 0.0, qs[k, idx])
    mag = np.where(inside, 0.0, ms[k, idx])
    with np.errstate(divide="ignore"):
        L = np.where(op > 0, np.log(255.0 * np.maximum(op, 1e-300)), -np.inf)
    return qmin, L, mag
