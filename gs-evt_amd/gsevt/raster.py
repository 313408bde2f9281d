"""Host orchestration of one render (forward and backward) over the C ABI.

This is the single-render ("plain") path behind the diff_gaussian_rasterization
drop-in: preprocess -> inclusive scan of tiles_touched -> one read of the
fragment count (the reference's rasterizer also reads num_rendered on the
host) -> key duplication -> stable radix sort -> tile ranges -> blend.  The
backward replays the blend in reverse, reduces the per-Gaussian 2D gradients,
chains them to view space and reduces the pose/velocity gradient on device.
All buffers come from torch's caching allocator; the library allocates
nothing.
"""
from __future__ import annotations

from contextlib import contextmanager
from dataclasses import dataclass

import torch

from . import _lib
from ._lib import call, ptr

GREY_WEIGHTS = (0.2989, 0.5870, 0.1140)   # utils/render_camera/frame.py:40

# Optional per-launch HIP-event timing of selected kernels (name -> list of
# (start, end, meta)); events are recorded on the stream the kernels use.
_TIMERS = None


def enable_kernel_timers(names):
    global _TIMERS
    _TIMERS = {n: [] for n in names}


def disable_kernel_timers():
    global _TIMERS
    t, _TIMERS = _TIMERS, None
    return t


@contextmanager
def timed(name, **meta):
    if _TIMERS is None or name not in _TIMERS:
        yield
        return
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    s.record()
    yield
    e.record()
    _TIMERS[name].append((s, e, meta))


@dataclass
class Camera:
    viewmatrix: torch.Tensor      # [4,4] transposed storage
    projmatrix: torch.Tensor
    projmatrix_raw: torch.Tensor
    campos: torch.Tensor          # [3]
    tanfovx: float
    tanfovy: float
    width: int
    height: int
    bg: torch.Tensor = None       # [3]
    tile_w: int = 16
    tile_h: int = 16

    @property
    def grid(self):
        return (self.width + self.tile_w - 1) // self.tile_w, (self.height + self.tile_h - 1) // self.tile_h


@dataclass
class GaussianMap:
    means3D: torch.Tensor
    opacities: torch.Tensor
    shs: torch.Tensor = None
    colors_precomp: torch.Tensor = None
    scales: torch.Tensor = None
    rotations: torch.Tensor = None
    cov3D_precomp: torch.Tensor = None
    sh_degree: int = 0
    scale_modifier: float = 1.0

    @property
    def P(self):
        return self.means3D.shape[0]

    @property
    def M(self):
        return self.shs.shape[1] if self.shs is not None else 1


def _f32(t):
    return None if t is None else t.contiguous().float()


def key_bits(n_tiles: int) -> int:
    """Sort only the bits that can be set: 32 depth bits + the tile id."""
    return 32 + max(1, (n_tiles - 1).bit_length())


def render_forward(gm: GaussianMap, cam: Camera, channels: int = 3, n_touched: bool = True):
    dev = gm.means3D.device
    P = gm.P
    W, H = cam.width, cam.height
    gx, gy = cam.grid
    n_tiles = gx * gy
    S = 3 if channels == 3 else 2
    stream = _lib.stream_of(dev)
    i32 = dict(dtype=torch.int32, device=dev)
    f32 = dict(dtype=torch.float32, device=dev)
    means = _f32(gm.means3D)
    opac = _f32(gm.opacities)
    shs = _f32(gm.shs)
    cpre = _f32(gm.colors_precomp)
    scales = _f32(gm.scales)
    rots = _f32(gm.rotations)
    cov = _f32(gm.cov3D_precomp)
    view = _f32(cam.viewmatrix)
    proj = _f32(cam.projmatrix)
    campos = _f32(cam.campos)
    bg = torch.zeros(3, **f32) if cam.bg is None else _f32(cam.bg)
    radii = torch.empty(P, **i32)
    tiles = torch.empty(P, **i32)
    splats = torch.empty(max(P, 1) * S * 4, **f32)
    call("gsevt_preprocess", P, gm.sh_degree, gm.M, ptr(means), ptr(scales), float(gm.scale_modifier), ptr(rots),
         ptr(opac), ptr(shs), ptr(cpre), ptr(cov), ptr(view), ptr(proj), ptr(campos), W, H, float(cam.tanfovx),
         float(cam.tanfovy), cam.tile_w, cam.tile_h, channels, None, ptr(radii), ptr(tiles), ptr(splats), stream)
    offsets = torch.empty(max(P, 1), **i32)
    total = torch.zeros(1, **i32)
    scan_scr = torch.empty(int(_lib.load().gsevt_scan_scratch_elems(P)), **i32)
    call("gsevt_scan_inclusive_u32", ptr(tiles), ptr(offsets), P, ptr(scan_scr), ptr(total), stream)
    n = int(total.item())
    cap = max(n, 1)
    keys = torch.empty(2 * cap, dtype=torch.int64, device=dev)
    vals = torch.empty(2 * cap, **i32)
    call("gsevt_duplicate_with_keys", P, ptr(offsets), ptr(radii), ptr(splats), channels, W, H, cam.tile_w,
         cam.tile_h, ptr(keys), ptr(vals), n, stream)
    if n > 0:
        sort_scr = torch.empty(int(_lib.load().gsevt_sort_scratch_elems(n)), **i32)
        call("gsevt_sort_pairs", ptr(keys), ptr(vals), ptr(keys[cap:]), ptr(vals[cap:]), ptr(total), n,
             key_bits(n_tiles), ptr(sort_scr), stream)
    ranges = torch.empty(2 * n_tiles, **i32)
    call("gsevt_tile_ranges", ptr(total), n, ptr(keys), n_tiles, ptr(ranges), stream)
    color = torch.empty(channels, H, W, **f32)
    depth = torch.empty(1, H, W, **f32)
    opacity = torch.empty(1, H, W, **f32)
    final_T = torch.empty(H, W, **f32)
    n_contrib = torch.empty(H, W, **i32)
    touched = torch.zeros(P, **i32) if n_touched else None
    with timed("blend_forward", n=n, W=W, H=H, radii=radii, channels=channels):
        call("gsevt_blend_forward", W, H, cam.tile_w, cam.tile_h, ptr(ranges), ptr(vals), ptr(splats), channels,
             ptr(bg), ptr(color), ptr(depth), ptr(opacity), ptr(final_T), ptr(n_contrib), ptr(touched), stream)
    state = dict(radii=radii, splats=splats, ranges=ranges, vals=vals[:cap], keys=keys[:cap], n=n,
                 final_T=final_T, n_contrib=n_contrib, bg=bg, channels=channels, tiles=tiles)
    if touched is None:
        touched = torch.zeros(P, **i32)
    return dict(color=color, radii=radii, depth=depth, opacity=opacity, n_touched=touched), state


def render_backward(gm: GaussianMap, cam: Camera, state, grad_color, grad_depth=None, grad_opacity=None,
                    vel_transform=None, delta_time=0.0, linear_vel=None, angular_vel=None, map_grads=None):
    """Returns (pose12 [12] float32 device tensor: dtheta, drho, dw, dv; dict of
    per-Gaussian gradients requested in map_grads)."""
    dev = gm.means3D.device
    P = gm.P
    W, H = cam.width, cam.height
    stream = _lib.stream_of(dev)
    f32 = dict(dtype=torch.float32, device=dev)
    channels = state["channels"]
    grad10 = torch.zeros(max(P, 1) * 10, **f32)
    gc = grad_color.contiguous().float()
    gd = None if grad_depth is None else grad_depth.contiguous().float()
    go = None if grad_opacity is None else grad_opacity.contiguous().float()
    with timed("blend_backward", n=state["n"], W=W, H=H, radii=state["radii"], channels=channels):
        call("gsevt_blend_backward", W, H, cam.tile_w, cam.tile_h, ptr(state["ranges"]), ptr(state["vals"]),
             ptr(state["splats"]), channels, ptr(state["bg"]), ptr(state["final_T"]), ptr(state["n_contrib"]),
             ptr(gc), ptr(gd), ptr(go), ptr(grad10), stream)
    map_grads = map_grads or {}
    M = gm.M
    outs = {}
    shapes = dict(means3D=(P, 3), scales=(P, 3), rotations=(P, 4), shs=(P, M, 3), opacities=(P, 1),
                  means2D=(P, 3), colors_precomp=(P, 3), cov3D_precomp=(P, 6))
    for k, want in map_grads.items():
        if want:
            outs[k] = torch.zeros(*shapes[k], **f32)
    nparts = int(_lib.load().gsevt_pose_partials_count(P))
    partials = torch.empty(nparts * 18, **f32)
    # converted copies must outlive the launches: a temporary freed right after
    # ptr() would hand its block to the next allocation before the kernel runs
    means, scales, rots = _f32(gm.means3D), _f32(gm.scales), _f32(gm.rotations)
    shs, cpre, cov = _f32(gm.shs), _f32(gm.colors_precomp), _f32(gm.cov3D_precomp)
    view, proj, praw, campos = (_f32(cam.viewmatrix), _f32(cam.projmatrix), _f32(cam.projmatrix_raw),
                                _f32(cam.campos))
    call("gsevt_preprocess_backward", P, gm.sh_degree, M, ptr(means), ptr(scales), float(gm.scale_modifier),
         ptr(rots), ptr(shs), ptr(cpre), ptr(cov), ptr(view), ptr(proj), ptr(praw), ptr(campos), W, H,
         float(cam.tanfovx), float(cam.tanfovy), cam.tile_w, cam.tile_h, ptr(state["radii"]), ptr(grad10), channels,
         None, ptr(partials), ptr(outs.get("means3D")), ptr(outs.get("scales")), ptr(outs.get("rotations")),
         ptr(outs.get("shs")), ptr(outs.get("opacities")), ptr(outs.get("means2D")), ptr(outs.get("colors_precomp")),
         ptr(outs.get("cov3D_precomp")), stream)
    pose12 = torch.empty(12, **f32)
    vel = None if vel_transform is None else _f32(vel_transform)
    lin = None if linear_vel is None else _f32(linear_vel)
    ang = None if angular_vel is None else _f32(angular_vel)
    call("gsevt_pose_finalize", ptr(partials), nparts, ptr(view), ptr(vel), float(delta_time), ptr(lin), ptr(ang),
         ptr(pose12), 0, stream)
    state["grad10"] = grad10
    return pose12, outs


def mark_visible(positions: torch.Tensor, viewmatrix: torch.Tensor, projmatrix: torch.Tensor) -> torch.Tensor:
    """Boolean [P] mask: False where the view-space z of a position is <= 0.2
    (the near-plane test of the rasterizer's mark_visible)."""
    dev = positions.device
    if dev.type != "cuda":
        raise RuntimeError("mark_visible runs on the GPU; positions must be a CUDA tensor")
    pos = positions.detach().to(torch.float32).contiguous()
    if pos.dim() != 2 or pos.shape[1] != 3:
        raise ValueError(f"positions must be [P, 3] (got {tuple(pos.shape)})")
    P = pos.shape[0]
    present = torch.zeros(P, dtype=torch.bool, device=dev)
    if P == 0:
        return present
    view = viewmatrix.detach().to(dev, torch.float32).contiguous()
    proj = projmatrix.detach().to(dev, torch.float32).contiguous()
    call("gsevt_mark_visible", P, ptr(pos), ptr(view), ptr(proj), ptr(present), _lib.stream_of(dev))
    return present
