"""Batched, fused tracking: many (pose hypothesis, event window) items per
launch, the whole iteration on the device, capturable in a HIP graph.

One iteration of every item = render of the last and next camera of its
window (grey intensity only), the event loss of utils/tracker.py:93-103 on the
normalised intensity change of utils/render_camera/frame.py:86-92, the exact
adjoint through the rasterizer to (theta, rho, w, v), torch-default Adam and
the update of utils/render_camera/camera.py:129-155 (velocity first, then the
left-multiplicative pose update).  The host issues a fixed sequence of
launches with no synchronisation; buffers are allocated once (the fragment
capacity is measured once and guarded by a device overflow flag).
"""
from __future__ import annotations

import ctypes
import math
import time

import torch

from . import _lib
from ._lib import call, ptr
from .camera import ZFAR, ZNEAR, projection_matrix

GREY = (0.2989, 0.5870, 0.1140)
# 256-pixel tiles made of four 8x8 wave blocks
TILE_SHAPES = ((16, 16), (8, 32), (32, 8))
BLEND_SCHEDULES = ("grid", "queue", "queue_xcd")


class ItemState(ctypes.Structure):
    _fields_ = [("pose", ctypes.c_void_p), ("lin", ctypes.c_void_p), ("ang", ctypes.c_void_p),
                ("adam_m", ctypes.c_void_p), ("adam_v", ctypes.c_void_p), ("adam_t", ctypes.c_void_p),
                ("iter", ctypes.c_void_p), ("dtau", ctypes.c_void_p), ("window", ctypes.c_void_p),
                ("loss", ctypes.c_void_p), ("loss_hist", ctypes.c_void_p), ("hist_len", ctypes.c_int),
                ("grad12", ctypes.c_void_p), ("stage", ctypes.c_void_p), ("start_vel", ctypes.c_void_p)]


_EXTRA = {
    "gsevt_tr_camera_record_bytes": (ctypes.c_int, []),
    "gsevt_tr_map_cov": (ctypes.c_int, [ctypes.c_int, ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_void_p]),
    "gsevt_tr_cams": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ItemState), ctypes.c_void_p, ctypes.c_float,
                                     ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "gsevt_tr_preprocess": (ctypes.c_int, [ctypes.c_int] * 4 + [ctypes.c_void_p] * 11
                            + [ctypes.c_int, ctypes.c_void_p]),
    "gsevt_seg_sort_scratch_elems": (ctypes.c_longlong, [ctypes.c_int, ctypes.c_longlong]),
    "gsevt_seg_sort_u32": (ctypes.c_int, [ctypes.c_int, ctypes.c_longlong, ctypes.c_void_p, ctypes.c_longlong,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    "gsevt_tr_bin_scratch_elems": (ctypes.c_longlong, [ctypes.c_int, ctypes.c_int]),
    "gsevt_tr_emit": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_longlong] + [ctypes.c_void_p] * 8),
    "gsevt_tr_ranges": (ctypes.c_int, [ctypes.c_int, ctypes.c_longlong, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    "gsevt_tr_blend_forward": (ctypes.c_int, [ctypes.c_int, ctypes.c_void_p, ctypes.c_float, ctypes.c_float]
                               + [ctypes.c_int] * 5 + [ctypes.c_longlong] + [ctypes.c_void_p] * 2
                               + [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_float] + [ctypes.c_void_p] * 4),
    "gsevt_tr_loss_partials_count": (ctypes.c_int, []),
    "gsevt_tr_loss": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.POINTER(ItemState), ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p]),
    "gsevt_tr_blend_backward": (ctypes.c_int, [ctypes.c_int, ctypes.c_void_p, ctypes.c_float, ctypes.c_float]
                                + [ctypes.c_int] * 5 + [ctypes.c_longlong] + [ctypes.c_void_p] * 2
                                + [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_float] + [ctypes.c_void_p] * 4
                                + [ctypes.POINTER(ItemState)]
                                + [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "gsevt_tr_pose_partials_count": (ctypes.c_int, [ctypes.c_int]),
    "gsevt_tr_blend_queue_count": (ctypes.c_int, []),
    "gsevt_tr_blend_forward_queue": (ctypes.c_int, [ctypes.c_int, ctypes.c_void_p, ctypes.c_float, ctypes.c_float]
                                     + [ctypes.c_int] * 5 + [ctypes.c_longlong] + [ctypes.c_void_p] * 2
                                     + [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_float] + [ctypes.c_void_p] * 4
                                     + [ctypes.c_int, ctypes.c_int, ctypes.c_void_p]),
    "gsevt_tr_blend_backward_queue": (ctypes.c_int, [ctypes.c_int, ctypes.c_void_p, ctypes.c_float, ctypes.c_float]
                                      + [ctypes.c_int] * 5 + [ctypes.c_longlong] + [ctypes.c_void_p] * 2
                                      + [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_float] + [ctypes.c_void_p] * 4
                                      + [ctypes.POINTER(ItemState)]
                                      + [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                         ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]),
    "gsevt_seg_sort_u32_bits": (ctypes.c_int, [ctypes.c_int, ctypes.c_longlong, ctypes.c_void_p, ctypes.c_longlong,
                                               ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                               ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                               ctypes.c_int, ctypes.c_void_p]),
    "gsevt_tr_emit_packed": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_void_p] + [ctypes.c_int] * 7
                             + [ctypes.c_longlong] + [ctypes.c_void_p] * 6 + [ctypes.c_void_p, ctypes.c_int,
                                                                                ctypes.c_void_p]),
    "gsevt_tr_ranges_packed": (ctypes.c_int, [ctypes.c_int, ctypes.c_longlong, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    "gsevt_tr_preprocess_backward": (ctypes.c_int, [ctypes.c_int] * 4 + [ctypes.c_void_p] * 8
                                     + [ctypes.c_int, ctypes.c_void_p]),
    "gsevt_tr_update": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ItemState),
                                       ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_void_p]),
    "gsevt_tr_update_staged": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ItemState),
                                              ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_void_p]),
}
_LOADED = False


def _lib_ext():
    global _LOADED
    L = _lib.load()
    if not _LOADED:
        for name, (res, args) in _EXTRA.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LOADED = True
    return L


def first_pass_bits(key_bits):
    """Digit bits of the first pass of gsevt_seg_sort_u32_bits over key_bits
    bits: an even number of passes of at most 8 bits, the first taking the
    remainder."""
    passes = (key_bits + 7) // 8
    passes += passes & 1
    dbits = (key_bits + passes - 1) // passes
    return max(1, key_bits - (passes - 1) * dbits)


class _HostFloats:
    """Host float array kept alive with its address (passed as const float*)."""

    def __init__(self, vals):
        self.arr = (ctypes.c_float * len(vals))(*[float(v) for v in vals])
        self.ptr = ctypes.cast(self.arr, ctypes.c_void_p)




class BatchedTracker:
    """B items tracked together on one GPU.  Items are described by their
    initial world->camera pose [B,4,4], velocities [B,3], window duration [B]
    and the index of their event frame in `events` ([n_windows, H, W])."""

    def __init__(self, gaussians, W, H, fx, fy, events, B, tile=(16, 16), lr=(0.004, 0.004, 0.002, 0.002),
                 max_optim_iter=200, hist_len=1024, device="cuda", bg=(0.0, 0.0, 0.0), chunk=None,
                 cull_tiles=True, schedule="grid", queue_workgroups=2048, map_from=None):
        self.L = _lib_ext()
        # blends: one workgroup per (render, tile) ("grid"), or a persistent
        # grid of queue_workgroups pulling items from one queue ("queue") or
        # from 8 per-XCD queues ("queue_xcd")
        if schedule not in BLEND_SCHEDULES:
            raise ValueError(f"schedule must be one of {BLEND_SCHEDULES}")
        self.schedule = schedule
        self.queue_workgroups = int(queue_workgroups)
        # packed fragments only: drop the tiles of a rect that the splat cannot
        # reach (alpha < 1/255 on the whole tile) before the tile sort
        self.cull_tiles = bool(cull_tiles)
        # packed fragments: the emission also writes the tile sort's
        # first-pass histogram (False: that pass counts the keys itself)
        self.fuse_first_hist = True
        dev = torch.device(device)
        self.dev = dev
        self.B = B
        # items go through the launches `chunk` at a time; scratch is sized for one chunk
        self.chunk = B if chunk is None else min(int(chunk), B)
        self.n_chunks = (B + self.chunk - 1) // self.chunk
        self.R = 2 * self.chunk
        self.W, self.H = W, H
        self.fovx = 2 * math.atan(W / (2 * fx))
        self.fovy = 2 * math.atan(H / (2 * fy))
        self.tanfovx = math.tan(self.fovx * 0.5)
        self.tanfovy = math.tan(self.fovy * 0.5)
        self._lr = _HostFloats(lr)
        self.lr = self._lr.ptr
        self.max_optim_iter = max_optim_iter
        self.bg_intensity = float(sum(w * b for w, b in zip(GREY, bg)))
        self._iw = _HostFloats(GREY)
        self.iw = self._iw.ptr
        praw = projection_matrix(ZNEAR, ZFAR, self.fovx, self.fovy).transpose(0, 1).contiguous()
        self._projraw = _HostFloats(praw.reshape(-1).tolist())
        self.projraw = self._projraw.ptr
        # frozen map, 3D covariance precomputed once (or the device copies of
        # another tracker of the same map: map_from)
        g = gaussians
        self.P = int(g.get_xyz.shape[0])
        self.gid_bits = max(1, (self.P - 1).bit_length())
        self.D = int(g.active_sh_degree)
        if map_from is not None:
            assert map_from.P == self.P and map_from.D == self.D
            self.means, self.opac, self.shs, self.cov, self.M = (map_from.means, map_from.opac, map_from.shs,
                                                                 map_from.cov, map_from.M)
        else:
            self.means = g.get_xyz.contiguous().float()
            self.opac = g.get_opacity.reshape(-1).contiguous().float()
            feats = g.get_features.float()
            self.M = int(feats.shape[1])
            K = (self.D + 1) ** 2
            # coefficient-major copy (row 3k + c) for lane-contiguous loads; the map is frozen
            self.shs = feats[:, :K, :].reshape(self.P, 3 * K).t().contiguous()
            scales = g.get_scaling.contiguous().float()
            rots = g.get_rotation.contiguous().float()
            self.cov = torch.empty(self.P, 6, dtype=torch.float32, device=dev)
            call("gsevt_tr_map_cov", self.P, ptr(scales), 1.0, ptr(rots), ptr(self.cov), self._s())
        self.events = events.reshape(-1, H, W).contiguous().float()
        # item state
        f32 = dict(dtype=torch.float32, device=dev)
        i32 = dict(dtype=torch.int32, device=dev)
        self.pose = torch.zeros(B, 12, dtype=torch.float64, device=dev)
        self.lin = torch.zeros(B, 3, **f32)
        self.ang = torch.zeros(B, 3, **f32)
        self.adam_m = torch.zeros(B, 12, **f32)
        self.adam_v = torch.zeros(B, 12, **f32)
        self.adam_t = torch.zeros(B, 4, **i32)
        self.iter = torch.zeros(B, **i32)
        self.dtau = torch.zeros(B, **f32)
        self.window = torch.zeros(B, **i32)
        self.loss = torch.zeros(B, **f32)
        self.hist_len = hist_len
        self.loss_hist = torch.zeros(B, hist_len, **f32)
        self.grad12 = torch.zeros(B, 12, **f32)
        # per-item stage of a pyramid level (staged steps): 0 coarse, 1 fine, 2 done
        self.stage = torch.ones(B, **i32)
        self.start_vel = torch.zeros(B, **i32)
        self.states = []
        for c in range(self.n_chunks):
            b0 = c * self.chunk
            o = lambda t: ptr(t[b0:]).value
            self.states.append((min(self.chunk, B - b0), ItemState(
                o(self.pose), o(self.lin), o(self.ang), o(self.adam_m), o(self.adam_v), o(self.adam_t),
                o(self.iter), o(self.dtau), o(self.window), o(self.loss), o(self.loss_hist), hist_len,
                o(self.grad12), o(self.stage), o(self.start_vel))))
        self.state = self.states[0][1]
        # per-render buffers (fragment capacity sized later)
        R, P = self.R, self.P
        cam_bytes = int(self.L.gsevt_tr_camera_record_bytes())
        self.cams = torch.empty(R * cam_bytes, dtype=torch.uint8, device=dev)
        self.vel = torch.empty(R, 16, **f32)
        self.dts = torch.empty(R, **f32)
        self.splats = torch.empty(R * P * 8, **f32)
        self.rects = None    # [R][P] uint2, allocated only for unpacked fragments (_rects)
        # tile counts in Gaussian and in depth order: only the unpacked emission
        # reads them (packed, they travel in the depth-sort values); _tiles()
        self.tiles = None
        self.tsorted = None
        self.dkey = torch.empty(R * P, **i32)
        self.dval = torch.empty(R * P, **i32)
        self.dkey2 = torch.empty(R * P, **i32)
        self.dval2 = torch.empty(R * P, **i32)
        self.dsort_scr = torch.empty(int(self.L.gsevt_seg_sort_scratch_elems(R, P)), **i32)
        self.tincl = torch.empty(R * P, **i32)
        self.bin_scr = torch.empty(int(self.L.gsevt_tr_bin_scratch_elems(R, P)), **i32)
        self.nfrag = torch.zeros(R, **i32)     # slots of every rect (capacity check)
        self.nkept = torch.zeros(R, **i32)     # fragments after the tile cull
        self.overflow = torch.zeros(1, **i32)
        HW = W * H
        self.inten = torch.empty(R * HW, **f32)
        self.final_T = torch.empty(R * HW, **f32)
        self.n_contrib = torch.empty(R * HW, **i32)
        # zero from the start: every iteration's preprocess backward sets the
        # entries it read back to zero, so the backward blend needs no clear
        self.grad6 = torch.zeros(R * P * 6, **f32)
        self.nblk = int(self.L.gsevt_tr_pose_partials_count(P))
        self.partials = torch.empty(R * self.nblk * 18, **f32)
        self.loss_part = torch.empty(B * int(self.L.gsevt_tr_loss_partials_count()) * 3, dtype=torch.float64,
                                     device=dev)
        self.scal = torch.empty(B * 4, **f32)
        self.queue = torch.zeros(int(self.L.gsevt_tr_blend_queue_count()), **i32)
        # tile tables; fragments are packed as tile << gid_bits | gid when
        # both fit in 32 bits
        self.set_tile(tile)

    def _s(self):
        return _lib.stream_of(self.dev)

    def _tiles(self):
        """Tile counts per (render, Gaussian) and their depth-ordered copy,
        made on first use by the unpacked emission."""
        if self.tiles is None:
            self.tiles = torch.empty(self.R * self.P, dtype=torch.int32, device=self.dev)
            self.tsorted = torch.empty(self.R * self.P, dtype=torch.int32, device=self.dev)
        return self.tiles

    def _rects(self):
        """Tile rects per (render, Gaussian); only the unpacked emission reads
        them (packed records carry the rect), so the buffer is made on first use."""
        if self.rects is None:
            self.rects = torch.empty(self.R * self.P * 2, dtype=torch.int32, device=self.dev)
        return self.rects

    # ------------------------------------------------------------ map layout
    def reorder_map(self, perm):
        """Lay the frozen map out in the order perm (new index i holds the
        Gaussian at old index perm[i]).  Fragment values and per-Gaussian
        buffers then use the new indices; self.map_perm maps them back to the
        model's order.  Renders equal those of the model with its Gaussians
        listed in that order (equal depths keep index order, as in the
        reference's stable sort).  Drops the captured graph."""
        perm = torch.as_tensor(perm, device=self.dev).long()
        assert perm.numel() == self.P
        self.means = self.means[perm].contiguous()
        self.opac = self.opac[perm].contiguous()
        self.cov = self.cov[perm].contiguous()
        self.shs = self.shs[:, perm].contiguous()
        self.map_perm = perm if getattr(self, "map_perm", None) is None else self.map_perm[perm]
        self.graph = None

    def contributors(self, render=0):
        """Bool mask [P] over the tracker's Gaussians: those that receive a
        gradient in `render` of the first chunk of items (blended into some
        pixel before it saturates).  Runs one iteration without an update;
        the item state is restored."""
        assert self.cap > 0, "call measure_capacity() first"
        nb, state = self.states[0]
        assert 0 <= render < 2 * nb
        snap = {k: getattr(self, k).clone() for k in self._STATE}
        self._step_chunk(nb, state, False, True, False, 0, keep_grad6=True)
        active = (self.grad6.view(-1, self.P, 6)[render] != 0).any(1)
        self.grad6.zero_()
        for k, v in snap.items():
            getattr(self, k).copy_(v)
        return active

    def reorder_map_by_depth(self, render=0, contributors_first=False):
        """Reorder the map front to back as seen by one render of the current
        items (k_tr_emit_packed then reads the splat records of nearby depth
        ranks from nearby addresses in every render whose view is close to
        that one, instead of one cache line per record).

        contributors_first: the Gaussians that receive a gradient in that
        render (contributors()) come first, each part front to back.  Renders
        whose views are close to that one then find their contributing
        Gaussians together: the preprocess backward evaluates a Gaussian's
        pose partials only where its gradients are non-zero, and a wave's
        lanes are then mostly all busy or all idle instead of a few busy lanes
        holding the whole wave (measured: 8.95 -> 6.93 ms per 256-item launch
        at 1M Gaussians)."""
        assert self.cap > 0, "call measure_capacity() first"
        nb, state = self.states[0]
        assert 0 <= render < 2 * nb
        self._forward_geometry(self.cap, nb, state)
        order = self.dval[render * self.P:(render + 1) * self.P].long() & self.gid_mask
        if contributors_first:
            active = self.contributors(render)[order]
            order = torch.cat([order[active], order[~active]])
        self.reorder_map(order)

    # ------------------------------------------------------------ tile shape
    def set_tile(self, tile):
        """Switch the tile shape (16x16, 8x32 or 32x8).  Resizes the tile
        tables and drops the fragment buffers and the captured graph:
        measure_capacity() and capture() must follow."""
        tile = (int(tile[0]), int(tile[1]))
        if tile not in TILE_SHAPES:
            raise ValueError(f"tile shape {tile} not supported; use one of {TILE_SHAPES}")
        self.graph = None
        self.cap = 0
        for name in ("fgid", "fgid2", "ftile", "ftile2", "fsort_scr"):
            if hasattr(self, name):
                delattr(self, name)
        self.tile = tile
        self.gx = (self.W + tile[0] - 1) // tile[0]
        self.gy = (self.H + tile[1] - 1) // tile[1]
        self.n_tiles = self.gx * self.gy
        self.tile_bits = max(1, (self.n_tiles - 1).bit_length())
        self.first_tile_bits = first_pass_bits(self.tile_bits)
        self.packed = (self.n_tiles << self.gid_bits) < (1 << 32)
        self.gid_mask = (1 << self.gid_bits) - 1 if self.packed else 0xFFFFFFFF
        self.ranges = torch.empty(self.R * self.n_tiles * 2, dtype=torch.int32, device=self.dev)

    _STATE = ("pose", "lin", "ang", "adam_m", "adam_v", "adam_t", "iter", "loss", "loss_hist", "grad12", "stage",
              "start_vel")

    def autotune(self, shapes=None, iters=3, warmup=1, schedules=None, **kw):
        """Pick the tile shape (and, with `schedules`, the blend schedule) with
        the fastest graph-captured iteration on the current items.  Each
        candidate is timed over `iters` replays after `warmup`; the item state
        is restored after every trial, so tuning does not advance the
        optimisation.  Leaves the tracker on the best candidate with capacity
        measured and returns (best, {candidate: ms per iteration}); candidates
        are tile shapes, or (shape, schedule) pairs when schedules are given."""
        shapes = list(TILE_SHAPES if shapes is None else shapes)
        scheds = [self.schedule] if schedules is None else list(schedules)
        snap = {k: getattr(self, k).clone() for k in self._STATE}

        def restore():
            for k, v in snap.items():
                getattr(self, k).copy_(v)

        times = {}
        for shp in shapes:
            self.set_tile(shp)
            restore()
            self.measure_capacity()
            for sch in scheds:
                if sch not in BLEND_SCHEDULES:
                    raise ValueError(f"schedule must be one of {BLEND_SCHEDULES}")
                self.schedule = sch
                self.capture(**kw)
                for _ in range(warmup):
                    restore()
                    self.graph.replay()
                restore()
                torch.cuda.synchronize(self.dev)
                t0 = time.perf_counter()
                for _ in range(iters):
                    self.graph.replay()
                torch.cuda.synchronize(self.dev)
                times[shp if schedules is None else (shp, sch)] = 1000.0 * (time.perf_counter() - t0) / iters
                self.check_overflow()
        best = min(times, key=times.get)
        if schedules is None:
            self.set_tile(best)
        else:
            self.set_tile(best[0])
            self.schedule = best[1]
        restore()
        self.measure_capacity()
        self.tune_times = times
        return best, times

    # ---------------------------------------------------------------- items
    def set_items(self, poses, lin, ang, dtau, window):
        """Initial world->camera poses [B,4,4], velocities [B,3], window
        durations [B] (seconds) and event-frame indices [B]; resets Adam."""
        poses = torch.as_tensor(poses, dtype=torch.float64, device=self.dev).reshape(self.B, 4, 4)
        self.pose.copy_(torch.cat([poses[:, :3, :3].reshape(self.B, 9), poses[:, :3, 3]], 1))
        self.lin.copy_(torch.as_tensor(lin, dtype=torch.float32, device=self.dev).reshape(self.B, 3))
        self.ang.copy_(torch.as_tensor(ang, dtype=torch.float32, device=self.dev).reshape(self.B, 3))
        self.dtau.copy_(torch.as_tensor(dtau, dtype=torch.float32, device=self.dev).reshape(self.B))
        self.window.copy_(torch.as_tensor(window, dtype=torch.int32, device=self.dev).reshape(self.B))
        self.reset_optimizer()

    def reset_optimizer(self):
        self.adam_m.zero_()
        self.adam_v.zero_()
        self.adam_t.zero_()
        self.iter.zero_()
        self.loss_hist.zero_()

    def poses(self):
        p = self.pose.double()
        T = torch.zeros(self.B, 4, 4, dtype=torch.float64, device=self.dev)
        T[:, :3, :3] = p[:, :9].reshape(self.B, 3, 3)
        T[:, :3, 3] = p[:, 9:]
        T[:, 3, 3] = 1
        return T

    # --------------------------------------------------------------- stages
    def _forward_geometry(self, cap, nb, state):
        L, s = self.L, self._s()
        R, P = 2 * nb, self.P
        call("gsevt_tr_cams", nb, ctypes.byref(state), self.projraw, self.tanfovx, self.tanfovy, self.W,
             self.H, self.tile[0], self.tile[1], ptr(self.cams), ptr(self.vel), ptr(self.dts), s)
        # packed fragments: the depth-sorted values also carry each Gaussian's
        # tile count (the same fit: n_tiles << gid_bits < 2^32)
        call("gsevt_tr_preprocess", R, P, self.D, self.M, ptr(self.means), ptr(self.cov), ptr(self.opac),
             ptr(self.shs), self.iw, ptr(self.cams), ptr(self.splats), None if self.packed else ptr(self._rects()),
             None if self.packed else ptr(self._tiles()), ptr(self.dkey), ptr(self.dval),
             self.gid_bits if self.packed else 0, s)
        call("gsevt_seg_sort_u32", R, P, None, P, ptr(self.dkey), ptr(self.dval), ptr(self.dkey2), ptr(self.dval2),
             32, ptr(self.dsort_scr), s)
        if self.packed:
            call("gsevt_tr_emit_packed", R, P, ptr(self.dval), None, ptr(self.splats),
                 int(self.cull_tiles), self.W, self.H, self.tile[0], self.tile[1], self.gid_bits, 1, cap,
                 None, ptr(self.tincl), ptr(self.bin_scr), ptr(self.nfrag), ptr(self.overflow),
                 ptr(self.fgid) if cap > 1 else ptr(self._dummy),
                 ptr(self.fsort_scr) if self._emit_hist(cap) else None, self.first_tile_bits, s)
        else:
            call("gsevt_tr_emit", R, P, ptr(self.dval), ptr(self._tiles()), ptr(self._rects()), self.W, self.tile[0], cap,
                 ptr(self.tsorted), ptr(self.tincl), ptr(self.bin_scr), ptr(self.nfrag), ptr(self.overflow),
                 ptr(self.ftile) if cap > 1 else ptr(self._dummy), ptr(self.fgid) if cap > 1 else ptr(self._dummy),
                 s)

    def _emit_hist(self, cap):
        """Whether the packed emission also counts the tile sort's first-pass
        histogram (so that pass reads no keys to count them)."""
        return self.fuse_first_hist and self.packed and cap > 1 and self.first_tile_bits <= 6

    def measure_capacity(self, margin=1.25, extra=1 << 16):
        """Fragment counts of the current items; sizes the fragment buffers."""
        self._dummy = torch.empty(self.R, dtype=torch.int32, device=self.dev)   # one slot per render (cap 1)
        nmax = 0
        for nb, state in self.states:
            self._forward_geometry(1, nb, state)
            torch.cuda.synchronize(self.dev)
            nmax = max(nmax, int(self.nfrag[:2 * nb].max().item()))
        self.alloc_fragments(int(nmax * margin) + extra)
        self.overflow.zero_()
        return nmax

    def alloc_fragments(self, cap):
        self.cap = int(cap)
        i32 = dict(dtype=torch.int32, device=self.dev)
        n = self.R * self.cap
        self.fgid = torch.empty(n, **i32)
        self.fgid2 = torch.empty(n, **i32)
        if not self.packed:
            self.ftile = torch.empty(n, **i32)
            self.ftile2 = torch.empty(n, **i32)
        self.fsort_scr = torch.empty(int(self.L.gsevt_seg_sort_scratch_elems(self.R, self.cap)), **i32)
        self.graph = None

    def step(self, do_step=True, opt_vel=True, unsigned=False, start_vel_iter=0, staged=False,
             converged_threshold=1e-4):
        """One iteration of every item (launches only, no synchronisation).
        staged: each item follows its own stage (self.stage: 0 coarse, 1
        fine, 2 done) and the device switches stages on convergence and at
        the caps of a pyramid level (gsevt_tr_update_staged); opt_vel,
        unsigned and start_vel_iter are then per item."""
        assert self.cap > 0, "call measure_capacity() first"
        for nb, state in self.states:
            self._step_chunk(nb, state, do_step, opt_vel, unsigned, start_vel_iter, staged, converged_threshold)

    def _step_chunk(self, nb, state, do_step, opt_vel, unsigned, start_vel_iter, staged=False,
                    converged_threshold=1e-4, keep_grad6=False):
        """keep_grad6: leave this iteration's per-Gaussian gradients in
        self.grad6 (the caller zeroes it afterwards)."""
        if staged:
            unsigned = -1
        s = self._s()
        R, P, cap = 2 * nb, self.P, self.cap
        self._forward_geometry(cap, nb, state)
        if self.packed:
            kept = self.nkept if self.cull_tiles else self.nfrag
            call("gsevt_seg_sort_u32_bits", R, cap, ptr(self.nfrag), 0, ptr(self.fgid), None, ptr(self.fgid2), None,
                 self.gid_bits, self.gid_bits + self.tile_bits, ptr(self.fsort_scr),
                 ptr(self.nkept) if self.cull_tiles else None, int(self._emit_hist(cap)), s)
            call("gsevt_tr_ranges_packed", R, cap, ptr(self.fgid), ptr(kept), self.gid_bits, self.n_tiles,
                 ptr(self.ranges), s)
        else:
            call("gsevt_seg_sort_u32", R, cap, ptr(self.nfrag), 0, ptr(self.ftile), ptr(self.fgid),
                 ptr(self.ftile2), ptr(self.fgid2), self.tile_bits, ptr(self.fsort_scr), s)
            call("gsevt_tr_ranges", R, cap, ptr(self.ftile), ptr(self.nfrag), self.n_tiles, ptr(self.ranges), s)
        fwd_args = (R, self.projraw, self.tanfovx, self.tanfovy, self.W, self.H, self.tile[0], self.tile[1], P, cap,
                    ptr(self.ranges), ptr(self.fgid), self.gid_mask, ptr(self.splats), self.bg_intensity,
                    ptr(self.inten), ptr(self.final_T), ptr(self.n_contrib))
        if self.schedule == "grid":
            call("gsevt_tr_blend_forward", *fwd_args, s)
        else:
            call("gsevt_tr_blend_forward_queue", *fwd_args, ptr(self.queue), self.queue_workgroups,
                 int(self.schedule == "queue_xcd"), s)
        call("gsevt_tr_loss", nb, self.W, self.H, ptr(self.inten), ptr(self.events), ctypes.byref(state),
             int(unsigned), ptr(self.loss_part), ptr(self.scal), s)
        bwd_args = (R, self.projraw, self.tanfovx, self.tanfovy, self.W, self.H, self.tile[0], self.tile[1], P, cap,
                    ptr(self.ranges), ptr(self.fgid), self.gid_mask, ptr(self.splats), self.bg_intensity,
                    ptr(self.inten), ptr(self.final_T), ptr(self.n_contrib), ptr(self.events), ctypes.byref(state),
                    ptr(self.scal), int(unsigned), ptr(self.grad6), 0)
        if self.schedule == "grid":
            call("gsevt_tr_blend_backward", *bwd_args, s)
        else:
            call("gsevt_tr_blend_backward_queue", *bwd_args, ptr(self.queue), self.queue_workgroups,
                 int(self.schedule == "queue_xcd"), s)
        call("gsevt_tr_preprocess_backward", R, P, self.D, self.M, ptr(self.means), ptr(self.cov), ptr(self.opac),
             ptr(self.shs), self.iw, ptr(self.cams), ptr(self.grad6), ptr(self.partials), int(not keep_grad6), s)
        if staged:
            call("gsevt_tr_update_staged", nb, self.nblk, ptr(self.partials), ptr(self.cams), ptr(self.vel),
                 ptr(self.dts), ctypes.byref(state), self.lr, self.max_optim_iter, float(converged_threshold), s)
        else:
            call("gsevt_tr_update", nb, self.nblk, ptr(self.partials), ptr(self.cams), ptr(self.vel),
                 ptr(self.dts), ctypes.byref(state), self.lr, self.max_optim_iter, int(opt_vel),
                 int(start_vel_iter), int(do_step), s)

    def capture(self, steps=1, **kw):
        """Capture `steps` consecutive iterations in one graph (replayed by
        run())."""
        st = torch.cuda.Stream(device=self.dev)
        st.wait_stream(torch.cuda.current_stream(self.dev))
        g = torch.cuda.CUDAGraph()
        with torch.cuda.stream(st):
            with torch.cuda.graph(g, stream=st):
                for _ in range(steps):
                    self.step(**kw)
        torch.cuda.current_stream(self.dev).wait_stream(st)
        self.graph = g
        self.graph_kw = kw
        return g

    def run(self, n, use_graph=True, **kw):
        for _ in range(n):
            if use_graph and self.graph is not None:
                self.graph.replay()
            else:
                self.step(**kw)

    def device_bytes(self):
        """Device memory held by the tracker's own buffers (map copy, records,
        sort and fragment buffers, gradients, images, queues), in bytes."""
        seen, total = set(), 0
        for v in vars(self).values():
            if torch.is_tensor(v) and v.is_cuda:
                st = v.untyped_storage()
                if st.data_ptr() not in seen:
                    seen.add(st.data_ptr())
                    total += st.nbytes()
        return total

    def check_overflow(self):
        if int(self.overflow.item()) != 0:
            raise RuntimeError(f"fragment capacity {self.cap} exceeded (max {int(self.nfrag.max().item())}); "
                               "call measure_capacity() again")
