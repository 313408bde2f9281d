"""The reference's per-window optimisation over the event pyramid on the
batched fused path.

utils/tracker.py:149-240 tracks one window in three pyramid levels, coarse
to fine.  At the top level (2) the coarse stage optimises the pose only
against the unsigned images; when it converges (tracker.py:65-76) the fine
stage takes over at the same level, optimising pose and velocity against the
signed images with the learning-rate ramp from that iteration; levels 1 and 0
run the fine stage from their first iteration.  A stage ends on convergence
or at its iteration cap (max_optim_iter, counted from the fine stage's start
for the fine stage), and a coarse stage that hits its cap ends the level.
One Adam state carries across the three levels; the loss history and the
iteration count restart at each level.

Here B items (pose hypotheses of one window) run that control flow together:
each level is a BatchedTracker at the level's resolution (focal lengths and
image size scaled by 2^-level, frame.py:61-83), the event frame at the level
is the nearest-resized one (tracker.py:78-91), and the stage of every item
(coarse, fine, done) lives on the device, switched by the update kernel after
each step.  The host replays a captured graph of `chunk_iters` iterations and
reads one flag per chunk to see whether every item is done.
"""
from __future__ import annotations

import torch

from .frame import image_pyramid
from .fused import BatchedTracker


class PyramidTracker:
    LEVELS = 3

    def __init__(self, gaussians, W, H, fx, fy, B, lr=(0.004, 0.004, 0.002, 0.002), max_optim_iter=200,
                 converged_threshold=1e-4, chunk_iters=10, tile=(16, 16), device="cuda"):
        self.dev = torch.device(device)
        self.B = B
        self.max_optim_iter = int(max_optim_iter)
        self.threshold = float(converged_threshold)
        self.chunk_iters = int(chunk_iters)
        self.trackers = []
        for lvl in range(self.LEVELS):
            s = 0.5 ** lvl
            w, h = int(W * s), int(H * s)
            ev = torch.zeros(1, h, w, dtype=torch.float32, device=self.dev)
            # a level holds at most the coarse and the fine stage, each capped
            hist = 2 * self.max_optim_iter + 2 * self.chunk_iters + 8
            self.trackers.append(BatchedTracker(gaussians, w, h, fx * s, fy * s, ev, B, tile=tile, lr=lr,
                                                max_optim_iter=self.max_optim_iter, hist_len=hist, device=self.dev,
                                                map_from=self.trackers[0] if self.trackers else None))

    _CARRY = ("pose", "lin", "ang", "adam_m", "adam_v", "adam_t")

    def _run_level(self, tr, lvl):
        """Steps the level's items until every one is done; returns the
        iterations run per item (the reference's optim_iter + 1)."""
        snap = {k: getattr(tr, k).clone() for k in tr._STATE}
        kw = dict(staged=True, converged_threshold=self.threshold)
        limit = 2 * self.max_optim_iter + 2
        for _attempt in range(3):
            if tr.cap == 0:
                tr.measure_capacity(margin=1.5)
            if tr.graph is None or tr.graph_kw != kw:
                tr.capture(steps=self.chunk_iters, **kw)
                for k, v in snap.items():     # capture itself does not run, but keep the state exact
                    getattr(tr, k).copy_(v)
            done = 0
            while done < limit:
                tr.graph.replay()
                done += self.chunk_iters
                if bool((tr.stage == 2).all()):
                    break
            if int(tr.overflow.item()) == 0:
                return tr.iter.clone()
            # fragment capacity exceeded in this level: size again and rerun it
            tr.overflow.zero_()
            for k, v in snap.items():
                getattr(tr, k).copy_(v)
            tr.measure_capacity(margin=1.5)
        tr.check_overflow()
        return tr.iter.clone()

    def track(self, poses, lin, ang, dtau, sign_frame):
        """Tracks the B items of one window: poses [B,4,4] world->camera,
        velocities [B,3], window duration dtau (s), signed event frame
        [H, W] at level 0.  Returns per level (2, 1, 0) the iterations each
        item ran, the iteration its fine stage started at, the stage it ended
        in and its losses; the trackers hold the final state
        (self.trackers[0].poses(), .lin, .ang)."""
        pyr = image_pyramid(sign_frame.reshape(1, *sign_frame.shape[-2:]).float(), self.LEVELS)
        B = self.B
        info = []
        prev = None
        for lvl in range(self.LEVELS - 1, -1, -1):
            tr = self.trackers[lvl]
            tr.events[0].copy_(pyr[lvl][0])
            if prev is None:
                tr.set_items(poses, torch.as_tensor(lin).reshape(-1, 3).expand(B, 3),
                             torch.as_tensor(ang).reshape(-1, 3).expand(B, 3), [dtau] * B, [0] * B)
            else:
                for k in self._CARRY:
                    getattr(tr, k).copy_(getattr(prev, k))
                tr.dtau.fill_(float(dtau))
                tr.window.zero_()
            tr.iter.zero_()
            tr.loss_hist.zero_()
            tr.start_vel.zero_()
            tr.stage.fill_(0 if lvl == self.LEVELS - 1 else 1)
            it = self._run_level(tr, lvl)
            info.append(dict(level=lvl, iterations=it.cpu().tolist(), start_vel=tr.start_vel.cpu().tolist(),
                             stage=tr.stage.cpu().tolist(),
                             losses=[tr.loss_hist[b, :int(it[b])].cpu().tolist() for b in range(B)]))
            prev = tr
        return info
