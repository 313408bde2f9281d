"""Trajectory evaluation against a ground-truth TUM file: absolute trajectory
error (ATE) of the camera centres after a rigid or similarity alignment.

The tracker writes one TUM row per window, "t tx ty tz qx qy qz qw", holding
the world->camera pose (viewpoint.T and viewpoint.R, utils/tracker.py:257-265).
Ground-truth TUM files usually hold the camera->world pose; each file's
convention is an argument here.

* ``read_tum``: rows of a TUM file (blank lines and "#" comments skipped).
* ``associate``: timestamp matching of the TUM RGB-D benchmark tools
  (associate.py, Sturm et al. 2012): every pair closer than ``max_diff``
  (after ``offset``) is a candidate; candidates are taken in order of
  increasing difference, each timestamp used at most once; the matches are
  returned in the order of the first list.
* ``align_umeyama``: least-squares s, R, t with dst ~ s R src + t (Umeyama,
  IEEE TPAMI 13(4), 1991), det R = +1.
* ``ate``: errors |s R c_est + t - c_gt| of the associated camera centres.
"""
from __future__ import annotations

import numpy as np


def quat_xyzw_to_rot(q):
    """Rotation matrices [N,3,3] of unit quaternions [N,4] given as (x, y, z, w)
    (normalised first, as scipy's Rotation.from_quat does)."""
    q = np.asarray(q, np.float64).reshape(-1, 4)
    q = q / np.linalg.norm(q, axis=1, keepdims=True)
    x, y, z, w = q.T
    R = np.empty((q.shape[0], 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def read_tum(path):
    """(timestamps [N] float64, poses [N,4,4] float64) of a TUM trajectory
    file, each pose [R | t] exactly as its row stores it."""
    rows = []
    with open(path) as f:
        for n, line in enumerate(f, 1):
            s = line.strip()
            if not s or s.startswith("#"):
                continue
            v = s.replace(",", " ").split()
            if len(v) != 8:
                raise ValueError(f"{path}:{n}: a TUM row has 8 fields, this one {len(v)}")
            rows.append([float(a) for a in v])
    a = np.array(rows, np.float64).reshape(-1, 8)
    T = np.tile(np.eye(4), (a.shape[0], 1, 1))
    T[:, :3, :3] = quat_xyzw_to_rot(a[:, 4:8]) if a.shape[0] else np.zeros((0, 3, 3))
    T[:, :3, 3] = a[:, 1:4]
    return a[:, 0], T


def associate(t_first, t_second, max_diff=0.02, offset=0.0):
    """Index pairs (i, j) matching t_first[i] with t_second[j] + offset, as the
    TUM benchmark's associate.py does (greedy by increasing |difference|,
    each index used once, pairs at max_diff or more rejected), sorted by i."""
    a = np.asarray(t_first, np.float64)
    b = np.asarray(t_second, np.float64) + offset
    order = np.argsort(b, kind="stable")
    bs = b[order]
    cand = []
    for i, ta in enumerate(a):
        lo = np.searchsorted(bs, ta - max_diff, side="left")
        hi = np.searchsorted(bs, ta + max_diff, side="right")
        for k in range(lo, hi):
            d = abs(ta - bs[k])
            if d < max_diff:
                cand.append((d, i, int(order[k])))
    cand.sort()
    used_a, used_b, out = set(), set(), []
    for d, i, j in cand:
        if i in used_a or j in used_b:
            continue
        used_a.add(i)
        used_b.add(j)
        out.append((i, j))
    out.sort()
    return np.array(out, np.int64).reshape(-1, 2)


def align_umeyama(src, dst, with_scale=False):
    """(s, R, t) minimising sum |dst_k - (s R src_k + t)|^2 over rotations R
    (det +1), translations t and, with ``with_scale``, scales s > 0."""
    src = np.asarray(src, np.float64).reshape(-1, 3)
    dst = np.asarray(dst, np.float64).reshape(-1, 3)
    n = src.shape[0]
    if n < 3:
        raise ValueError("alignment needs at least 3 associated poses")
    ms, md = src.mean(0), dst.mean(0)
    xs, xd = src - ms, dst - md
    cov = xd.T @ xs / n
    U, D, Vt = np.linalg.svd(cov)
    S = np.eye(3)
    if np.linalg.det(U) * np.linalg.det(Vt) < 0:
        S[2, 2] = -1.0
    R = U @ S @ Vt
    s = float(np.trace(np.diag(D) @ S) / ((xs ** 2).sum() / n)) if with_scale else 1.0
    t = md - s * R @ ms
    return s, R, t


def camera_centres(T, world_to_camera):
    """Camera centres in world coordinates of poses [N,4,4]."""
    T = np.asarray(T, np.float64)
    if world_to_camera:
        return -np.einsum("nji,nj->ni", T[:, :3, :3], T[:, :3, 3])
    return T[:, :3, 3].copy()


def ate(est, gt, est_world_to_camera=True, gt_world_to_camera=False, align="se3", max_diff=0.02, offset=0.0):
    """Absolute trajectory error of an estimate against ground truth.

    est, gt: TUM file paths or (timestamps, poses [N,4,4]) pairs.  The
    estimate's rows are world->camera by default (what the tracker writes),
    the ground truth's camera->world.  align: "se3" (rotation + translation),
    "sim3" (also scale) or "none".  Returns a dict with rmse, mean, median,
    max (m), the per-pair errors, the pairs and the alignment (s, R, t)."""
    te, Te = read_tum(est) if isinstance(est, str) else (np.asarray(est[0], np.float64), np.asarray(est[1]))
    tg, Tg = read_tum(gt) if isinstance(gt, str) else (np.asarray(gt[0], np.float64), np.asarray(gt[1]))
    pairs = associate(te, tg, max_diff, offset)
    if pairs.shape[0] == 0:
        raise ValueError("no estimate timestamp is within max_diff of a ground-truth timestamp")
    ce = camera_centres(Te[pairs[:, 0]], est_world_to_camera)
    cg = camera_centres(Tg[pairs[:, 1]], gt_world_to_camera)
    if align == "none":
        s, R, t = 1.0, np.eye(3), np.zeros(3)
    elif align in ("se3", "sim3"):
        s, R, t = align_umeyama(ce, cg, with_scale=align == "sim3")
    else:
        raise ValueError(f"align must be 'se3', 'sim3' or 'none', not {align!r}")
    err = np.linalg.norm(s * ce @ R.T + t - cg, axis=1)
    return {"rmse": float(np.sqrt((err ** 2).mean())), "mean": float(err.mean()), "median": float(np.median(err)),
            "max": float(err.max()), "errors": err, "pairs": pairs, "scale": s, "R": R, "t": t}
