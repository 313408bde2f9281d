"""Trajectory evaluation (gsevt.evaluate) against closed forms and brute force.

* quaternion -> rotation equals scipy's Rotation.from_quat;
* Umeyama alignment recovers a known similarity / rigid transform exactly
  (to 1e-12) and, when the best orthogonal fit is a reflection (a mirrored
  cloud), the best proper rotation, equal to scipy's Kabsch solver;
* association equals a brute-force restatement of the TUM benchmark's
  associate.py (all pairs, sort, greedy) on random jittered stamps;
* ATE: zero for a noise-free copy of the ground truth in another frame after
  sim3 alignment (se3 when that frame change is rigid), the known 1 cm noise
  level with noise, metres with no alignment, and, with no alignment, equal to
  the per-window camera-centre error of gsevt.sequence.drift on the same
  poses written as TUM files by the tracker's row format;
* read_tum reads the rows SequenceResult.write_tum writes.
"""
import numpy as np
import pytest
import torch
from scipy.spatial.transform import Rotation

from gsevt import evaluate as ev
from gsevt.sequence import drift
from gsevt.tracker import rotation_to_quat_xyzw


def _random_poses(rng, n):
    R = Rotation.random(n, random_state=rng.integers(1 << 31)).as_matrix()
    T = np.tile(np.eye(4), (n, 1, 1))
    T[:, :3, :3] = R
    T[:, :3, 3] = rng.normal(0, 2, (n, 3))
    return T


def _write_tum(path, t, T):
    with open(path, "w") as f:
        f.write("# timestamp tx ty tz qx qy qz qw\n\n")
        for ti, Ti in zip(t, T):
            q = rotation_to_quat_xyzw(Ti[:3, :3])
            f.write(" ".join(str(v) for v in (ti, *Ti[:3, 3].tolist(), *q.tolist())) + "\n")


def test_quaternion_matches_scipy():
    rng = np.random.default_rng(0)
    q = rng.normal(size=(50, 4)) * rng.uniform(0.5, 3, (50, 1))     # not unit: normalised first
    np.testing.assert_allclose(ev.quat_xyzw_to_rot(q), Rotation.from_quat(q).as_matrix(), atol=1e-14)


def test_read_tum_roundtrip(tmp_path):
    rng = np.random.default_rng(1)
    t = np.sort(rng.uniform(0, 10, 20))
    T = _random_poses(rng, 20)
    _write_tum(str(tmp_path / "a.txt"), t, T)
    t2, T2 = ev.read_tum(str(tmp_path / "a.txt"))
    np.testing.assert_array_equal(t2, t)
    np.testing.assert_allclose(T2, T, atol=1e-14)
    (tmp_path / "bad.txt").write_text("1 2 3\n")
    with pytest.raises(ValueError, match="8 fields"):
        ev.read_tum(str(tmp_path / "bad.txt"))


@pytest.mark.parametrize("with_scale", [False, True])
def test_umeyama_recovers_known_transform(with_scale):
    rng = np.random.default_rng(2)
    src = rng.normal(size=(40, 3))
    R = Rotation.random(random_state=3).as_matrix()
    s = 2.5 if with_scale else 1.0
    t = np.array([0.3, -1.0, 4.0])
    dst = s * src @ R.T + t
    s2, R2, t2 = ev.align_umeyama(src, dst, with_scale)
    assert abs(s2 - s) < 1e-12
    np.testing.assert_allclose(R2, R, atol=1e-12)
    np.testing.assert_allclose(t2, t, atol=1e-12)


def test_umeyama_never_returns_a_reflection():
    """A cloud mirrored in x: the best orthogonal fit is that reflection, the
    best rotation is what scipy's Kabsch solver (align_vectors) returns."""
    rng = np.random.default_rng(4)
    src = rng.normal(size=(30, 3)) * np.array([3.0, 2.0, 1.0])
    dst = src * np.array([-1.0, 1.0, 1.0]) + rng.normal(0, 1e-3, (30, 3))
    _, R, _ = ev.align_umeyama(src, dst)
    assert abs(np.linalg.det(R) - 1.0) < 1e-12
    want = Rotation.align_vectors(dst - dst.mean(0), src - src.mean(0))[0].as_matrix()
    np.testing.assert_allclose(R, want, atol=1e-9)


def _associate_brute(a, b, max_diff, offset):
    """associate.py of the TUM RGB-D tools over index pairs."""
    pot = sorted((abs(a[i] - (b[j] + offset)), i, j) for i in range(len(a)) for j in range(len(b))
                 if abs(a[i] - (b[j] + offset)) < max_diff)
    ua, ub, out = set(), set(), []
    for d, i, j in pot:
        if i not in ua and j not in ub:
            ua.add(i)
            ub.add(j)
            out.append((i, j))
    return sorted(out)


@pytest.mark.parametrize("seed,offset", [(0, 0.0), (1, 0.05), (2, -0.013)])
def test_associate_matches_brute_force(seed, offset):
    rng = np.random.default_rng(seed)
    a = np.sort(rng.uniform(0, 5, 80))
    b = np.sort(np.concatenate([a[::2] + rng.normal(0, 0.01, 40) - offset, rng.uniform(0, 5, 30)]))
    got = ev.associate(a, b, 0.02, offset)
    assert [tuple(p) for p in got.tolist()] == _associate_brute(a, b, 0.02, offset)
    assert got.shape[0] > 20


def test_ate_alignment_cases(tmp_path):
    rng = np.random.default_rng(5)
    n = 60
    t = np.arange(n) * 0.01
    gt = _random_poses(rng, n)                        # camera->world
    # the estimate: world->camera poses of the ground truth in another frame
    # (similarity s, R, t), camera centres perturbed by 1 cm noise
    s, R, tt = 0.7, Rotation.random(random_state=6).as_matrix(), np.array([1.0, 2.0, -0.5])
    noise = rng.normal(0, 0.01, (n, 3))
    c_gt = gt[:, :3, 3]
    c_est = ((c_gt - tt) @ R) / s + noise            # inverse map of x -> s R x + t, plus noise
    est = np.tile(np.eye(4), (n, 1, 1))
    Rw2c = np.transpose(gt[:, :3, :3], (0, 2, 1))
    est[:, :3, :3] = Rw2c
    est[:, :3, 3] = -np.einsum("nij,nj->ni", Rw2c, c_est)
    _write_tum(str(tmp_path / "est.txt"), t, est)
    _write_tum(str(tmp_path / "gt.txt"), t, gt)
    r = ev.ate(str(tmp_path / "est.txt"), str(tmp_path / "gt.txt"), align="sim3")
    assert r["pairs"].shape[0] == n
    assert abs(r["scale"] - s) < 0.01 and 0.005 < r["rmse"] < 0.015
    exact = ev.ate((t, est), (t, gt), align="none")
    assert exact["rmse"] > 1.0
    noise_free = est.copy()
    noise_free[:, :3, 3] = -np.einsum("nij,nj->ni", Rw2c, ((c_gt - tt) @ R) / s)
    assert ev.ate((t, noise_free), (t, gt), align="sim3")["rmse"] < 1e-9
    assert ev.ate((t, noise_free), (t, gt), align="se3")["rmse"] > 0.1       # the scale is not rigid
    rigid = est.copy()
    rigid[:, :3, 3] = -np.einsum("nij,nj->ni", Rw2c, (c_gt - tt) @ R)
    assert ev.ate((t, rigid), (t, gt), align="se3")["rmse"] < 1e-9
    with pytest.raises(ValueError, match="no estimate timestamp"):
        ev.ate((t + 10.0, est), (t, gt))


def test_ate_without_alignment_equals_sequence_drift(tmp_path):
    rng = np.random.default_rng(7)
    n = 25
    gt = _random_poses(rng, n)                       # world->camera, as the sequence keeps them
    est = gt.copy()
    est[:, :3, 3] += rng.normal(0, 0.05, (n, 3))
    t = np.arange(n) * 0.01 + 0.005
    _write_tum(str(tmp_path / "e.txt"), t, est)
    _write_tum(str(tmp_path / "g.txt"), t, gt)
    r = ev.ate(str(tmp_path / "e.txt"), str(tmp_path / "g.txt"), gt_world_to_camera=True, align="none")
    e, _ = drift(torch.as_tensor(est), torch.as_tensor(gt))
    np.testing.assert_allclose(r["errors"], e.numpy(), rtol=1e-9, atol=1e-12)


def test_reads_the_sequence_tum_writer(tmp_path):
    """read_tum reads what SequenceResult.write_tum writes (the tracker's row
    format): mid-window timestamps and the world->camera poses."""
    from gsevt.sequence import SequenceResult
    rng = np.random.default_rng(8)
    est = _random_poses(rng, 12)
    res = SequenceResult(torch.as_tensor(est), [], [], 0, 1.0, 12)
    path = str(tmp_path / "seq.txt")
    res.write_tum(path, 0.01)
    t, T = ev.read_tum(path)
    np.testing.assert_allclose(t, np.arange(12) * 0.01 + 0.005, rtol=0, atol=1e-15)
    np.testing.assert_allclose(T, est, atol=1e-12)
    r = ev.ate(path, path, gt_world_to_camera=True)
    assert r["rmse"] < 1e-12 and r["pairs"].shape[0] == 12
