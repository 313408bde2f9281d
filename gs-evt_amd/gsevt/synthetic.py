"""Seeded synthetic scenes, cameras and event windows for the tracking
workloads (numpy PCG64, generated in float64 and cast to float32).

Workloads (all cameras world->camera = identity, principal point at the image
centre, which is what the reference render uses: camera.py:74-75):
  cpu_small     20k Gaussians uniform in a 2 m cube 3 m ahead, SH0 grey
                intensity, 346x260, fx = fy = 300, 8k uniform events
  patches_300k  300k anisotropic Gaussians on 64 planar patches, SH3,
                640x480, fx = fy = 550, 200k events per 10 ms window
  patches_1m    same generator at 1M Gaussians (gradient, batched and
                sequence workloads)
"""
from __future__ import annotations

import hashlib
import math
from dataclasses import dataclass

import numpy as np

SH_C0 = 0.28209479177387814

SEED_SCENE = 1000
SEED_EVENTS = 2000
SEED_INIT = 3000
SEED_HYP = 4000
SEED_TRAJ = 5000
SEED_FD = 6000


@dataclass
class RawScene:
    """Pre-activation parameters in the layout of gaussian_model.py:298-323."""
    xyz: np.ndarray            # [P,3]
    features_dc: np.ndarray    # [P,1,3]
    features_rest: np.ndarray  # [P,K-1,3]
    scaling: np.ndarray        # [P,3] log scales
    rotation: np.ndarray       # [P,4] unnormalised (w,x,y,z)
    opacity: np.ndarray        # [P,1] logits
    sh_degree: int

    @property
    def P(self):
        return self.xyz.shape[0]

    def activated(self):
        """numpy activations (float32), same formulas as the model."""
        q = self.rotation.astype(np.float64)
        q = q / np.maximum(np.linalg.norm(q, axis=1, keepdims=True), 1e-12)
        feats = np.concatenate([self.features_dc, self.features_rest], axis=1)
        return dict(
            means=self.xyz.astype(np.float32),
            scales=np.exp(self.scaling.astype(np.float64)).astype(np.float32),
            rots=q.astype(np.float32),
            opac=(1.0 / (1.0 + np.exp(-self.opacity[:, 0].astype(np.float64)))).astype(np.float32),
            shs=np.ascontiguousarray(feats.astype(np.float32)),
        )

    def sha256(self) -> str:
        h = hashlib.sha256()
        for a in (self.xyz, self.features_dc, self.features_rest, self.scaling, self.rotation, self.opacity):
            h.update(np.ascontiguousarray(a).tobytes())
        return h.hexdigest()


@dataclass
class Intrinsics:
    W: int
    H: int
    fx: float
    fy: float

    @property
    def fovx(self):
        return 2 * math.atan(self.W / (2 * self.fx))

    @property
    def fovy(self):
        return 2 * math.atan(self.H / (2 * self.fy))


CPU_SMALL_CAM = Intrinsics(346, 260, 300.0, 300.0)
PATCH_CAM = Intrinsics(640, 480, 550.0, 550.0)


def _random_quats(rng, P):
    q = rng.standard_normal((P, 4))
    return q / np.linalg.norm(q, axis=1, keepdims=True)


def cube_scene(P=20000, seed=SEED_SCENE + 1, sh_degree=0):
    """Means ~ U([-1,1]^3) + (0,0,3); log-scales N(-3.5, 0.5); random unit
    quaternions; opacity logits N(0,1); SH0 grey intensity U(0,1)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    xyz = rng.uniform(-1.0, 1.0, (P, 3)) + np.array([0.0, 0.0, 3.0])
    scaling = rng.normal(-3.5, 0.5, (P, 3))
    rot = _random_quats(rng, P)
    opacity = rng.normal(0.0, 1.0, (P, 1))
    inten = rng.uniform(0.0, 1.0, P)
    K = (sh_degree + 1) ** 2
    dc = np.repeat(((inten - 0.5) / SH_C0)[:, None, None], 3, axis=2)
    rest = np.zeros((P, K - 1, 3))
    f = np.float32
    return RawScene(xyz.astype(f), dc.astype(f), rest.astype(f), scaling.astype(f), rot.astype(f),
                    opacity.astype(f), sh_degree)


def patch_scene(P=300000, seed=SEED_SCENE + 2, sh_degree=3, n_patches=64, extent=0.5, jitter=0.01,
                log_scale_mean=-4.0, log_scale_std=0.7, sh_std=0.2):
    """Gaussians on 64 planar patches (centres uniform in the 2 m cube 3 m
    ahead, normals uniform on S^2, 0.5 x 0.5 m, N(0, 1 cm) normal jitter,
    round-robin assignment); log-scales N(-4, 0.7); all SH coefficients
    N(0, 0.2); opacity logits N(0, 1)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    centres = rng.uniform(-1.0, 1.0, (n_patches, 3)) + np.array([0.0, 0.0, 3.0])
    normals = rng.standard_normal((n_patches, 3))
    normals /= np.linalg.norm(normals, axis=1, keepdims=True)
    # orthonormal in-plane axes
    helper = np.where(np.abs(normals[:, :1]) < 0.9, np.array([[1.0, 0, 0]]), np.array([[0, 1.0, 0]]))
    u = np.cross(normals, helper)
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    v = np.cross(normals, u)
    pid = np.arange(P) % n_patches
    a = rng.uniform(-extent / 2, extent / 2, P)
    b = rng.uniform(-extent / 2, extent / 2, P)
    h = rng.normal(0.0, jitter, P)
    xyz = centres[pid] + a[:, None] * u[pid] + b[:, None] * v[pid] + h[:, None] * normals[pid]
    scaling = rng.normal(log_scale_mean, log_scale_std, (P, 3))
    rot = _random_quats(rng, P)
    opacity = rng.normal(0.0, 1.0, (P, 1))
    K = (sh_degree + 1) ** 2
    sh = rng.normal(0.0, sh_std, (P, K, 3))
    f = np.float32
    return RawScene(xyz.astype(f), sh[:, :1].astype(f), sh[:, 1:].astype(f), scaling.astype(f), rot.astype(f),
                    opacity.astype(f), sh_degree)


def small_test_scene(P=200, seed=7, sh_degree=3, W=64, H=48, spread=0.6, log_scale_mean=-3.0):
    """Tiny scene for oracle/brute-force and finite-difference tests."""
    rng = np.random.Generator(np.random.PCG64(seed))
    xyz = rng.uniform(-spread, spread, (P, 3)) * np.array([1.0, 0.75, 0.5]) + np.array([0.0, 0.0, 2.5])
    scaling = rng.normal(log_scale_mean, 0.4, (P, 3))
    rot = _random_quats(rng, P)
    opacity = rng.normal(0.0, 1.0, (P, 1))
    K = (sh_degree + 1) ** 2
    sh = rng.normal(0.0, 0.3, (P, K, 3))
    f = np.float32
    return RawScene(xyz.astype(f), sh[:, :1].astype(f), sh[:, 1:].astype(f), scaling.astype(f), rot.astype(f),
                    opacity.astype(f), sh_degree)


def random_twist(rng, trans_sigma, rot_sigma):
    """Perturbation [rho; theta] with N(0, sigma^2) components."""
    return np.concatenate([rng.normal(0, trans_sigma, 3), rng.normal(0, rot_sigma, 3)])


def fixed_norm_twist(rng, trans_norm, rot_norm):
    """[rho; theta] with |rho| = trans_norm, |theta| = rot_norm, random directions."""
    a = rng.standard_normal(3)
    b = rng.standard_normal(3)
    return np.concatenate([a / np.linalg.norm(a) * trans_norm, b / np.linalg.norm(b) * rot_norm])


def uniform_events(n, W, H, duration_us=10000, seed=SEED_EVENTS + 1, t0_us=0):
    """n events: t ~ U(0, 10 ms) sorted (integer microseconds), x, y uniform,
    polarity Bernoulli(1/2) in {0, 1}."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ts = np.sort(rng.integers(0, duration_us, n)) + t0_us
    x = rng.integers(0, W, n)
    y = rng.integers(0, H, n)
    p = rng.integers(0, 2, n)
    return dict(ts=ts.astype(np.int64), x=x.astype(np.int32), y=y.astype(np.int32), p=p.astype(np.int32))


def events_from_delta(delta_I: np.ndarray, n, noise_frac=0.1, duration_us=10000, seed=SEED_EVENTS + 2,
                      t0_us=0):
    """Events drawn from an intensity-change image: (1 - noise_frac) n
    pixels sampled with probability proportional to |dI|, polarity = sign(dI);
    noise_frac n uniform noise events with random polarity."""
    rng = np.random.Generator(np.random.PCG64(seed))
    H, W = delta_I.shape
    w = np.abs(delta_I.astype(np.float64)).reshape(-1)
    n_sig = int(round(n * (1 - noise_frac)))
    n_noise = n - n_sig
    if w.sum() <= 0:
        n_noise, n_sig = n, 0
    idx = rng.choice(W * H, size=n_sig, p=w / w.sum()) if n_sig else np.zeros(0, np.int64)
    pol = (delta_I.reshape(-1)[idx] > 0).astype(np.int32)
    nidx = rng.integers(0, W * H, n_noise)
    npol = rng.integers(0, 2, n_noise).astype(np.int32)
    allidx = np.concatenate([idx, nidx])
    allpol = np.concatenate([pol, npol])
    order = rng.permutation(n)
    allidx, allpol = allidx[order], allpol[order]
    ts = np.sort(rng.integers(0, duration_us, n)) + t0_us
    return dict(ts=ts.astype(np.int64), x=(allidx % W).astype(np.int32), y=(allidx // W).astype(np.int32),
                p=allpol)
