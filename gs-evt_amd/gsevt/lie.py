"""SO(3)/SE(3) maths of the tracking loop.

Restates utils/pose.py:13-91 of the reference: twist order [rho; theta],
Rodrigues exponential with the |theta| < 1e-5 series branch, the V matrix of
SE(3), and SO3_log with the acos clamp.  The reference branches with a Python
``if`` on a device scalar (pose.py:34,68), which forces a device-to-host sync
per call; these versions select the branch with ``torch.where`` so they never
synchronise, and produce the same values.
"""
from __future__ import annotations

import torch

SMALL_ANGLE = 1e-5


def skew(x: torch.Tensor) -> torch.Tensor:
    """[..., 3] -> [..., 3, 3] with skew(x) @ y = x cross y (pose.py:13-23)."""
    z = torch.zeros_like(x[..., 0])
    return torch.stack([
        torch.stack([z, -x[..., 2], x[..., 1]], -1),
        torch.stack([x[..., 2], z, -x[..., 0]], -1),
        torch.stack([-x[..., 1], x[..., 0], z], -1),
    ], -2)


def _eye(x: torch.Tensor) -> torch.Tensor:
    return torch.eye(3, device=x.device, dtype=x.dtype).expand(*x.shape[:-1], 3, 3)


def so3_exp(theta: torch.Tensor) -> torch.Tensor:
    """pose.py:26-41."""
    W = skew(theta)
    W2 = W @ W
    angle = torch.linalg.vector_norm(theta, dim=-1)[..., None, None]
    small = angle < SMALL_ANGLE
    a = torch.where(small, torch.ones_like(angle), angle)
    big = _eye(theta) + (torch.sin(a) / a) * W + ((1 - torch.cos(a)) / (a ** 2)) * W2
    tiny = _eye(theta) + W + 0.5 * W2
    return torch.where(small, tiny, big)


def so3_V(theta: torch.Tensor) -> torch.Tensor:
    """pose.py:60-75."""
    W = skew(theta)
    W2 = W @ W
    angle = torch.linalg.vector_norm(theta, dim=-1)[..., None, None]
    small = angle < SMALL_ANGLE
    a = torch.where(small, torch.ones_like(angle), angle)
    big = _eye(theta) + W * ((1.0 - torch.cos(a)) / (a ** 2)) + W2 * ((a - torch.sin(a)) / (a ** 3))
    tiny = _eye(theta) + 0.5 * W + (1.0 / 6.0) * W2
    return torch.where(small, tiny, big)


def se3_exp(xi: torch.Tensor) -> torch.Tensor:
    """[..., 6] twist [rho; theta] -> [..., 4, 4] (pose.py:78-91)."""
    rho, theta = xi[..., :3], xi[..., 3:]
    R = so3_exp(theta)
    t = (so3_V(theta) @ rho[..., None])[..., 0]
    T = torch.zeros(*xi.shape[:-1], 4, 4, device=xi.device, dtype=xi.dtype)
    T[..., :3, :3] = R
    T[..., :3, 3] = t
    T[..., 3, 3] = 1
    return T


def so3_log(R: torch.Tensor) -> torch.Tensor:
    """pose.py:44-57: rotation vector, zero when the angle is below 1e-5."""
    tr = R[..., 0, 0] + R[..., 1, 1] + R[..., 2, 2]
    c = torch.clamp((tr - 1) / 2.0, -1.0, 1.0)
    theta = torch.acos(c)
    small = theta.abs() < SMALL_ANGLE
    s = torch.where(small, torch.ones_like(theta), 2 * torch.sin(theta))
    axis = torch.stack([R[..., 2, 1] - R[..., 1, 2], R[..., 0, 2] - R[..., 2, 0], R[..., 1, 0] - R[..., 0, 1]], -1)
    out = theta[..., None] * axis / s[..., None]
    return torch.where(small[..., None], torch.zeros_like(out), out)


def rt_to_mat(R: torch.Tensor, t: torch.Tensor) -> torch.Tensor:
    T = torch.zeros(*R.shape[:-2], 4, 4, device=R.device, dtype=R.dtype)
    T[..., :3, :3] = R
    T[..., :3, 3] = t
    T[..., 3, 3] = 1
    return T
