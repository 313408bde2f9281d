"""Frozen Gaussian map as the tracker feeds it to the rasterizer.

Mirrors the activations of gaussian_splatting/scene/gaussian_model.py:52-60,
75-95 (exp scales, normalised (w,x,y,z) quaternion, sigmoid opacity,
features = cat(f_dc [P,1,3], f_rest [P,K-1,3]) -> [P,K,3]).  The map is frozen
during tracking (gaussian_model.py:298-323 create every tensor with
requires_grad False), so the activated tensors are computed once and cached
instead of on every property access.
"""
from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F


class GaussianModel:
    def __init__(self, sh_degree: int):
        self.active_sh_degree = sh_degree
        self.max_sh_degree = sh_degree
        self._cache = None

    @classmethod
    def from_raw(cls, xyz, features_dc, features_rest, scaling, rotation, opacity, sh_degree, device="cuda"):
        m = cls(sh_degree)
        t = lambda a: torch.as_tensor(np.asarray(a), dtype=torch.float32, device=device).contiguous()
        m._xyz = t(xyz)
        m._features_dc = t(features_dc)
        m._features_rest = t(features_rest)
        m._scaling = t(scaling)
        m._rotation = t(rotation)
        m._opacity = t(opacity)
        return m

    def _activate(self):
        if self._cache is None:
            self._cache = dict(
                xyz=self._xyz,
                scaling=torch.exp(self._scaling),
                rotation=F.normalize(self._rotation),
                opacity=torch.sigmoid(self._opacity),
                features=torch.cat((self._features_dc, self._features_rest), dim=1).contiguous(),
            )
        return self._cache

    @property
    def get_xyz(self):
        return self._activate()["xyz"]

    @property
    def get_scaling(self):
        return self._activate()["scaling"]

    @property
    def get_rotation(self):
        return self._activate()["rotation"]

    @property
    def get_opacity(self):
        return self._activate()["opacity"]

    @property
    def get_features(self):
        return self._activate()["features"]

    @property
    def num_points(self):
        return self._xyz.shape[0]
