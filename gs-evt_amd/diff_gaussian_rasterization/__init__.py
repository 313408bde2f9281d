"""Drop-in ``diff_gaussian_rasterization`` for GS-EVT, backed by gfx950 HIP.

Same Python surface as the module the reference imports
(gaussian_splatting/gaussian_renderer/__init__.py:15-18):

* ``GaussianRasterizationSettings`` - NamedTuple built with keywords
  (fields used at __init__.py:58-72 and 318-337, velocity fields optional so
  the legacy ``render`` call still works).
* ``GaussianRasterizer(raster_settings)(means3D=, means2D=, shs=,
  colors_precomp=, opacities=, scales=, rotations=, cov3D_precomp=, theta=,
  rho=, w=, v=)`` returning ``(color [3,H,W], radii [P] int32,
  depth [1,H,W], opacity [1,H,W], n_touched [P] int32)``
  (__init__.py:362-375).
* ``GaussianRasterizer.markVisible(positions)``.

Backward returns gradients for every tensor input that requires one.  The
gradients of theta / rho (the camera-pose delta) and w / v (the velocity
delta) are the exact adjoint of the render pose
``T_r = Exp(delta_time [v; w]) Exp([rho; theta]) T_c``; they are computed and
reduced on the device.

Definitions of outputs the reference module leaves open (its source is not in
the reference tree): ``opacity`` = sum_i alpha_i T_i (= 1 - final T);
``depth`` = sum_i z_i alpha_i T_i (no background term); ``n_touched[i]`` =
number of pixels Gaussian i was blended into.
"""
from __future__ import annotations

from typing import NamedTuple, Optional

import torch
import torch.nn as nn

from gsevt import raster as _raster


class GaussianRasterizationSettings(NamedTuple):
    image_height: int
    image_width: int
    tanfovx: float
    tanfovy: float
    bg: torch.Tensor
    scale_modifier: float
    viewmatrix: torch.Tensor
    projmatrix: torch.Tensor
    projmatrix_raw: torch.Tensor
    sh_degree: int
    campos: torch.Tensor
    prefiltered: bool
    angular_vel: Optional[torch.Tensor] = None
    linear_vel: Optional[torch.Tensor] = None
    vel_transofrm: Optional[torch.Tensor] = None
    vel_transofrm_inv: Optional[torch.Tensor] = None
    delta_time: float = 0.0
    debug: bool = False
    tile_w: int = 16
    tile_h: int = 16


def _cam(s: GaussianRasterizationSettings) -> _raster.Camera:
    return _raster.Camera(viewmatrix=s.viewmatrix, projmatrix=s.projmatrix, projmatrix_raw=s.projmatrix_raw,
                          campos=s.campos, tanfovx=float(s.tanfovx), tanfovy=float(s.tanfovy),
                          width=int(s.image_width), height=int(s.image_height), bg=s.bg, tile_w=int(s.tile_w),
                          tile_h=int(s.tile_h))


class _RasterizeGaussians(torch.autograd.Function):
    @staticmethod
    def forward(ctx, means3D, means2D, sh, colors_precomp, opacities, scales, rotations, cov3Ds_precomp, theta,
                rho, w, v, raster_settings):
        gm = _raster.GaussianMap(means3D=means3D.detach(), opacities=opacities.detach(),
                                 shs=None if sh is None else sh.detach(),
                                 colors_precomp=None if colors_precomp is None else colors_precomp.detach(),
                                 scales=None if scales is None else scales.detach(),
                                 rotations=None if rotations is None else rotations.detach(),
                                 cov3D_precomp=None if cov3Ds_precomp is None else cov3Ds_precomp.detach(),
                                 sh_degree=int(raster_settings.sh_degree),
                                 scale_modifier=float(raster_settings.scale_modifier))
        cam = _cam(raster_settings)
        out, state = _raster.render_forward(gm, cam, channels=3, n_touched=True)
        ctx.gm = gm
        ctx.cam = cam
        ctx.state = state
        ctx.settings = raster_settings
        ctx.has = dict(shs=sh is not None, colors_precomp=colors_precomp is not None, scales=scales is not None,
                       rotations=rotations is not None, cov3D_precomp=cov3Ds_precomp is not None)
        ctx.mark_non_differentiable(out["radii"], out["n_touched"])
        return out["color"], out["radii"], out["depth"], out["opacity"], out["n_touched"]

    @staticmethod
    def backward(ctx, grad_color, grad_radii, grad_depth, grad_opacity, grad_n_touched):
        s = ctx.settings
        need = ctx.needs_input_grad
        names = ["means3D", "means2D", "shs", "colors_precomp", "opacities", "scales", "rotations",
                 "cov3D_precomp"]
        want = {n: bool(need[i]) for i, n in enumerate(names) if n != "means2D"}
        want["means2D"] = bool(need[1])
        gc = grad_color if grad_color is not None else torch.zeros_like(ctx.state["final_T"]).expand(3, -1, -1)
        pose12, g = _raster.render_backward(
            ctx.gm, ctx.cam, ctx.state, gc,
            None if grad_depth is None else grad_depth.reshape(ctx.cam.height, ctx.cam.width),
            None if grad_opacity is None else grad_opacity.reshape(ctx.cam.height, ctx.cam.width),
            vel_transform=s.vel_transofrm, delta_time=float(s.delta_time or 0.0), linear_vel=s.linear_vel,
            angular_vel=s.angular_vel, map_grads=want)
        res = [g.get(n) for n in names]
        res[4] = None if res[4] is None else res[4].reshape(ctx.gm.opacities.shape)
        theta = pose12[0:3] if need[8] else None
        rho = pose12[3:6] if need[9] else None
        wg = pose12[6:9] if need[10] else None
        vg = pose12[9:12] if need[11] else None
        return (*res, theta, rho, wg, vg, None)


def rasterize_gaussians(means3D, means2D, sh, colors_precomp, opacities, scales, rotations, cov3Ds_precomp,
                        theta, rho, w, v, raster_settings):
    return _RasterizeGaussians.apply(means3D, means2D, sh, colors_precomp, opacities, scales, rotations,
                                     cov3Ds_precomp, theta, rho, w, v, raster_settings)


class GaussianRasterizer(nn.Module):
    def __init__(self, raster_settings: GaussianRasterizationSettings):
        super().__init__()
        self.raster_settings = raster_settings

    def markVisible(self, positions: torch.Tensor) -> torch.Tensor:
        """Boolean mask of Gaussians not behind the near plane (False where
        view z <= 0.2), computed by the HIP kernel."""
        with torch.no_grad():
            s = self.raster_settings
            return _raster.mark_visible(positions, s.viewmatrix, s.projmatrix)

    def forward(self, means3D, means2D, opacities, shs=None, colors_precomp=None, scales=None, rotations=None,
                cov3D_precomp=None, theta=None, rho=None, w=None, v=None):
        s = self.raster_settings
        if (shs is None and colors_precomp is None) or (shs is not None and colors_precomp is not None):
            raise Exception("Please provide exactly one of either SHs or precomputed colors!")
        if ((scales is None or rotations is None) and cov3D_precomp is None) or (
                (scales is not None or rotations is not None) and cov3D_precomp is not None):
            raise Exception("Please provide exactly one of either scale/rotation pair or precomputed 3D covariance!")
        dev = means3D.device
        z3 = lambda t: torch.zeros(3, device=dev) if t is None else t
        return rasterize_gaussians(means3D, means2D, shs, colors_precomp, opacities, scales, rotations,
                                   cov3D_precomp, z3(theta), z3(rho), z3(w), z3(v), s)


__all__ = ["GaussianRasterizationSettings", "GaussianRasterizer", "rasterize_gaussians"]
