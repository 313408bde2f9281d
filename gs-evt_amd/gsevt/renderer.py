"""Renderer glue: counterparts of render1 / render2 / build_rasterizer /
run_rasterizer (gaussian_splatting/gaussian_renderer/__init__.py:155-384).

They build GaussianRasterizationSettings from the Camera state exactly as
the reference does (transposed velocity transforms, delta_time = -/+ tau/2,
camera matrices from the Camera properties) and call the drop-in rasterizer.
"""
from __future__ import annotations

import math

import torch

from diff_gaussian_rasterization import GaussianRasterizationSettings, GaussianRasterizer


def build_rasterizer(cam, vel, vel_inv, delta_time, pc, bg, scaling_modifier=1.0, tile=(16, 16)):
    """__init__.py:313-339."""
    s = GaussianRasterizationSettings(
        image_height=int(cam.image_height), image_width=int(cam.image_width),
        tanfovx=math.tan(cam.FoVx * 0.5), tanfovy=math.tan(cam.FoVy * 0.5), bg=bg,
        scale_modifier=scaling_modifier, viewmatrix=cam.world_view_transform, projmatrix=cam.full_proj_transform,
        projmatrix_raw=cam.projection_matrix, sh_degree=pc.active_sh_degree, campos=cam.camera_center,
        prefiltered=False, angular_vel=cam.angular_vel, linear_vel=cam.linear_vel, vel_transofrm=vel,
        vel_transofrm_inv=vel_inv, delta_time=delta_time, debug=False, tile_w=tile[0], tile_h=tile[1])
    return GaussianRasterizer(raster_settings=s)


def run_rasterizer(rasterizer, means3D, means2D, shs, colors_precomp, opacity, scales, rotations, cov3D_precomp,
                   theta, rho, w, v, mask=None):
    """__init__.py:340-384.  With a mask, every per-Gaussian input is indexed
    by it and n_touched is None, as in the reference's mask branch (which
    unpacks four results where the rasterizer returns five; here the fifth is
    dropped, so the branch works)."""
    if mask is not None:
        sel = lambda t: None if t is None else t[mask]
        img, radii, depth, opac, _ = rasterizer(
            means3D=means3D[mask], means2D=means2D[mask], shs=sel(shs), colors_precomp=sel(colors_precomp),
            opacities=opacity[mask], scales=sel(scales), rotations=sel(rotations),
            cov3D_precomp=sel(cov3D_precomp), theta=theta, rho=rho, w=w, v=v)
        return {"render": img, "radii": radii, "depth": depth, "opacity": opac, "n_touched": None}
    img, radii, depth, opac, n_touched = rasterizer(
        means3D=means3D, means2D=means2D, shs=shs, colors_precomp=colors_precomp, opacities=opacity, scales=scales,
        rotations=rotations, cov3D_precomp=cov3D_precomp, theta=theta, rho=rho, w=w, v=v)
    return {"render": img, "radii": radii, "depth": depth, "opacity": opac, "n_touched": n_touched}


def _map_inputs(pc):
    scales = pc.get_scaling
    if scales.shape[-1] == 1:
        scales = scales.repeat(1, 3)
    return pc.get_xyz, pc.get_opacity, scales, pc.get_rotation, pc.get_features


def render2(last_cam, curr_cam, next_cam, pc, bg, scaling_modifier=1.0, tile=(16, 16)):
    """__init__.py:234-310."""
    if pc.get_xyz.shape[0] == 0:
        return None
    screenspace = torch.zeros_like(pc.get_xyz, dtype=pc.get_xyz.dtype, requires_grad=True,
                                   device=pc.get_xyz.device) + 0
    last_vel = curr_cam.last_vel_transform.t()
    last_vel_inv = curr_cam.last_vel_transform_inv.t()
    next_vel = curr_cam.next_vel_transform.t()
    next_vel_inv = curr_cam.next_vel_transform_inv.t()
    last_r = build_rasterizer(last_cam, last_vel, last_vel_inv, -curr_cam.delta_tau / 2, pc, bg, scaling_modifier,
                              tile)
    next_r = build_rasterizer(next_cam, next_vel, next_vel_inv, curr_cam.delta_tau / 2, pc, bg, scaling_modifier,
                              tile)
    means3D, opacity, scales, rotations, shs = _map_inputs(pc)
    th, rh, w, v = curr_cam.cam_rot_delta, curr_cam.cam_trans_delta, curr_cam.cam_w_delta, curr_cam.cam_v_delta
    last = run_rasterizer(last_r, means3D, screenspace, shs, None, opacity, scales, rotations, None, th, rh, w, v)
    nxt = run_rasterizer(next_r, means3D, screenspace, shs, None, opacity, scales, rotations, None, th, rh, w, v)
    return last, nxt


def render1(cam, pc, bg, scaling_modifier=1.0, tile=(16, 16), mask=None):
    """__init__.py:155-231: single view, identity velocity transform, delta_time 0
    (mask: render only the selected Gaussians, n_touched None)."""
    if pc.get_xyz.shape[0] == 0:
        return None
    screenspace = torch.zeros_like(pc.get_xyz, dtype=pc.get_xyz.dtype, requires_grad=True,
                                   device=pc.get_xyz.device) + 0
    eye = torch.eye(4, device=pc.get_xyz.device, dtype=pc.get_xyz.dtype)
    r = build_rasterizer(cam, eye, eye, 0, pc, bg, scaling_modifier, tile)
    means3D, opacity, scales, rotations, shs = _map_inputs(pc)
    return run_rasterizer(r, means3D, screenspace, shs, None, opacity, scales, rotations, None, cam.cam_rot_delta,
                          cam.cam_trans_delta, cam.cam_w_delta, cam.cam_v_delta, mask=mask)
