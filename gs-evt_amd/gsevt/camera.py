"""Camera state of the tracking loop and the matrices the rasterizer reads.

Mirrors utils/render_camera/camera.py:9-218 and
gaussian_splatting/utils/graphics_utils.py:33-69,100-101 of the reference:
world-to-camera (R, T), linear/angular velocity, the four optimised deltas
(cam_rot_delta theta, cam_trans_delta rho, cam_w_delta, cam_v_delta), the
symmetric-frustum projection (principal point ignored, camera.py:74-75) and the
left-multiplicative pose update (camera.py:133-147).  Matrices are returned in
the transposed storage the rasterizer settings carry.
"""
from __future__ import annotations

import math

import torch
from torch import nn

from .lie import se3_exp, so3_log, rt_to_mat

ZNEAR = 0.01
ZFAR = 100.0


def focal2fov(focal: float, pixels: int) -> float:
    """graphics_utils.py:100-101."""
    return 2 * math.atan(pixels / (2 * focal))


def fov2focal(fov: float, pixels: int) -> float:
    return pixels / (2 * math.tan(fov / 2))


def projection_matrix(znear: float, zfar: float, fovx: float, fovy: float) -> torch.Tensor:
    """graphics_utils.py:49-69 (values computed in Python float, stored f32)."""
    tan_y = math.tan(fovy / 2)
    tan_x = math.tan(fovx / 2)
    top = tan_y * znear
    bottom = -top
    right = tan_x * znear
    left = -right
    P = torch.zeros(4, 4)
    P[0, 0] = 2.0 * znear / (right - left)
    P[1, 1] = 2.0 * znear / (top - bottom)
    P[0, 2] = (right + left) / (right - left)
    P[1, 2] = (top + bottom) / (top - bottom)
    P[3, 2] = 1.0
    P[2, 2] = zfar / (zfar - znear)
    P[2, 3] = -(zfar * znear) / (zfar - znear)
    return P


def world2view(R: torch.Tensor, t: torch.Tensor, exact_reference_ops: bool = True) -> torch.Tensor:
    """graphics_utils.py:33-46.  The reference inverts twice (translate = 0,
    scale = 1), which is the identity up to rounding; exact_reference_ops keeps
    those two inversions so the plain path computes what the reference does."""
    Rt = torch.zeros((4, 4), device=R.device, dtype=R.dtype)
    Rt[:3, :3] = R
    Rt[:3, 3] = t
    Rt[3, 3] = 1.0
    if not exact_reference_ops:
        return Rt
    C2W = torch.linalg.inv(Rt)
    return torch.linalg.inv(C2W)


class Camera(nn.Module):
    """utils/render_camera/camera.py:9-218."""

    def __init__(self, R, t, angular_vel, linear_vel, fovx, fovy, image_width, image_height, delta_tau=0,
                 device="cuda"):
        super().__init__()
        self.device = torch.device(device)
        self.R = R.to(device)
        self.T = t.to(device)
        self.angular_vel = angular_vel
        self.linear_vel = linear_vel
        self.FoVx = fovx
        self.FoVy = fovy
        self.image_width = image_width
        self.image_height = image_height
        self.zfar = ZFAR
        self.znear = ZNEAR
        self.delta_tau = delta_tau
        self.cam_rot_delta = nn.Parameter(torch.zeros(3, device=self.device), requires_grad=True)
        self.cam_trans_delta = nn.Parameter(torch.zeros(3, device=self.device), requires_grad=True)
        self.cam_w_delta = nn.Parameter(torch.zeros(3, device=self.device), requires_grad=False)
        self.cam_v_delta = nn.Parameter(torch.zeros(3, device=self.device), requires_grad=False)

    # camera.py:68-91
    @property
    def projection_matrix(self):
        return projection_matrix(self.znear, self.zfar, self.FoVx, self.FoVy).transpose(0, 1).to(self.device)

    @property
    def world_view_transform(self):
        return world2view(self.R, self.T).transpose(0, 1)

    @property
    def full_proj_transform(self):
        return (self.world_view_transform.unsqueeze(0).bmm(self.projection_matrix.unsqueeze(0))).squeeze(0)

    @property
    def camera_center(self):
        return self.world_view_transform.inverse()[3, :3]

    @property
    def curr_pose(self):
        return rt_to_mat(self.R, self.T)

    # camera.py:100-123
    def compute_motion_vectors(self):
        return self.angular_vel * (self.delta_tau / 2), self.linear_vel * (self.delta_tau / 2)

    @property
    def last_vel_transform(self):
        rot, trans = self.compute_motion_vectors()
        return se3_exp(torch.cat([-trans, -rot], 0))

    @property
    def last_vel_transform_inv(self):
        return torch.linalg.inv(self.last_vel_transform)

    @property
    def next_vel_transform(self):
        rot, trans = self.compute_motion_vectors()
        return se3_exp(torch.cat([trans, rot], 0))

    @property
    def next_vel_transform_inv(self):
        return torch.linalg.inv(self.next_vel_transform)

    # camera.py:125-155
    def update_RT(self, R, t):
        self.R = R.to(device=self.device)
        self.T = t.to(device=self.device)

    def update_vwRT(self):
        self.update_velocity()
        self.update_pose()

    def update_pose(self):
        delta = torch.cat([self.cam_trans_delta, self.cam_rot_delta], 0)
        new = se3_exp(delta) @ rt_to_mat(self.R, self.T)
        self.update_RT(new[0:3, 0:3], new[0:3, 3])
        self.cam_rot_delta.data.fill_(0)
        self.cam_trans_delta.data.fill_(0)

    def update_velocity(self):
        self.angular_vel += self.cam_w_delta
        self.linear_vel += self.cam_v_delta
        self.cam_w_delta.data.fill_(0)
        self.cam_v_delta.data.fill_(0)

    # camera.py:157-181
    def cal_weighted_velocity(self, last_data, delta_tau, weight):
        last_T, last_R = last_data[0], last_data[1]
        last_pose = rt_to_mat(last_R, last_T)
        curr_pose = rt_to_mat(self.R.detach(), self.T.detach())
        delta = curr_pose @ torch.inverse(last_pose)
        lin = delta[0:3, 3] / delta_tau
        ang = so3_log(delta[0:3, 0:3]) / delta_tau
        self.linear_vel = weight * lin + (1 - weight) * self.linear_vel
        self.angular_vel = weight * ang + (1 - weight) * self.angular_vel

    # camera.py:183-201
    def const_vel_model(self, tau):
        xi = torch.cat([self.linear_vel * tau, self.angular_vel * tau], 0)
        new = se3_exp(xi) @ rt_to_mat(self.R, self.T)
        self.last_R = self.R.clone()
        self.last_T = self.T.clone()
        self.update_RT(new[:3, :3], new[:3, 3])

    @staticmethod
    def init_from_yaml(config):
        """camera.py:203-218 (plain dict config)."""
        g = config["Gaussian"]
        W, H = g["img_width"], g["img_height"]
        device = g["model_params"]["device"]
        fx, fy = g["calib_params"]["fx"], g["calib_params"]["fy"]
        R = torch.tensor(config["Tracking"]["initial_pose"]["rot"]["data"]).reshape(3, 3)
        t = torch.tensor(config["Tracking"]["initial_pose"]["trans"]["data"]).reshape(3)
        lin = torch.tensor(config["Tracking"]["initial_vel"]["linear_vel"], device=device, dtype=torch.float32)
        ang = torch.tensor(config["Tracking"]["initial_vel"]["angular_vel"], device=device, dtype=torch.float32)
        cam = Camera(R, t, ang, lin, focal2fov(fx, W), focal2fov(fy, H), W, H, device=device)
        cam.fx = fx
        cam.fy = fy
        cam.update_pose()
        return cam


def render_matrices(R: torch.Tensor, t: torch.Tensor, fovx: float, fovy: float):
    """Direct (inversion-free) viewmatrix / projmatrix / projmatrix_raw /
    campos for a world-to-camera pose, in settings storage.  Mathematically
    equal to Camera's properties; used by the fused tracker and tests."""
    Rt = rt_to_mat(R, t)
    view = Rt.transpose(-1, -2)
    praw = projection_matrix(ZNEAR, ZFAR, fovx, fovy).to(R.device, R.dtype).transpose(0, 1)
    proj = view @ praw
    campos = -(R.transpose(-1, -2) @ t[..., None])[..., 0]
    return view, proj, praw, campos
