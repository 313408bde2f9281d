"""CPU tests of the host-side mirror (Lie maths, camera matrices, tracker
control flow, Adam) and of the C ABI surface (no GPU compute)."""
import ctypes
import os
import pathlib
import re

import numpy as np
import pytest
import torch
from scipy.linalg import expm

from helpers import ROOT, skew

from gsevt import lie
from gsevt.camera import Camera, focal2fov, projection_matrix, render_matrices
from gsevt.tracker import check_convergence, lr_fraction, tracking_loss
from oracle import tracking_step


def _twist_matrix(xi):
    M = np.zeros((4, 4))
    M[:3, :3] = skew(xi[3:])
    M[:3, 3] = xi[:3]
    return M


@pytest.mark.parametrize("scale", [0.0, 1e-7, 1e-4, 0.3, 2.0])
def test_se3_exp_matches_matrix_exponential(scale):
    rng = np.random.default_rng(int(scale * 1e7) % 1000)
    xi = rng.normal(0, 1, 6) * scale
    got = lie.se3_exp(torch.tensor(xi, dtype=torch.float64)).numpy()
    np.testing.assert_allclose(got, expm(_twist_matrix(xi)), atol=1e-10)


def test_so3_log_inverts_exp_and_small_angle_is_zero():
    rng = np.random.default_rng(3)
    for _ in range(20):
        th = rng.normal(0, 0.8, 3)
        R = lie.so3_exp(torch.tensor(th, dtype=torch.float64))
        np.testing.assert_allclose(lie.so3_log(R).numpy(), th, atol=1e-9)
    R = lie.so3_exp(torch.tensor([1e-7, 0, 0], dtype=torch.float64))
    assert torch.all(lie.so3_log(R) == 0)      # pose.py:48-49


def test_batched_lie_matches_single():
    xi = torch.randn(5, 6, dtype=torch.float64) * 0.5
    batch = lie.se3_exp(xi)
    for k in range(5):
        torch.testing.assert_close(batch[k], lie.se3_exp(xi[k]))


def test_projection_matrix_closed_form():
    fx, W = 550.0, 640
    fov = focal2fov(fx, W)
    P = projection_matrix(0.01, 100.0, fov, fov).double().numpy()
    # x_ndc = P00 x / z with P00 = 2 f / W; z_ndc maps near -> 0, far -> 1
    assert abs(P[0, 0] - 2 * fx / W) < 1e-6
    z = np.array([0.01, 100.0])
    zn = (P[2, 2] * z + P[2, 3]) / z
    np.testing.assert_allclose(zn, [0.0, 1.0], atol=1e-6)
    assert P[0, 2] == 0 and P[1, 2] == 0 and P[3, 2] == 1


def test_camera_matrices_and_pose_update():
    rng = np.random.default_rng(5)
    T = expm(_twist_matrix(rng.normal(0, 0.3, 6)))
    R = torch.tensor(T[:3, :3], dtype=torch.float32)
    t = torch.tensor(T[:3, 3], dtype=torch.float32)
    cam = Camera(R, t, torch.zeros(3), torch.zeros(3), 1.0, 0.8, 64, 48, device="cpu")
    wv = cam.world_view_transform.double().numpy()
    np.testing.assert_allclose(wv.T, T, atol=1e-5)
    np.testing.assert_allclose(cam.camera_center.double().numpy(), -T[:3, :3].T @ T[:3, 3], atol=1e-5)
    v, p, praw, c = render_matrices(R, t, 1.0, 0.8)
    np.testing.assert_allclose(p.numpy(), cam.full_proj_transform.numpy(), atol=1e-5)
    # left-multiplicative update (camera.py:133-147)
    d = np.array([0.01, -0.02, 0.03, 0.02, 0.01, -0.01])
    cam.cam_trans_delta.data[:] = torch.tensor(d[:3], dtype=torch.float32)
    cam.cam_rot_delta.data[:] = torch.tensor(d[3:], dtype=torch.float32)
    with torch.no_grad():        # the tracker updates under no_grad (tracker.py:215-222)
        cam.update_pose()
    Tn = np.eye(4)
    Tn[:3, :3] = cam.R.numpy()
    Tn[:3, 3] = cam.T.numpy()
    np.testing.assert_allclose(Tn, expm(_twist_matrix(d)) @ T, atol=1e-5)
    assert torch.all(cam.cam_rot_delta == 0) and torch.all(cam.cam_trans_delta == 0)


def test_velocity_transforms_and_weighted_velocity():
    lin = torch.tensor([0.3, -0.1, 0.2])
    ang = torch.tensor([0.1, 0.2, -0.3])
    cam = Camera(torch.eye(3), torch.zeros(3), ang.clone(), lin.clone(), 1.0, 0.8, 64, 48, delta_tau=0.01,
                 device="cpu")
    xi = np.concatenate([lin.numpy(), ang.numpy()]) * 0.005
    np.testing.assert_allclose(cam.next_vel_transform.double().numpy(), expm(_twist_matrix(xi)), atol=1e-6)
    np.testing.assert_allclose(cam.last_vel_transform.double().numpy(), expm(_twist_matrix(-xi)), atol=1e-6)
    # const_vel_model then cal_weighted_velocity with the same tau recovers the velocity
    last = [cam.T.clone(), cam.R.clone()]
    cam.const_vel_model(0.01)
    cam.cal_weighted_velocity(last, 0.01, 0.5)
    np.testing.assert_allclose(cam.angular_vel.numpy(), ang.numpy(), atol=1e-4)


def test_lr_fraction_quirk_and_ramp():
    # tracker.py:188-191: fraction = 1 at it - start = 0 (velocity LR 0), ramps 1/100 .. 1
    assert lr_fraction(0, 0, 200) == 1
    assert lr_fraction(1, 0, 200) == pytest.approx(0.01)
    assert lr_fraction(100, 0, 200) == pytest.approx(1.0)
    assert lr_fraction(101, 0, 200) == 1
    assert lr_fraction(37, 30, 200) == pytest.approx(0.07)


def test_check_convergence():
    assert not check_convergence([1.0] * 10)
    assert check_convergence([1.0] * 11, 1e-4)
    assert not check_convergence(list(np.linspace(1, 0, 11)), 1e-4)
    assert check_convergence(list(1 + 1e-6 * np.arange(20)), 1e-4)


def test_tracking_loss_and_its_gradient_restatement():
    rng = np.random.default_rng(0)
    cl = rng.uniform(0, 1, (3, 12, 10))
    cn = rng.uniform(0, 1, (3, 12, 10))
    e = rng.normal(0, 0.1, (12, 10))
    L, gl, gn, r = tracking_step.delta_ir_and_grad(cl, cn, e)
    tl = torch.tensor(cl, requires_grad=True)
    tn = torch.tensor(cn, requires_grad=True)
    w = torch.tensor([0.2989, 0.5870, 0.1140], dtype=torch.float64).view(3, 1, 1)
    d = (tn * w).sum(0) - (tl * w).sum(0)
    lt = tracking_loss(d / torch.norm(d), torch.tensor(e))
    lt.backward()
    assert abs(float(lt.detach()) - L) < 1e-12
    np.testing.assert_allclose(gl, tl.grad.numpy(), atol=1e-12)
    np.testing.assert_allclose(gn, tn.grad.numpy(), atol=1e-12)


def test_adam_restatement_matches_torch():
    rng = np.random.default_rng(1)
    lrs = dict(theta=0.004, rho=0.004, w=0.002, v=0.002)
    params = {k: torch.zeros(3, dtype=torch.float64, requires_grad=True) for k in lrs}
    opt = torch.optim.Adam([{"params": [params[k]], "lr": lrs[k]} for k in lrs])
    mine = tracking_step.Adam(lrs)
    cur = {k: np.zeros(3) for k in lrs}
    for it in range(12):
        g = {k: rng.normal(0, 1, 3) for k in lrs}
        if it < 3:
            g["w"] = g["v"] = None   # coarse stage: no velocity gradient
        for k in lrs:
            params[k].grad = None if g[k] is None else torch.tensor(g[k])
        opt.step()
        cur = mine.step(cur, g)
        for k in lrs:
            np.testing.assert_allclose(cur[k], params[k].detach().numpy(), atol=1e-12)
            cur[k] = np.zeros(3)
            params[k].data.zero_()   # deltas are reset after every step (camera.py:146-155)


def test_events_scatter_oracle_and_frame_normalisation():
    from oracle import events as oev
    rng = np.random.default_rng(2)
    W, H, n = 40, 30, 5000
    x, y, p = rng.integers(0, W, n), rng.integers(0, H, n), rng.integers(0, 2, n)
    c = oev.scatter(x, y, p, W, H)
    ref = np.zeros((H, W), np.int64)
    for xi, yi, pi in zip(x, y, p):
        ref[yi, xi] += 1 if pi else -1
    np.testing.assert_array_equal(c, ref)
    K = np.array([[35.0, 0, 20.0], [0, 35.0, 15.0], [0, 0, 1]])
    # zero distortion: the remap is the identity
    np.testing.assert_array_equal(oev.undistort(c.astype(np.float32), K, np.zeros(5)), c.astype(np.float32))
    k = oev.gaussian_kernel(9, 0.0)
    assert abs(k.sum() - 1) < 1e-6 and np.all(k[:4] == k[::-1][:4])
    s, a = oev.event_frame(x, y, p, W, H, K, np.array([-0.03, 0.04, -0.0005, -0.001, 0.0]))
    assert abs(np.linalg.norm(s) - 1) < 1e-5
    np.testing.assert_array_equal(a, np.abs(s))


def test_ctypes_signatures_match_header():
    """Every ctypes signature the Python side binds (the plain table and the
    batched tracker's) has the header declaration's argument count, and each
    argument its kind: pointer, int, long long, float, double or uint32_t."""
    import ctypes
    from gsevt import _lib, fused
    hdr = pathlib.Path(ROOT, "include", "gsevt.h").read_text()
    body = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    decls = {name: params for name, params in
             re.findall(r"\b(gsevt_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", body)}
    kinds = {"int": {ctypes.c_int}, "long long": {ctypes.c_longlong}, "float": {ctypes.c_float},
             "double": {ctypes.c_double}, "uint32_t": {ctypes.c_uint32, ctypes.c_uint}}
    tables = dict(_lib.SIGNATURES)
    tables.update(fused._EXTRA)
    checked = 0
    for name, (_, args) in tables.items():
        assert name in decls, f"{name} is bound but not declared"
        params = decls[name].strip()
        ps = [] if params in ("", "void") else [q.strip() for q in params.split(",")]
        assert len(args) == len(ps), f"{name}: {len(args)} ctypes arguments, {len(ps)} declared"
        for i, (a, q) in enumerate(zip(args, ps)):
            if "*" in q:
                ok = a in (ctypes.c_void_p, ctypes.c_char_p) or issubclass(a, ctypes._Pointer)
            else:
                ok = a in kinds[re.sub(r"\s+\w+$", "", q)]
            assert ok, f"{name} argument {i} ({q}) bound as {a.__name__}"
        checked += 1
    assert checked > 40


def test_c_abi_exports_every_declared_symbol():
    hdr = pathlib.Path(ROOT, "include", "gsevt.h").read_text()
    body = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    names = sorted(set(re.findall(r"\b(gsevt_[a-z0-9_]+)\s*\(", body)))
    assert len(names) > 30
    lib = os.path.join(ROOT, "gs-evt_amd", "gsevt", "libgsevt_hip.so")
    if not os.path.exists(lib):
        import __graft_entry__
        __graft_entry__.build()
    L = ctypes.CDLL(lib)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert L.gsevt_abi_version() == 1
    # argument errors are reported, not crashed on (no GPU work is launched)
    from gsevt import _lib
    Lp = _lib.load()
    rc = Lp.gsevt_preprocess(10, 5, 16, None, None, 1.0, None, None, None, None, None, None, None, None, 64, 48,
                             0.5, 0.5, 16, 16, 3, None, None, None, None, None)
    assert rc != 0 and b"sh_degree" in Lp.gsevt_last_error()
    rc = Lp.gsevt_blend_forward(64, 48, 10, 10, None, None, None, 3, None, None, None, None, None, None, None, None)
    assert rc != 0 and b"tile" in Lp.gsevt_last_error()
    # the allocation-free event frame rejects frames beyond its static
    # scratch (1280 x 1024 pixels) before touching a device; the pointers are
    # never dereferenced
    fake = ctypes.c_void_p(256)
    rc = Lp.gsevt_event_frame_noalloc(1281, 1024, fake, fake, 9, 0.0, fake, None, None)
    assert rc != 0 and b"static scratch limit" in Lp.gsevt_last_error()
