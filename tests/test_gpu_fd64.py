"""Gradients of the tracking loss against FP64 central differences at full
size (the differentiation configuration: 1M Gaussians, SH degree 3,
640x480, 32 seeded poses).

The differences come from the GPU FP64 engine in oracle/fd64.hip, whose
discrete decisions are frozen at the base point (taken in FP64, or those of
the FP32 forward being differentiated).  It is pinned first at small
size: its record-mode render equals the CPU FP64 oracle's render, its frozen
render at the base point equals its record render bit for bit, and its
difference quotients equal the CPU FP64 oracle's adjoint (itself checked by
FP64 central differences in test_oracle_cpu.py).  At 1M Gaussians the 12-float
gradient (theta, rho, w, v) of the batched fused path (all 32 poses in one
launch) and of the plain drop-in path (autograd through
diff_gaussian_rasterization) are then compared with the FP64 differences of
the function with the FP32 forward's decisions, step h = 1e-6, for every
pose (gaussian_renderer/__init__.py:301-308 routes these gradients): the
velocity entries (w, v) within 1e-3 relative to their largest one on every
pose; the whole gradient within 5e-4 relative (to its largest entry) at the
median over the poses and within 5e-3 on every pose.

Why not 1e-3 on every pose: the loss gradient is the difference of the two
renders' gradients, which cancel (on rotation about the optical axis by up
to ~200x at these poses), and every FP32 evaluation of a render's gradient
carries rounding of ~1e-4 of its size.  Evaluations that are equally valid
in FP32 therefore spread by ~1e-3 of the total on that entry: the FP32
oracle, whose sums are all FP64, is itself 4e-4, 2e-4 and 1.9e-3 from the
differences on poses 0, 9 and 14 (printed by the test), the fused path with
FP32 atomics varies between runs by ~6e-4 there, and FP64 per-Gaussian sums
on the device move it by as much without bringing it closer.  An error in
the adjoint (a dropped term, a wrong sign or index) moves the gradient by
the size of a render's gradient, 10^2-10^4 times the per-pose bound; a 1%
error in a velocity entry fails the velocity bound, one of 1% in a pose
entry the per-pose bound where that entry is the largest.
"""
import math

import numpy as np
import pytest
import torch

from helpers import (orc, synthetic, make_cam, oracle_scene, scene_astype, se3_exp, oracle_scene_from_model,
                     device_cam_records)
from oracle import tracking_step
from oracle.fd64 import FD64

pytestmark = pytest.mark.gpu

GREY64 = np.array([0.2989, 0.5870, 0.1140])


def _engine(s, W, H, f, cuda):
    return FD64(s.means, s.scales, s.rots, s.opac, s.shs, s.sh_degree, W, H, f, f, device=cuda)


def _small(seed=31):
    raw = synthetic.small_test_scene(P=2000, seed=seed, sh_degree=3, spread=0.8, log_scale_mean=-3.5)
    return oracle_scene(raw), (160, 120, 140.0)


def test_fd64_record_render_matches_cpu_fp64_oracle(cuda):
    s, (W, H, f) = _small()
    T = se3_exp([0.02, -0.01, 0.03, 0.02, -0.015, 0.01])
    eng = _engine(s, W, H, f, cuda)
    img, st = eng.record(T)
    s64 = scene_astype(s, np.float64)
    cam = make_cam(T, W, H, f, f, dtype=np.float64)
    pre = orc.preprocess(s64, cam, np.float64)
    b = orc.bin(pre, cam, np.float64)
    I = pre["rgb"].reshape(-1, 3) @ GREY64
    fw = orc.blend_forward(pre, b, cam, colors=I, C=1, dtype=np.float64, n_touched=False)
    got = img.cpu().numpy()
    err = np.abs(got - fw["color"][0]).max()
    print(f"[fd64] record render vs CPU FP64 oracle: max |d| = {err:.2e}, "
          f"{int(st['vals'].numel())} fragments (oracle {b['n']})", flush=True)
    assert int(st["vals"].numel()) == b["n"]
    assert err < 1e-10
    again = eng.frozen(T, st)
    assert torch.equal(again, img), "frozen render at the base point differs from the record render"


def test_fd64_fp32_decisions_reproduce_the_fp32_render(cuda):
    """Decisions read from an FP32 forward of the drop-in path (its binning,
    splat records and per-pixel last fragment): the frozen FP64 render with
    them equals that forward's grey render within FP32 rounding (1e-5; where
    FP32 and FP64 decide differently the FP64 record render differs by up to
    ~1e-3, printed)."""
    from gsevt import raster
    from helpers import gpu_cam, gpu_map
    s, (W, H, f) = _small(33)
    T = se3_exp([0.01, 0.02, -0.01, -0.01, 0.02, 0.01])
    cam = make_cam(T, W, H, f, f)
    out, st = raster.render_forward(gpu_map(s, cuda), gpu_cam(cam, cuda), channels=3, n_touched=False)
    grey = (torch.tensor(GREY64, device=cuda).reshape(3, 1, 1) * out["color"].double()).sum(0)
    eng = _engine(s, W, H, f, cuda)
    dec = eng.record_from_forward(T, st["ranges"], st["vals"], st["n"], st["splats"], 3, st["n_contrib"])
    img = eng.frozen(T, dec)
    rec, _ = eng.record(T)
    d32 = float((img - grey).abs().max())
    d64 = float((rec - grey).abs().max())
    print(f"[fd64] frozen render with FP32 decisions vs the FP32 render: {d32:.2e}; with FP64 decisions {d64:.2e}",
          flush=True)
    assert d32 < 1e-5


@pytest.mark.parametrize("seed", [0, 1])
def test_fd64_differences_match_cpu_fp64_adjoint(cuda, seed):
    s, (W, H, f) = _small(40 + seed)
    rng = np.random.default_rng(seed)
    Tc = se3_exp(rng.normal(0, [0.03] * 3 + [0.02] * 3))
    lin, ang, dtau = rng.normal(0, 0.5, 3), rng.normal(0, 0.5, 3), 0.01
    ev = rng.normal(0, 1.0, (H, W))
    ev /= np.linalg.norm(ev)
    eng = _engine(s, W, H, f, cuda)
    L0, fd = eng.loss_and_fd(Tc, lin, ang, dtau, ev)
    s64 = scene_astype(s, np.float64)
    st, cams = [], []
    for sgn in (-1.0, 1.0):
        E = se3_exp(sgn * 0.5 * dtau * np.concatenate([lin, ang]))
        c = make_cam(E @ Tc, W, H, f, f, dtype=np.float64, vel=E, delta_time=sgn * 0.5 * dtau, lin=lin, ang=ang)
        cams.append(c)
        st.append(orc.render(s64, c, dtype=np.float64, n_touched=False))
    L, g_l, g_n, _ = tracking_step.delta_ir_and_grad(st[0]["fwd"]["color"], st[1]["fwd"]["color"], ev)
    bl = orc.render_backward(s64, cams[0], st[0], g_l, dtype=np.float64)["pose"]
    bn = orc.render_backward(s64, cams[1], st[1], g_n, dtype=np.float64)["pose"]
    an = np.concatenate([bl[k] + bn[k] for k in ("theta", "rho", "w", "v")])
    rel = np.abs(fd - an).max() / np.abs(an).max()
    print(f"[fd64] loss {L0:.10f} (CPU FP64 {L:.10f}); FD vs CPU FP64 adjoint rel err {rel:.2e}", flush=True)
    assert abs(L0 - L) < 1e-9
    assert rel < 1e-6


def _c3_poses(T_gt, n=32):
    """Seeds 6000..6031: Exp(xi) T_gt, xi ~ N(0, diag((5 cm)^2 I, (3 deg)^2 I))."""
    out = []
    for k in range(n):
        rng = np.random.Generator(np.random.PCG64(6000 + k))
        xi = np.concatenate([rng.normal(0, 0.05, 3), rng.normal(0, math.radians(3.0), 3)])
        out.append(se3_exp(xi) @ T_gt)
    return np.stack(out)


@pytest.mark.timeout(1800)
def test_adjoint_vs_fp64_differences_1m_32_poses(cuda, monkeypatch):
    """Per pose, the FP64 differences of the loss with the decisions of the
    FP32 forward frozen (which fragment is blended into which pixel, as that
    forward decided; the plain and the fused forwards each use their own
    float32 camera matrices and binning, so each gradient is checked against
    its own forward's decisions): the derivative of the function the FP32
    adjoint differentiates.  Both HIP gradients must meet the bounds of the
    module docstring against it, the FP64 loss with those decisions within
    1e-5 of the FP32 loss.  On poses 0, 9 and 14 the FP32 oracle
    (FP32 terms, FP64 sums) on the fused path's camera records is compared
    with the same differences, as the measure of the FP32 floor.
    The differences with decisions taken in FP64 are printed beside them: a
    fragment whose alpha or transmittance test comes out differently in FP32
    and FP64 changes a pixel by up to ~4e-3, which moves the FP64 function's
    gradient by ~1e-3 relative without any error in the adjoint."""
    from gsevt import raster, workloads
    from gsevt.fused import BatchedTracker
    from gsevt.frame import RenderFrame
    from gsevt.tracker import tracking_loss
    prob = workloads.build("patches_1m", device=cuda)
    intr = prob.intr
    gm = prob.gaussians
    assert gm.num_points == 1_000_000 and gm.active_sh_degree == 3 and (intr.W, intr.H) == (640, 480)
    T_gt = prob.gt_pose.double().cpu().numpy()
    poses = _c3_poses(T_gt)
    lin = prob.gt_lin.double().cpu().numpy()
    ang = prob.gt_ang.double().cpu().numpy()
    dtau = float(np.float32(prob.delta_tau))
    B = len(poses)
    # fused: all 32 poses as the items of one launch
    tr = BatchedTracker(gm, intr.W, intr.H, intr.fx, intr.fy, prob.sign_Ie, B, device=cuda)
    tr.set_items(poses, np.tile(lin, (B, 1)), np.tile(ang, (B, 1)), [dtau] * B, [0] * B)
    tr.measure_capacity()
    tr.step(do_step=False)
    torch.cuda.synchronize()
    tr.check_overflow()
    g_fused = tr.grad12.cpu().numpy().astype(np.float64)
    loss_fused = tr.loss.cpu().numpy()
    R, HW, gid_mask = tr.R, intr.W * intr.H, tr.gid_mask
    fg, rg = tr.fgid.view(R, tr.cap), tr.ranges.view(R, -1)
    sp, nc, nk = tr.splats.view(R, -1), tr.n_contrib.view(R, HW), tr.nkept.cpu().numpy()
    s = oracle_scene_from_model(gm)
    eng = FD64(s.means, s.scales, s.rots, s.opac, s.shs, s.sh_degree, intr.W, intr.H, intr.fx, intr.fy, device=cuda)
    ev = prob.sign_Ie[0].double()
    states = []
    orig_rf = raster.render_forward

    def rec_forward(*args, **kw):
        out, state = orig_rf(*args, **kw)
        states.append(state)
        return out, state

    monkeypatch.setattr(raster, "render_forward", rec_forward)
    worst = dict(fused=0.0, plain=0.0, fused64=0.0)
    errs_f, errs_p = [], []
    ocams = device_cam_records(tr)
    vel, dts = tr.vel.cpu().numpy(), tr.dts.cpu().numpy()
    for k in range(B):
        # plain drop-in path at the same pose; its two forwards' states give the decisions
        states.clear()
        Tk = torch.as_tensor(poses[k], dtype=torch.float32, device=cuda)
        cam = prob.camera(pose=Tk, lin=prob.gt_lin, ang=prob.gt_ang)
        for p in (cam.cam_rot_delta, cam.cam_trans_delta, cam.cam_w_delta, cam.cam_v_delta):
            p.requires_grad_(True)
        rf = RenderFrame(cam, gm, None, prob.background, 0)
        loss = tracking_loss(rf.sign_delta_Ir, prob.sign_Ie)
        loss.backward()
        g_plain = torch.cat([cam.cam_rot_delta.grad, cam.cam_trans_delta.grad, cam.cam_w_delta.grad,
                             cam.cam_v_delta.grad]).double().cpu().numpy()
        assert len(states) == 2
        Tl, Tn = eng.poses(poses[k], lin, ang, dtau)
        # the plain forward's decisions (its own binning and float32 matrices)
        dec_p = tuple(eng.record_from_forward(T, st["ranges"], st["vals"], st["n"], st["splats"], 3,
                                              st["n_contrib"]) for T, st in zip((Tl, Tn), states))
        # the fused forward's (tile-culled packed lists, its own camera records)
        dec_f = tuple(eng.record_from_forward(T, rg[r], fg[r], int(nk[r]), sp[r], 2, nc[r], gid_mask=gid_mask)
                      for T, r in zip((Tl, Tn), (2 * k, 2 * k + 1)))
        Lp, fd_p = eng.loss_and_fd(poses[k], lin, ang, dtau, ev, decisions=dec_p)
        Lf, fd_f = eng.loss_and_fd(poses[k], lin, ang, dtau, ev, decisions=dec_f)
        _, fd64 = eng.loss_and_fd(poses[k], lin, ang, dtau, ev)
        ef = np.abs(g_fused[k] - fd_f).max() / np.abs(fd_f).max()
        ep = np.abs(g_plain - fd_p).max() / np.abs(fd_p).max()
        e64 = np.abs(g_fused[k] - fd64).max() / np.abs(fd64).max()
        worst = dict(fused=max(worst["fused"], ef), plain=max(worst["plain"], ep), fused64=max(worst["fused64"], e64))
        print(f"[c3] pose {k:2d}: loss fused {loss_fused[k]:.6f} (FP64 {Lf:.6f}), plain {float(loss):.6f} "
              f"(FP64 {Lp:.6f}); max|FD| {np.abs(fd_f).max():.3e}; rel err fused {ef:.2e}, plain {ep:.2e}; "
              f"(fused vs FD with FP64 decisions {e64:.2e})", flush=True)
        if k in (0, 9, 14):
            # the FP32 floor: the oracle's FP32 gradient (FP64 sums) on the same camera records
            sts = []
            for r in (2 * k, 2 * k + 1):
                c = ocams[r]
                c.vel, c.delta_time, c.lin, c.ang = vel[r].astype(np.float64), float(dts[r]), lin, ang
                sts.append((c, orc.render(s, c, n_touched=False)))
            (cl, sl), (cn, sn) = sts
            _, g_l, g_n, _ = tracking_step.delta_ir_and_grad(sl["fwd"]["color"], sn["fwd"]["color"],
                                                             ev.cpu().numpy())
            bl = orc.render_backward(s, cl, sl, g_l.astype(np.float32))["pose"]
            bn = orc.render_backward(s, cn, sn, g_n.astype(np.float32))["pose"]
            go = np.concatenate([bl[q] + bn[q] for q in ("theta", "rho", "w", "v")])
            gl = np.concatenate([bl[q] for q in ("theta", "rho", "w", "v")])
            eo = np.abs(go - fd_f).max() / np.abs(fd_f).max()
            print(f"[c3] pose {k:2d}: FP32 oracle (FP64 sums) rel err {eo:.2e}; largest entry of one render's "
                  f"gradient {np.abs(gl).max():.3e} against {np.abs(fd_f).max():.3e} for the loss", flush=True)
            assert eo < 5e-3
        assert abs(Lf - loss_fused[k]) < 1e-5 and abs(Lp - float(loss)) < 1e-5
        vf = np.abs(g_fused[k][6:] - fd_f[6:]).max() / np.abs(fd_f[6:]).max()
        vp = np.abs(g_plain[6:] - fd_p[6:]).max() / np.abs(fd_p[6:]).max()
        print(f"[c3] pose {k:2d}: velocity entries rel err fused {vf:.2e}, plain {vp:.2e}", flush=True)
        assert vf < 1e-3 and vp < 1e-3, (k, fd_f, g_fused[k], fd_p, g_plain)
        assert ef < 5e-3 and ep < 5e-3, (k, fd_f, g_fused[k], fd_p, g_plain)
        errs_f.append(ef)
        errs_p.append(ep)
    n1f = sum(e < 1e-3 for e in errs_f)
    n1p = sum(e < 1e-3 for e in errs_p)
    print(f"[c3] 32 poses: worst rel err fused {worst['fused']:.2e} (median {np.median(errs_f):.2e}, {n1f} of 32 "
          f"within 1e-3), plain {worst['plain']:.2e} (median {np.median(errs_p):.2e}, {n1p} within 1e-3); "
          f"fused vs FD with FP64 decisions {worst['fused64']:.2e}", flush=True)
    assert np.median(errs_f) < 5e-4 and np.median(errs_p) < 5e-4
