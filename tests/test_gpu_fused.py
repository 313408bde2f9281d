"""GPU parity of the batched, fused tracking path.

* binning (depth sort + tile sort) and the grey-intensity render of every
  item's two renders are bit-exact against the FP32 oracle run on the camera
  records the device computed; with the tile cull, the fragment lists are the
  oracle's minus fragments that reach no pixel of their tile;
* the per-item loss and 12-float gradient (theta, rho, w, v) match the oracle
  chain (loss 1e-5 absolute, gradient 1e-3 relative);
* at the same state the fused gradient equals the one autograd returns
  through the drop-in rasterizer, and on a ground-truth-motion problem the
  fused poses stay within 1e-4 m / 1e-3 rad of the plain drop-in path after
  the same Adam iterations;
* a captured graph replays to the same results as eager launches.
"""
import math

import numpy as np
import pytest
import torch

from helpers import orc, synthetic, se3_exp
from oracle import tracking_step

pytestmark = pytest.mark.gpu

GREY = np.array([0.2989, 0.5870, 0.1140], np.float32)


def _problem(cuda, B=2, P=4000, W=160, H=120, f=140.0, seed=0, log_scale_mean=-4.0):
    from gsevt.model import GaussianModel
    raw = synthetic.patch_scene(P=P, sh_degree=3, seed=1000 + seed, n_patches=16, log_scale_mean=log_scale_mean)
    gm = GaussianModel.from_raw(raw.xyz, raw.features_dc, raw.features_rest, raw.scaling, raw.rotation,
                                raw.opacity, raw.sh_degree, device=cuda)
    rng = np.random.default_rng(seed)
    poses = np.stack([se3_exp(rng.normal(0, [0.02] * 3 + [0.01] * 3)) for _ in range(B)])
    lin = rng.normal(0, 0.3, (B, 3))
    ang = rng.normal(0, 0.3, (B, 3))
    ev = torch.from_numpy(rng.normal(0, 1.0 / math.sqrt(W * H), (2, H, W)).astype(np.float32)).to(cuda)
    return gm, raw, poses, lin, ang, ev, (W, H, f)


def _tracker(cuda, gm, ev, B, W, H, f, **kw):
    from gsevt.fused import BatchedTracker
    return BatchedTracker(gm, W, H, f, f, ev, B, device=cuda, **kw)


def _cam_records(tr):
    raw = tr.cams.cpu().numpy().view(np.uint8).reshape(tr.R, -1)
    out = []
    for r in range(tr.R):
        fl = raw[r, :53 * 4].view(np.float32)
        out.append(orc.Cam(view=fl[0:16].copy(), proj=fl[16:32].copy(), projraw=fl[32:48].copy(),
                           campos=fl[48:51].copy(), W=tr.W, H=tr.H, tanfovx=float(fl[51]), tanfovy=float(fl[52]),
                           bg=np.zeros(3, np.float32), tile_w=tr.tile[0], tile_h=tr.tile[1]))
    return out


def _oracle_scene(gm):
    """The exact activated tensors the device used (torch's on-device
    exp / normalize / sigmoid), so the oracle sees the same inputs."""
    c = lambda t: np.ascontiguousarray(t.detach().float().cpu().numpy())
    return orc.Scene(means=c(gm.get_xyz), scales=c(gm.get_scaling), rots=c(gm.get_rotation),
                     opac=c(gm.get_opacity).reshape(-1), shs=c(gm.get_features), sh_degree=gm.active_sh_degree)


@pytest.mark.parametrize("mode", ["all_slots", "tile_culled", "unpacked"])
@pytest.mark.parametrize("tile", [(16, 16), (32, 8)])
def test_fused_binning_and_render_bitexact(cuda, tile, mode):
    """Without the tile cull the fragment lists equal the oracle's binning
    (packed tile|gaussian words, or separate tile and Gaussian arrays, the
    form taken when the two do not fit 32 bits); with it they are the
    oracle's lists minus fragments that reach no pixel of their tile
    (check_culled_binning).  The render is bit-exact every way."""
    from helpers import check_culled_binning
    gm, raw, poses, lin, ang, ev, (W, H, f) = _problem(cuda)
    cull = mode == "tile_culled"
    tr = _tracker(cuda, gm, ev, 2, W, H, f, tile=tile, cull_tiles=cull)
    if mode == "unpacked":
        tr.packed = False
        tr.gid_mask = 0xFFFFFFFF
    tr.set_items(poses, lin, ang, [0.01, 0.008], [0, 1])
    tr.measure_capacity()
    tr.step(do_step=False)
    torch.cuda.synchronize()
    tr.check_overflow()
    scene = _oracle_scene(gm)
    cams = _cam_records(tr)
    HW = W * H
    inten = tr.inten.cpu().numpy().reshape(tr.R, H, W)
    fgid = tr.fgid.cpu().numpy().view(np.uint32).reshape(tr.R, tr.cap) & np.uint32(tr.gid_mask)
    ranges = tr.ranges.cpu().numpy().view(np.uint32).reshape(tr.R, -1)
    nfrag = tr.nfrag.cpu().numpy()
    nkept = tr.nkept.cpu().numpy()
    for r, cam in enumerate(cams):
        print(f"[fused] render {r} tile {tile} {mode} nfrag {nfrag[r]} kept {nkept[r]}", flush=True)
        pre = orc.preprocess(scene, cam)
        b = orc.bin(pre, cam)
        assert nfrag[r] == b["n"]
        if cull:
            kept, _ = check_culled_binning(fgid[r, :nkept[r]], ranges[r], b, pre, cam)
            assert kept == nkept[r] < b["n"]
        else:
            np.testing.assert_array_equal(fgid[r, :b["n"]], b["vals"])
            np.testing.assert_array_equal(ranges[r], b["ranges"])
        rgb = pre["rgb"].reshape(-1, 3)
        I = (GREY[0] * rgb[:, 0] + GREY[1] * rgb[:, 1]) + GREY[2] * rgb[:, 2]
        fw = orc.blend_forward(pre, b, cam, colors=I.astype(np.float32), C=1, n_touched=False)
        assert np.abs(inten[r] - fw["color"][0]).max() <= 1e-5
        assert np.array_equal(inten[r].view(np.uint32), fw["color"][0].view(np.uint32))


@pytest.mark.parametrize("cull", [True, False])
@pytest.mark.parametrize("shape", [(160, 120, 140.0, (16, 16), -4.0), (320, 240, 280.0, (16, 16), -4.0),
                                   (640, 480, 550.0, (32, 8), -4.0), (640, 480, 550.0, (16, 16), -1.0)])
def test_emission_first_pass_histogram(cuda, cull, shape):
    """The packed emission's count of the tile sort's first digit (3, 4 and 5
    bits for 80, 300 and 1200 tiles) equals a count of the keys it emitted:
    kept fragments by the low bits of their tile, per 4096-slot block.  The
    large splats of the last case (log-scale mean -1) give each workgroup's
    256 Gaussians several 4096-slot blocks, so its counts leave the four-block
    window and are flushed before the workgroup ends.  The
    sort on that histogram gives the same fragment lists, ranges, kept counts
    and renders as the sort that counts the keys itself."""
    W, H, f, tile, log_scale = shape
    gm, raw, poses, lin, ang, ev, _ = _problem(cuda, B=2, W=W, H=H, f=f, seed=7, log_scale_mean=log_scale)
    tr = _tracker(cuda, gm, ev, 2, W, H, f, tile=tile, cull_tiles=cull)
    tr.set_items(poses, lin, ang, [0.01, 0.008], [0, 1])
    tr.measure_capacity()
    assert tr.packed and tr._emit_hist(tr.cap)
    # packed: the tile counts travel in the depth-sort values, no count buffers
    assert tr.tiles is None and tr.tsorted is None
    assert tr.device_bytes() >= 4 * (tr.splats.numel() + tr.fgid.numel() + tr.fgid2.numel())
    nb_items, state = tr.states[0]
    tr._forward_geometry(tr.cap, nb_items, state)
    torch.cuda.synchronize()
    R, cap, bits = tr.R, tr.cap, tr.first_tile_bits
    nblk = -(-cap // 4096)
    hist = tr.fsort_scr[:R * (1 << bits) * nblk].cpu().numpy().view(np.uint32).reshape(R, 1 << bits, nblk)
    keys = tr.fgid.cpu().numpy().view(np.uint32).reshape(R, cap)
    nfrag = tr.nfrag.cpu().numpy()
    for r in range(R):
        k = keys[r, :min(int(nfrag[r]), cap)]
        kept = k != np.uint32(0xFFFFFFFF)
        assert cull or kept.all()
        digit = (k >> np.uint32(tr.gid_bits)) & np.uint32((1 << bits) - 1)
        blk = np.arange(k.size) // 4096
        want = np.zeros((1 << bits, nblk), np.int64)
        np.add.at(want, (digit[kept].astype(np.int64), blk[kept]), 1)
        np.testing.assert_array_equal(hist[r].astype(np.int64), want)
    # sort blocks met by each emission workgroup's 256 depth-consecutive Gaussians
    counts = (tr.dval[:R * tr.P].cpu().numpy().view(np.uint32) >> np.uint32(tr.gid_bits)).reshape(R, tr.P)
    incl = np.cumsum(counts.astype(np.int64), axis=1)
    starts = np.concatenate([np.zeros((R, 1), np.int64), incl[:, :-1]], axis=1)[:, ::256]
    ends = incl[:, np.minimum(np.arange(255, tr.P + 255, 256), tr.P - 1)]
    met = np.where(ends > starts, (ends - 1) // 4096 - starts // 4096 + 1, 0).max()
    if log_scale > -2.0:
        assert met > 4, met   # some workgroup flushes its window before it ends
    print(f"[fused] {W}x{H} tile {tile} cull {cull}: {bits}-bit first digit over {nblk} blocks, "
          f"{int(hist.sum())} kept fragments counted, up to {met} blocks per workgroup", flush=True)
    outs = []
    for fuse in (True, False):
        tr.fuse_first_hist = fuse
        tr.step(do_step=False)
        torch.cuda.synchronize()
        kept_n = (tr.nkept if cull else tr.nfrag).cpu().numpy().copy()
        fg = tr.fgid.cpu().numpy().view(np.uint32).reshape(R, cap)
        outs.append((kept_n, [fg[r, :kept_n[r]].copy() for r in range(R)], tr.ranges.cpu().numpy().copy(),
                     tr.inten.cpu().numpy().copy()))
    (k1, f1, r1, i1), (k2, f2, r2, i2) = outs
    np.testing.assert_array_equal(k1, k2)
    for a, b in zip(f1, f2):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(r1, r2)
    assert np.array_equal(i1.view(np.uint32), i2.view(np.uint32))


def test_fused_loss_and_gradient_vs_oracle(cuda):
    gm, raw, poses, lin, ang, ev, (W, H, f) = _problem(cuda, seed=1)
    tr = _tracker(cuda, gm, ev, 2, W, H, f)
    tr.set_items(poses, lin, ang, [0.01, 0.012], [1, 0])
    tr.measure_capacity()
    tr.step(do_step=False)
    torch.cuda.synchronize()
    scene = _oracle_scene(gm)
    cams = _cam_records(tr)
    vel = tr.vel.cpu().numpy()
    dts = tr.dts.cpu().numpy()
    loss = tr.loss.cpu().numpy()
    g12 = tr.grad12.cpu().numpy()
    evn = ev.cpu().numpy()
    for b in range(2):
        cl, cn = cams[2 * b], cams[2 * b + 1]
        for side, c in ((0, cl), (1, cn)):
            c.vel = vel[2 * b + side].astype(np.float64)
            c.delta_time = float(dts[2 * b + side])
            c.lin = lin[b]
            c.ang = ang[b]
        st_l = orc.render(scene, cl, n_touched=False)
        st_n = orc.render(scene, cn, n_touched=False)
        e = evn[[1, 0][b]]
        L, g_l, g_n, _ = tracking_step.delta_ir_and_grad(st_l["fwd"]["color"], st_n["fwd"]["color"], e)
        assert abs(L - loss[b]) < 1e-5, (L, loss[b])
        bl = orc.render_backward(scene, cl, st_l, g_l.astype(np.float32))
        bn = orc.render_backward(scene, cn, st_n, g_n.astype(np.float32))
        ref = np.concatenate([bl["pose"][k] + bn["pose"][k] for k in ("theta", "rho", "w", "v")])
        rel = np.abs(g12[b] - ref).max() / np.abs(ref).max()
        print(f"[fused] item {b}: loss {loss[b]:.6f} (oracle {L:.6f}), grad rel err {rel:.2e}", flush=True)
        assert rel < 1e-3


def _plain_camera(cuda, pose, lin, ang, W, H, f, dtau):
    from gsevt.camera import Camera
    fov = 2 * math.atan(W / (2 * f)), 2 * math.atan(H / (2 * f))
    T0 = torch.as_tensor(np.asarray(pose), dtype=torch.float32, device=cuda)
    cam = Camera(T0[:3, :3].clone(), T0[:3, 3].clone(), torch.as_tensor(np.asarray(ang), dtype=torch.float32,
                 device=cuda).clone(), torch.as_tensor(np.asarray(lin), dtype=torch.float32, device=cuda).clone(),
                 fov[0], fov[1], W, H, delta_tau=dtau, device=cuda)
    cam.fx = cam.fy = f
    return cam


def test_fused_gradient_matches_plain_dropin_path(cuda):
    """Same state, one iteration: the fused 12-vector gradient equals the
    gradient autograd returns through the drop-in rasterizer (RGB renders,
    Camera properties, render2, RenderFrame)."""
    from gsevt.frame import RenderFrame
    from gsevt.tracker import tracking_loss
    gm, raw, poses, lin, ang, ev, (W, H, f) = _problem(cuda, B=1, seed=4)
    tr = _tracker(cuda, gm, ev[:1], 1, W, H, f)
    tr.set_items(poses, lin, ang, [0.01], [0])
    tr.measure_capacity()
    tr.step(do_step=False)
    torch.cuda.synchronize()
    g_fused = tr.grad12[0].cpu().numpy()
    cam = _plain_camera(cuda, poses[0], lin[0], ang[0], W, H, f, 0.01)
    for p in (cam.cam_rot_delta, cam.cam_trans_delta, cam.cam_w_delta, cam.cam_v_delta):
        p.requires_grad_(True)
    rf = RenderFrame(cam, gm, None, torch.zeros(3, device=cuda), 0)
    loss = tracking_loss(rf.sign_delta_Ir, ev[0:1])
    loss.backward()
    g_plain = torch.cat([cam.cam_rot_delta.grad, cam.cam_trans_delta.grad, cam.cam_w_delta.grad,
                         cam.cam_v_delta.grad]).cpu().numpy()
    rel = np.abs(g_fused - g_plain).max() / np.abs(g_plain).max()
    print(f"[fused vs plain] loss {float(tr.loss[0]):.6f} / {float(loss):.6f}; grad rel err {rel:.2e}")
    print("  fused", np.array2string(g_fused, precision=4))
    print("  plain", np.array2string(g_plain, precision=4))
    assert abs(float(tr.loss[0]) - float(loss)) < 1e-5
    assert rel < 1e-3


def test_fused_tracks_like_plain_path(cuda):
    """Events rendered from a ground-truth motion, initial pose 2 cm / 1 deg
    off: after the same number of Adam iterations both paths agree within
    1e-4 m and 1e-3 rad."""
    from gsevt import workloads
    from gsevt.fused import BatchedTracker
    from gsevt.tracker import plain_fine_iterations
    prob = workloads.build("cpu_small", device=cuda)
    n_it = 30
    intr = prob.intr
    T0 = prob.init_pose.double().cpu().numpy()
    tr = BatchedTracker(prob.gaussians, intr.W, intr.H, intr.fx, intr.fy, prob.sign_Ie, 1, device=cuda)
    tr.set_items(T0[None], prob.init_lin[None], prob.init_ang[None], [prob.delta_tau], [0])
    tr.measure_capacity()
    tr.run(n_it, use_graph=False)
    torch.cuda.synchronize()
    Tf = tr.poses()[0].cpu().numpy()
    cam = prob.camera()
    plain_fine_iterations(cam, prob.gaussians, prob.background, prob.sign_Ie, n_it, (0.004, 0.004, 0.002, 0.002))
    Tp = np.eye(4)
    Tp[:3, :3] = cam.R.detach().cpu().numpy()
    Tp[:3, 3] = cam.T.detach().cpu().numpy()
    dt = np.abs(Tf[:3, 3] - Tp[:3, 3]).max()
    dR = Tf[:3, :3] @ Tp[:3, :3].T
    ang_err = math.acos(max(-1.0, min(1.0, (np.trace(dR) - 1) / 2)))
    moved = np.abs(Tf[:3, 3] - T0[:3, 3]).max()
    print(f"[fused vs plain] {n_it} it on cpu_small: |dt| {dt:.2e} m, angle {ang_err:.2e} rad "
          f"(moved {moved:.2e} m); losses {float(tr.loss[0]):.5f}")
    assert moved > 1e-4
    assert dt < 1e-4 and ang_err < 1e-3


def test_fused_graph_replay_matches_eager(cuda):
    gm, raw, poses, lin, ang, ev, (W, H, f) = _problem(cuda, B=3, seed=3)
    a = _tracker(cuda, gm, torch.cat([ev, ev[:1]]), 3, W, H, f)
    a.set_items(poses, lin, ang, [0.01, 0.009, 0.011], [0, 1, 2])
    a.measure_capacity()
    a.run(4, use_graph=False)
    torch.cuda.synchronize()
    ref_hist = a.loss_hist[:, :4].cpu().numpy()
    ref_pose = a.poses().cpu().numpy()
    a.set_items(poses, lin, ang, [0.01, 0.009, 0.011], [0, 1, 2])
    a.capture()
    a.set_items(poses, lin, ang, [0.01, 0.009, 0.011], [0, 1, 2])
    for k in range(4):
        a.graph.replay()
        torch.cuda.synchronize()
        x = torch.ones(1000, device=cuda).sum()   # unrelated eager work between replays
        assert float(x) == 1000.0
    hist = a.loss_hist[:, :4].cpu().numpy()
    np.testing.assert_allclose(hist, ref_hist, rtol=1e-4, atol=1e-6)
    np.testing.assert_allclose(a.poses().cpu().numpy(), ref_pose, atol=1e-6)
    a.check_overflow()


def test_autotune_restores_state_and_keeps_tracking(cuda):
    from gsevt.fused import TILE_SHAPES
    gm, raw, poses, lin, ang, ev, (W, H, f) = _problem(cuda, B=3, seed=4)
    evs = torch.cat([ev, ev[:1]])
    a = _tracker(cuda, gm, evs, 3, W, H, f)
    a.set_items(poses, lin, ang, [0.01, 0.009, 0.011], [0, 1, 2])
    a.measure_capacity()
    before = {k: getattr(a, k).clone() for k in a._STATE}
    best, times = a.autotune(iters=2, warmup=1)
    print(f"[autotune] {times} -> {best}", flush=True)
    assert best in TILE_SHAPES and set(times) == set(TILE_SHAPES) and a.tile == best
    for k, v in before.items():
        assert torch.equal(getattr(a, k), v), k
    a.run(3, use_graph=False)
    b = _tracker(cuda, gm, evs, 3, W, H, f, tile=best)
    b.set_items(poses, lin, ang, [0.01, 0.009, 0.011], [0, 1, 2])
    b.measure_capacity()
    b.run(3, use_graph=False)
    torch.cuda.synchronize()
    np.testing.assert_allclose(a.loss_hist[:, :3].cpu().numpy(), b.loss_hist[:, :3].cpu().numpy(), rtol=1e-5,
                               atol=1e-7)
    np.testing.assert_allclose(a.poses().cpu().numpy(), b.poses().cpu().numpy(), atol=1e-6)


def test_fused_loss_when_render_does_not_change(cuda):
    """Zero velocity: the last and next cameras coincide, dI = 0 exactly.
    The fused loss is then ||e|| (dI / ||dI|| taken as 0; the reference's
    0 / 0 gives NaN) and the gradient is zero."""
    gm, raw, poses, lin, ang, ev, (W, H, f) = _problem(cuda, B=1, seed=5)
    tr = _tracker(cuda, gm, ev[:1], 1, W, H, f)
    tr.set_items(poses, np.zeros((1, 3)), np.zeros((1, 3)), [0.01], [0])
    tr.measure_capacity()
    tr.step(do_step=False)
    torch.cuda.synchronize()
    inten = tr.inten.view(2, -1)
    assert torch.equal(inten[0], inten[1])
    e = float(ev[0].double().norm())
    print(f"[fused] unchanged render: loss {float(tr.loss[0]):.7f}, ||e|| {e:.7f}", flush=True)
    assert abs(float(tr.loss[0]) - e) <= 1e-6 * max(1.0, e)
    assert torch.all(tr.grad12 == 0)


def test_fused_coarse_then_fine_stage_tracks_like_plain_path(cuda):
    """The coarse stage of utils/tracker.py:149-229 (unsigned images, pose
    only: velocity parameters without gradients, so their Adam groups do not
    step; update_pose) for 8 iterations, then the fine stage (signed images,
    pose and velocity, learning-rate ramp from start_vel_opt_iter = 7 as when
    the coarse stage converges at iteration 7) for 8 more, on one Adam state.
    The fused path (unsigned=True, opt_vel=False; then start_vel_iter=7) and
    the plain drop-in path agree per iteration (loss) and at the end (pose
    within 1e-4 m / 1e-3 rad, velocity within 1e-4)."""
    from gsevt import workloads
    from gsevt.fused import BatchedTracker
    from gsevt.frame import RenderFrame
    from gsevt.tracker import lr_fraction, tracking_loss
    prob = workloads.build("cpu_small", device=cuda)
    n_coarse, n_fine, max_iter = 8, 8, 200
    lr = (0.004, 0.004, 0.002, 0.002)
    intr = prob.intr
    T0 = prob.init_pose.double().cpu().numpy()
    evs = torch.stack([prob.unsign_Ie[0], prob.sign_Ie[0]])
    tr = BatchedTracker(prob.gaussians, intr.W, intr.H, intr.fx, intr.fy, evs, 1, device=cuda,
                        max_optim_iter=max_iter)
    tr.set_items(T0[None], prob.init_lin[None], prob.init_ang[None], [prob.delta_tau], [0])
    tr.measure_capacity()
    for _ in range(n_coarse):
        tr.step(opt_vel=False, unsigned=True)
    tr.window.fill_(1)    # the signed frame
    for _ in range(n_fine):
        tr.step(opt_vel=True, unsigned=False, start_vel_iter=n_coarse - 1)
    torch.cuda.synchronize()
    hist = tr.loss_hist[0, :n_coarse + n_fine].cpu().numpy()
    Tf = tr.poses()[0].cpu().numpy()
    # plain path with one torch Adam across both stages
    cam = prob.camera()
    opt = torch.optim.Adam([{"params": [cam.cam_rot_delta], "lr": lr[0]}, {"params": [cam.cam_trans_delta], "lr": lr[1]},
                            {"params": [cam.cam_w_delta], "lr": lr[2]}, {"params": [cam.cam_v_delta], "lr": lr[3]}])
    losses = []
    for it in range(n_coarse + n_fine):
        coarse = it < n_coarse
        cam.cam_rot_delta.requires_grad_(True)
        cam.cam_trans_delta.requires_grad_(True)
        cam.cam_w_delta.requires_grad_(not coarse)
        cam.cam_v_delta.requires_grad_(not coarse)
        if not coarse:
            fr = lr_fraction(it, n_coarse - 1, max_iter)
            for g, b, s in zip(opt.param_groups, lr, (fr, fr, 1 - fr, 1 - fr)):
                g["lr"] = b * s
        rf = RenderFrame(cam, prob.gaussians, None, prob.background, 0)
        loss = tracking_loss(rf.unsign_delta_Ir if coarse else rf.sign_delta_Ir,
                             prob.unsign_Ie if coarse else prob.sign_Ie)
        loss.backward()
        losses.append(float(loss))
        with torch.no_grad():
            opt.step()
            if coarse:
                cam.update_pose()
            else:
                cam.update_vwRT()
            opt.zero_grad()
    Tp = np.eye(4)
    Tp[:3, :3] = cam.R.detach().cpu().numpy()
    Tp[:3, 3] = cam.T.detach().cpu().numpy()
    dt = np.abs(Tf[:3, 3] - Tp[:3, 3]).max()
    dR = Tf[:3, :3] @ Tp[:3, :3].T
    ang_err = math.acos(max(-1.0, min(1.0, (np.trace(dR) - 1) / 2)))
    dv = max(float((tr.lin[0] - cam.linear_vel).abs().max()), float((tr.ang[0] - cam.angular_vel).abs().max()))
    moved_v = float((cam.linear_vel - prob.init_lin).abs().max())
    print(f"[fused coarse->fine] losses fused {np.array2string(hist, precision=5)}", flush=True)
    print(f"[fused coarse->fine] losses plain {np.array2string(np.array(losses), precision=5)}", flush=True)
    print(f"[fused coarse->fine] |dt| {dt:.2e} m, angle {ang_err:.2e} rad, |dvel| {dv:.2e} (moved {moved_v:.2e})",
          flush=True)
    np.testing.assert_allclose(hist, losses, rtol=1e-4, atol=1e-5)
    assert moved_v > 1e-6
    assert dt < 1e-4 and ang_err < 1e-3 and dv < 1e-4


@pytest.mark.parametrize("schedule", ["queue", "queue_xcd"])
def test_fused_persistent_blend_schedules_match_grid(cuda, schedule):
    """The persistent work-queue blends (a fixed grid of workgroups pulling
    (render, tile) items from one queue, or from 8 per-XCD queues) render the
    same images bit for bit as one workgroup per item, and give the same loss
    and gradients within float-atomic rounding (1e-4); fewer workgroups than items
    (64) and more (4096) both drain the queues."""
    gm, raw, poses, lin, ang, ev, (W, H, f) = _problem(cuda, B=3, seed=6)
    evs = torch.cat([ev, ev[:1]])
    ref = _tracker(cuda, gm, evs, 3, W, H, f)
    ref.set_items(poses, lin, ang, [0.01, 0.009, 0.011], [0, 1, 2])
    ref.measure_capacity()
    ref.step(do_step=False)
    torch.cuda.synchronize()
    for wg in (64, 4096):
        q = _tracker(cuda, gm, evs, 3, W, H, f, schedule=schedule, queue_workgroups=wg)
        q.set_items(poses, lin, ang, [0.01, 0.009, 0.011], [0, 1, 2])
        q.measure_capacity()
        q.step(do_step=False)
        torch.cuda.synchronize()
        assert torch.equal(q.inten, ref.inten) and torch.equal(q.final_T, ref.final_T)
        assert torch.equal(q.n_contrib, ref.n_contrib)
        np.testing.assert_allclose(q.loss.cpu().numpy(), ref.loss.cpu().numpy(), rtol=1e-6)
        g, gr = q.grad12.cpu().numpy(), ref.grad12.cpu().numpy()
        rel = np.abs(g - gr).max() / np.abs(gr).max()
        print(f"[fused] {schedule} with {wg} workgroups: renders bit-exact, gradient rel diff {rel:.1e}", flush=True)
        # another order of the backward's float atomics; the pose gradient is
        # a difference of the two renders' (nearly equal) gradients, which
        # magnifies rounding to ~1e-5 (two grid launches differ as much)
        assert rel < 1e-4


@pytest.mark.parametrize("contributors_first", [False, True])
def test_reordered_map_matches_permuted_oracle(cuda, contributors_first):
    """With the frozen map laid out front to back for one render
    (reorder_map_by_depth; with contributors_first, the Gaussians with a
    gradient in that render first), the tracker renders what the oracle
    renders for the model with its Gaussians listed in that order: culled
    binning checked, grey renders bit-exact; loss and gradient equal those of
    the tracker in model order (within float-atomic rounding)."""
    from helpers import check_culled_binning
    gm, raw, poses, lin, ang, ev, (W, H, f) = _problem(cuda, B=2, seed=7)
    ref = _tracker(cuda, gm, ev, 2, W, H, f)
    ref.set_items(poses, lin, ang, [0.01, 0.012], [0, 1])
    ref.measure_capacity()
    ref.step(do_step=False)
    tr = _tracker(cuda, gm, ev, 2, W, H, f)
    tr.set_items(poses, lin, ang, [0.01, 0.012], [0, 1])
    tr.measure_capacity()
    tr.reorder_map_by_depth(0, contributors_first=contributors_first)
    tr.measure_capacity()
    tr.step(do_step=False)
    torch.cuda.synchronize()
    perm = tr.map_perm.cpu().numpy()
    assert np.array_equal(np.sort(perm), np.arange(tr.P)) and not np.array_equal(perm, np.arange(tr.P))
    if contributors_first:
        # the Gaussians with a gradient in render 0 lead, the rest follow
        act = tr.contributors(0).cpu().numpy()
        n_act = int(act.sum())
        assert 0 < n_act < tr.P and act[:n_act].all() and not act[n_act:].any()
    s0 = _oracle_scene(gm)
    scene = orc.Scene(means=s0.means[perm], scales=s0.scales[perm], rots=s0.rots[perm], opac=s0.opac[perm],
                      shs=s0.shs[perm], sh_degree=s0.sh_degree)
    cams = _cam_records(tr)
    inten = tr.inten.cpu().numpy().reshape(tr.R, H, W)
    fgid = tr.fgid.cpu().numpy().view(np.uint32).reshape(tr.R, tr.cap) & np.uint32(tr.gid_mask)
    ranges = tr.ranges.cpu().numpy().view(np.uint32).reshape(tr.R, -1)
    nkept = tr.nkept.cpu().numpy()
    for r, cam in enumerate(cams):
        pre = orc.preprocess(scene, cam)
        b = orc.bin(pre, cam)
        check_culled_binning(fgid[r, :nkept[r]], ranges[r], b, pre, cam)
        rgb = pre["rgb"].reshape(-1, 3)
        I = ((GREY[0] * rgb[:, 0] + GREY[1] * rgb[:, 1]) + GREY[2] * rgb[:, 2]).astype(np.float32)
        fw = orc.blend_forward(pre, b, cam, colors=I, C=1, n_touched=False)
        assert np.array_equal(inten[r].view(np.uint32), fw["color"][0].view(np.uint32))
    np.testing.assert_allclose(tr.loss.cpu().numpy(), ref.loss.cpu().numpy(), rtol=1e-6)
    g, gr = tr.grad12.cpu().numpy(), ref.grad12.cpu().numpy()
    rel = np.abs(g - gr).max() / np.abs(gr).max()
    print(f"[fused] reordered map: renders bit-exact vs permuted oracle, gradient rel diff {rel:.1e}", flush=True)
    assert rel < 1e-5
