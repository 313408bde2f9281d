"""Parity at the benchmark's own shape: 1M Gaussians (SH3, 64 planar
patches), 640x480, 64 pose hypotheses x 4 event windows = 256 items, i.e.
512 renders in one launch per stage (the configuration bench.py times).

For items 0, 63, 64, 191 and 255 (first and last hypothesis of windows 0, 1
and 3) and both renders of each, against the FP32 oracle on the camera
records the device computed:
* the tile-culled fragment lists and ranges are the oracle's binning minus
  fragments that reach no pixel of their tile (check_culled_binning);
* the grey render is bit-exact;
* the loss is within 1e-5 and the 12-float gradient (theta, rho, w, v) within
  1e-3 relative of the oracle chain.
Then 5 graph-replayed Adam iterations of all 256 items leave those items'
poses within 1e-6 of single-item trackers run on the same items alone
(utils/tracker.py:176-240 per item), losses within 1e-4.
"""
import numpy as np
import pytest
import torch

from helpers import orc, device_cam_records, oracle_scene_from_model, check_culled_binning
from oracle import tracking_step

pytestmark = pytest.mark.gpu

GREY = np.array([0.2989, 0.5870, 0.1140], np.float32)
ITEMS = (0, 63, 64, 191, 255)


@pytest.fixture(scope="module")
def prob(cuda):
    from gsevt import workloads
    return workloads.build_batched(P=1_000_000, n_hyp=64, n_windows=4, device=cuda)


def _items(prob):
    B = len(prob.item_window)
    lin = prob.gt_lin.expand(B, 3)
    ang = prob.gt_ang.expand(B, 3)
    dtau = [prob.dtau[w] for w in prob.item_window]
    return B, lin, ang, dtau


@pytest.mark.timeout(1200)
def test_256_items_one_launch_vs_oracle(cuda, prob):
    from gsevt.fused import BatchedTracker
    intr = prob.intr
    B, lin, ang, dtau = _items(prob)
    assert B == 256
    tr = BatchedTracker(prob.gaussians, intr.W, intr.H, intr.fx, intr.fy, prob.events, B, device=cuda)
    assert tr.n_chunks == 1 and tr.R == 512
    tr.set_items(prob.init_poses, lin, ang, dtau, prob.item_window)
    nmax = tr.measure_capacity()
    print(f"\n[headline] 256 items, 512 renders per launch, capacity {tr.cap} (max {nmax})", flush=True)
    tr.step(do_step=False)
    torch.cuda.synchronize()
    tr.check_overflow()
    scene = oracle_scene_from_model(prob.gaussians)
    cams = device_cam_records(tr)
    W, H = intr.W, intr.H
    nfrag = tr.nfrag.cpu().numpy()
    nkept = tr.nkept.cpu().numpy()
    vel, dts = tr.vel.cpu().numpy(), tr.dts.cpu().numpy()
    loss, g12 = tr.loss.cpu().numpy(), tr.grad12.cpu().numpy()
    fg = tr.fgid.view(tr.R, tr.cap)
    rg = tr.ranges.view(tr.R, -1)
    im = tr.inten.view(tr.R, H * W)
    lin_np, ang_np = lin.cpu().numpy(), ang.cpu().numpy()
    for b in ITEMS:
        states = []
        for side in (0, 1):
            r = 2 * b + side
            cam = cams[r]
            cam.vel = vel[r].astype(np.float64)
            cam.delta_time = float(dts[r])
            cam.lin, cam.ang = lin_np[b], ang_np[b]
            st = orc.render(scene, cam, n_touched=False)
            bn = st["bin"]
            assert nfrag[r] == bn["n"]
            gids = fg[r, :nkept[r]].cpu().numpy().view(np.uint32) & np.uint32(tr.gid_mask)
            kept, _ = check_culled_binning(gids, rg[r].cpu().numpy().view(np.uint32), bn, st["pre"], cam)
            assert kept == nkept[r]
            rgb = st["pre"]["rgb"].reshape(-1, 3)
            I = ((GREY[0] * rgb[:, 0] + GREY[1] * rgb[:, 1]) + GREY[2] * rgb[:, 2]).astype(np.float32)
            fw = orc.blend_forward(st["pre"], bn, cam, colors=I, C=1, n_touched=False)
            got = im[r].cpu().numpy().reshape(H, W)
            assert np.array_equal(got.view(np.uint32), fw["color"][0].view(np.uint32)), f"render {r}"
            states.append((cam, st))
            print(f"[headline] item {b} render {r}: {bn['n']} slots, {nkept[r]} kept, binning checked, "
                  "grey render bit-exact", flush=True)
        (cl, st_l), (cn, st_n) = states
        e = prob.events[prob.item_window[b]].cpu().numpy()
        L, g_l, g_n, _ = tracking_step.delta_ir_and_grad(st_l["fwd"]["color"], st_n["fwd"]["color"], e)
        assert abs(L - float(loss[b])) < 1e-5, (b, L, loss[b])
        bl = orc.render_backward(scene, cl, st_l, g_l.astype(np.float32))
        bnx = orc.render_backward(scene, cn, st_n, g_n.astype(np.float32))
        ref = np.concatenate([bl["pose"][k] + bnx["pose"][k] for k in ("theta", "rho", "w", "v")])
        rel = np.abs(g12[b] - ref).max() / np.abs(ref).max()
        print(f"[headline] item {b}: loss {loss[b]:.6f} (oracle {L:.6f}), gradient rel err {rel:.2e}", flush=True)
        assert rel < 1e-3
        del states, st_l, st_n, bl, bnx


@pytest.mark.timeout(1200)
def test_256_items_graph_replay_matches_single_item_trackers(cuda, prob):
    from gsevt.fused import BatchedTracker
    intr = prob.intr
    B, lin, ang, dtau = _items(prob)
    n_it = 5
    tr = BatchedTracker(prob.gaussians, intr.W, intr.H, intr.fx, intr.fy, prob.events, B, device=cuda)
    tr.set_items(prob.init_poses, lin, ang, dtau, prob.item_window)
    tr.measure_capacity()
    tr.capture()
    tr.set_items(prob.init_poses, lin, ang, dtau, prob.item_window)
    for _ in range(n_it):
        tr.graph.replay()
    torch.cuda.synchronize()
    tr.check_overflow()
    poses = tr.poses().cpu().numpy()
    hist = tr.loss_hist[:, :n_it].cpu().numpy()
    moved = np.abs(poses[:, :3, 3] - prob.init_poses.double().cpu().numpy()[:, :3, 3]).max()
    del tr
    torch.cuda.empty_cache()
    for b in ITEMS:
        w = prob.item_window[b]
        one = BatchedTracker(prob.gaussians, intr.W, intr.H, intr.fx, intr.fy, prob.events[w:w + 1], 1,
                             device=cuda)
        one.set_items(prob.init_poses[b:b + 1], lin[b:b + 1], ang[b:b + 1], [dtau[b]], [0])
        one.measure_capacity()
        one.capture()
        one.set_items(prob.init_poses[b:b + 1], lin[b:b + 1], ang[b:b + 1], [dtau[b]], [0])
        for _ in range(n_it):
            one.graph.replay()
        torch.cuda.synchronize()
        one.check_overflow()
        p1 = one.poses()[0].cpu().numpy()
        d = np.abs(p1 - poses[b]).max()
        print(f"[headline] item {b}: |pose(256-item launch) - pose(alone)| = {d:.2e} after {n_it} iterations; "
              f"losses {hist[b, 0]:.5f} -> {hist[b, -1]:.5f}", flush=True)
        # the backward's float atomics add in a different order in the two
        # launches; the resulting ~1e-8 pose differences can move a fragment
        # across the alpha or transmittance threshold of one pixel, which
        # changes a loss by up to ~1e-5 (measured 1.7e-5) without moving poses
        np.testing.assert_allclose(one.loss_hist[0, :n_it].cpu().numpy(), hist[b], rtol=0, atol=1e-4)
        assert d < 1e-6
        del one
    assert moved > 1e-4
