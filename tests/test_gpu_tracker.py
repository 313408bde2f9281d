"""The plain tracker's full control flow (utils/tracker.py:105-286) on a small
seeded problem: 3 pyramid levels, coarse stage switching to fine, iteration
caps, velocity blend, TUM rows; and that tracking reduces the pose error."""
import math
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _config(W, H, fx, fy, max_iter):
    return {
        "Event": {"img_width": W, "img_height": H, "gaussian_kernel_size": 9,
                  "intrinsic": {"data": [fx, 0, W / 2, 0, fy, H / 2, 0, 0, 1]},
                  "distortion_factors": [0, 0, 0, 0, 0]},
        "Optimizer": {"cam_rot_delta": 0.004, "cam_trans_delta": 0.004, "cam_w_delta": 0.002,
                      "cam_v_delta": 0.002, "converged_threshold": 1e-4, "max_optim_iter": max_iter},
    }


def test_plain_tracker_control_flow(cuda):
    from gsevt import workloads
    from gsevt.tracker import Tracker
    prob = workloads.build("cpu_small", device=cuda)
    print(f"[tracker] cpu_small: {prob.gaussians.num_points} Gaussians, {prob.events.size()} events", flush=True)
    cam = prob.camera()
    cfg = _config(prob.intr.W, prob.intr.H, prob.intr.fx, prob.intr.fy, max_iter=12)
    tr = Tracker(cfg, [prob.events, prob.events], cam, prob.gaussians, None, prob.background, device=cuda)
    traj = tr.tracking()
    assert len(traj) == 2 and all(len(r) == 8 for r in traj)
    for info in tr.window_log:
        lv = [l["level"] for l in info["levels"]]
        assert lv == [2, 1, 0]
        for l in info["levels"]:
            # a stage ends on convergence or at its cap (max_optim_iter past the fine-stage start)
            assert l["iters"] <= l["start_vel"] + 12
            assert math.isfinite(l["final_loss"])
    q = np.array(traj[-1][4:8])
    assert abs(np.linalg.norm(q) - 1) < 1e-5
    assert tr.iterations > 0


def test_plain_tracker_writes_run_outputs(cuda, tmp_path):
    """With Tracking.save_path a run writes the reference's outputs
    (tracker.py:105-288): log, TUM file, one overlay frame per window, GIF."""
    from PIL import Image
    from gsevt import workloads
    from gsevt.tracker import Tracker
    prob = workloads.build("cpu_small", device=cuda)
    cfg = _config(prob.intr.W, prob.intr.H, prob.intr.fx, prob.intr.fy, max_iter=4)
    cfg["Tracking"] = {"save_path": str(tmp_path / "run")}
    tr = Tracker(cfg, [prob.events, prob.events], prob.camera(), prob.gaussians, None, prob.background,
                 device=cuda)
    traj = tr.tracking()
    out = tmp_path / "run"
    rows = (out / "tracking_pose_tum.txt").read_text().splitlines()
    assert len(rows) == 2 and [float(v) for v in rows[1].split()] == [float(v) for v in traj[1]]
    log_text = (out / "tracking_log.log").read_text()
    assert "frame_idx:\t1 / 2" in log_text and "optim_iter:" in log_text and "save gif in" in log_text
    frames = sorted(os.listdir(out / "tracking_frames"))
    assert frames == ["frame_0.png", "frame_1.png"]
    im = Image.open(out / "tracking_frames" / "frame_1.png")
    assert im.size == (prob.intr.W, prob.intr.H) and im.mode == "RGB"
    assert Image.open(out / "tracking_video.gif").n_frames == 2
    print(f"[tracker] run outputs in {out}: {frames}", flush=True)


def test_plain_fine_iterations_reduce_pose_error(cuda):
    from gsevt import workloads
    from gsevt.tracker import plain_fine_iterations
    prob = workloads.build("cpu_small", device=cuda)
    cam = prob.camera()
    e0 = float((cam.T - prob.gt_pose[:3, 3]).norm())
    losses = []
    plain_fine_iterations(cam, prob.gaussians, prob.background, prob.sign_Ie, 40, (0.004, 0.004, 0.002, 0.002),
                          losses=losses)
    e1 = float((cam.T - prob.gt_pose[:3, 3]).norm())
    print(f"[tracker] translation error {e0:.4f} -> {e1:.4f} m, loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    assert losses[-1] < losses[0]


def test_render_mask_branch_equals_subset_render(cuda):
    """gaussian_renderer/__init__.py:344-357: a mask renders only the selected
    Gaussians (n_touched None); gradients reach only the selected ones."""
    import torch
    from gsevt import workloads
    from gsevt.renderer import render1
    from gsevt.model import GaussianModel
    prob = workloads.build("cpu_small", device=cuda)
    g = prob.gaussians
    cam = prob.camera()
    bg = prob.background
    mask = torch.zeros(g.get_xyz.shape[0], dtype=torch.bool, device=cuda)
    mask[::3] = True
    out_m = render1(cam, g, bg, mask=mask)
    c = lambda t: t[mask].cpu().numpy()
    sub = GaussianModel.from_raw(c(g._xyz), c(g._features_dc), c(g._features_rest), c(g._scaling), c(g._rotation),
                                 c(g._opacity), g.active_sh_degree, device=cuda)
    out_s = render1(cam, sub, bg)
    assert out_m["n_touched"] is None
    assert torch.equal(out_m["render"], out_s["render"]) and torch.equal(out_m["radii"], out_s["radii"])
    assert torch.equal(out_m["depth"], out_s["depth"])
