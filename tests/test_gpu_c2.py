"""The drop-in baseline configuration on the GPU: 300k anisotropic Gaussians
on 64 planar patches (SH degree 3), 640x480, fx = fy = 550, 200k events of
one 10 ms window drawn from the rendered intensity change under a
ground-truth twist (0.3 m/s, 20 deg/s).

One tracking iteration of the plain path (RenderFrame -> tracking_loss ->
autograd through diff_gaussian_rasterization, as utils/tracker.py:204-213
runs it) against the FP32 oracle on the exact float32 settings each of the
two rasterizer calls received: radii, 64-bit keys, values and tile ranges
bit-exact, the colour render bit-exact (the bar is 1e-5), the loss within
1e-5 and the 12-float gradient (theta, rho, w, v) within 1e-3 relative.
"""
import numpy as np
import pytest
import torch

from helpers import orc, oracle_scene_from_model
from oracle import tracking_step

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(900)
def test_plain_dropin_300k_iteration_vs_oracle(cuda, monkeypatch):
    import diff_gaussian_rasterization as dgr
    from gsevt import raster, workloads
    from gsevt.frame import RenderFrame
    from gsevt.tracker import tracking_loss
    prob = workloads.build("patches_300k", device=cuda)
    assert prob.gaussians.num_points == 300000 and prob.events.size() == 200000
    calls = []
    orig_rg, orig_rf = dgr.rasterize_gaussians, raster.render_forward

    def rec_settings(*args):
        calls.append({"settings": args[-1]})
        return orig_rg(*args)

    def rec_forward(*args, **kw):
        out, state = orig_rf(*args, **kw)
        calls[-1]["out"], calls[-1]["state"] = out, state
        return out, state

    monkeypatch.setattr(dgr, "rasterize_gaussians", rec_settings)
    monkeypatch.setattr(raster, "render_forward", rec_forward)
    cam = prob.camera()
    for p in (cam.cam_rot_delta, cam.cam_trans_delta, cam.cam_w_delta, cam.cam_v_delta):
        p.requires_grad_(True)
    rf = RenderFrame(cam, prob.gaussians, None, prob.background, 0)
    loss = tracking_loss(rf.sign_delta_Ir, prob.sign_Ie)
    loss.backward()
    torch.cuda.synchronize()
    g12 = torch.cat([cam.cam_rot_delta.grad, cam.cam_trans_delta.grad, cam.cam_w_delta.grad,
                     cam.cam_v_delta.grad]).cpu().numpy()
    assert len(calls) == 2
    scene = oracle_scene_from_model(prob.gaussians)
    f = lambda t: t.detach().float().cpu().numpy().reshape(-1)
    states = []
    for k, c in enumerate(calls):
        s = c["settings"]
        ocam = orc.Cam(view=f(s.viewmatrix), proj=f(s.projmatrix), projraw=f(s.projmatrix_raw), campos=f(s.campos),
                       W=s.image_width, H=s.image_height, tanfovx=float(np.float32(s.tanfovx)),
                       tanfovy=float(np.float32(s.tanfovy)), bg=f(s.bg), tile_w=16, tile_h=16,
                       vel=f(s.vel_transofrm).astype(np.float64), delta_time=float(s.delta_time),
                       lin=f(s.linear_vel).astype(np.float64), ang=f(s.angular_vel).astype(np.float64))
        assert (float(s.delta_time) < 0) == (k == 0), "the first call renders the last pose"
        st = orc.render(scene, ocam, n_touched=False)
        pre, b, fw = st["pre"], st["bin"], st["fwd"]
        out, gst = c["out"], c["state"]
        np.testing.assert_array_equal(out["radii"].detach().cpu().numpy(), pre["radii"])
        assert gst["n"] == b["n"]
        np.testing.assert_array_equal(gst["keys"][:b["n"]].cpu().numpy().view(np.uint64), b["keys"])
        np.testing.assert_array_equal(gst["vals"][:b["n"]].cpu().numpy().view(np.uint32), b["vals"])
        np.testing.assert_array_equal(gst["ranges"].cpu().numpy().view(np.uint32), b["ranges"])
        col = out["color"].detach().cpu().numpy()
        assert np.abs(col - fw["color"]).max() <= 1e-5
        assert np.array_equal(col.view(np.uint32), fw["color"].view(np.uint32))
        print(f"[c2] render {k}: N_f = {b['n']}, visible {int((pre['radii'] > 0).sum())}, binning and render "
              "bit-exact", flush=True)
        states.append((ocam, st))
    (cl, st_l), (cn, st_n) = states
    L, g_l, g_n, _ = tracking_step.delta_ir_and_grad(st_l["fwd"]["color"], st_n["fwd"]["color"],
                                                     prob.sign_Ie[0].cpu().numpy())
    assert abs(L - float(loss)) < 1e-5, (L, float(loss))
    bl = orc.render_backward(scene, cl, st_l, g_l.astype(np.float32))
    bn = orc.render_backward(scene, cn, st_n, g_n.astype(np.float32))
    ref = np.concatenate([bl["pose"][k] + bn["pose"][k] for k in ("theta", "rho", "w", "v")])
    rel = np.abs(g12 - ref).max() / np.abs(ref).max()
    print(f"[c2] loss {float(loss):.6f} (oracle {L:.6f}), 12-float gradient rel err {rel:.2e}", flush=True)
    assert rel < 1e-3
