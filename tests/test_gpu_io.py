"""On-disk inputs through the GPU path: a 3DGS PLY map written in the
reference's attribute order renders bit-identically to the scene it was
written from (drop-in rasterizer), and events read from events.txt give the
same integer event-frame counts as the numpy scatter of the same events."""
import os
import sys

import numpy as np
import pytest
import torch

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from helpers import gpu_cam, make_cam          # noqa: E402
from ply_writer import write_ply, gaussian_vertex   # noqa: E402
from oracle import events as oev               # noqa: E402

pytestmark = pytest.mark.gpu


def test_ply_map_renders_like_source_scene(cuda, tmp_path):
    from gsevt import io as gio, raster, synthetic
    from gsevt.model import GaussianModel
    raw = synthetic.patch_scene(P=50000, sh_degree=3, seed=77)
    path = str(tmp_path / "map.ply")
    write_ply(path, gaussian_vertex(raw))
    a = gio.load_ply(path, max_sh_degree=3, device=cuda)
    b = GaussianModel.from_raw(raw.xyz, raw.features_dc, raw.features_rest, raw.scaling, raw.rotation,
                               raw.opacity, raw.sh_degree, device=cuda)
    cam = gpu_cam(make_cam(np.eye(4), 320, 240, 275.0, 275.0), cuda)

    def render(gm):
        m = raster.GaussianMap(means3D=gm.get_xyz, opacities=gm.get_opacity, shs=gm.get_features,
                               colors_precomp=None, scales=gm.get_scaling, rotations=gm.get_rotation,
                               cov3D_precomp=None, sh_degree=gm.active_sh_degree, scale_modifier=1.0)
        out, _ = raster.render_forward(m, cam, channels=3, n_touched=False)
        return out["color"].cpu().numpy()
    ca, cb = render(a), render(b)
    print(f"[io] PLY map of {raw.P} Gaussians: render bit-identical = {np.array_equal(ca, cb)}", flush=True)
    assert np.array_equal(ca.view(np.uint32), cb.view(np.uint32))


def test_events_txt_frame_counts(cuda, tmp_path):
    from gsevt import io as gio
    from gsevt.frame import event_frame
    W, H = 346, 260
    rng = np.random.default_rng(21)
    n = 30000
    ts = np.sort(rng.integers(0, 10000, n)) + 1_000_000
    x = rng.integers(0, W, n)
    y = rng.integers(0, H, n)
    p = rng.integers(0, 2, n)
    path = str(tmp_path / "events.txt")
    with open(path, "w") as f:
        f.writelines(f"{a} {b} {c} {d}\n" for a, b, c, d in zip(ts, x, y, p))
    wins = gio.load_events_from_txt(path, 10000)
    assert len(wins) == 3
    K = np.array([[300.0, 0, W / 2], [0, 300.0, H / 2], [0, 0, 1]])
    for k, w in enumerate(wins):
        _, _, count = event_frame(w.x, w.y, w.p, W, H, K, np.zeros(5), 9, cuda)
        torch.cuda.synchronize()
        sl = slice(k * 10000, (k + 1) * 10000)
        np.testing.assert_array_equal(count.cpu().numpy(), oev.scatter(x[sl], y[sl], p[sl], W, H))


def test_bag_events_to_gpu_event_frame(cuda, tmp_path):
    """Events read from a ROS1 bag (utils/event_camera/event.py:42-69, format
    2.0 written by tests/bagfile.py with bz2 chunks) through the reference's
    EventFrame class on the GPU: each window's integer counts equal the numpy
    scatter of the same events, and the frame is unit-norm."""
    from bagfile import BagWriter, EVENT_ARRAY_DEF, encode_event_array
    from gsevt import io as gio
    from gsevt.frame import EventFrame
    W, H = 346, 260
    rng = np.random.default_rng(5)
    path = str(tmp_path / "events.bag")
    w = BagWriter(path, "bz2", 4096)
    conn = w.add_connection(gio.BAG_TOPIC, "dvs_msgs/EventArray", EVENT_ARRAY_DEF.format(pkg="dvs_msgs"))
    allx, ally, allp = [], [], []
    for i in range(40):
        k = int(rng.integers(100, 400))
        xs, ys, ps = rng.integers(0, W, k), rng.integers(0, H, k), rng.integers(0, 2, k)
        ev = [(int(a), int(b), 100 + i // 10, (i % 10) * 1000000 + j * 10, int(c)) for j, (a, b, c) in
              enumerate(zip(xs, ys, ps))]
        w.write(conn, 100 + i // 10, (i % 10) * 1000000, encode_event_array(i, 100 + i // 10, (i % 10) * 1000000, ev,
                                                                             height=H, width=W))
        allx.append(xs), ally.append(ys), allp.append(ps)
    w.close()
    x, y, p = np.concatenate(allx), np.concatenate(ally), np.concatenate(allp)
    M = 2500
    wins = gio.load_events_from_bag(path, M)
    assert len(wins) == x.shape[0] // M
    K = np.array([[300.0, 0, W / 2], [0, 300.0, H / 2], [0, 0, 1]])
    for k, win in enumerate(wins):
        ef = EventFrame(W, H, K, np.zeros(5), 9, win, device=cuda)
        torch.cuda.synchronize()
        sl = slice(k * M, (k + 1) * M)
        np.testing.assert_array_equal(ef.count.cpu().numpy(), oev.scatter(x[sl], y[sl], p[sl], W, H))
        assert abs(float(ef.sign_delta_Ie.double().norm()) - 1.0) < 1e-5
    print(f"[io] bag: {x.shape[0]} events in 40 messages -> {len(wins)} windows of {M}, GPU counts exact", flush=True)
