"""GPU event frames vs the numpy oracle: integer counts bit-exact (uneven
windows, out-of-image events rejected), undistort + blur + normalise within
1e-6 absolute, nearest pyramid bit-exact."""
import numpy as np
import pytest
import torch

from oracle import events as oev

pytestmark = pytest.mark.gpu

K_VECTOR = np.array([[327.32749, 0, 304.97749], [0, 327.46184, 235.37621], [0, 0, 1.0]])
DIST_VECTOR = np.array([-0.031982, 0.041966, -0.000507, -0.001031, 0.0])   # configs/VECTOR/desk_normal1_config.yaml:3-8


@pytest.mark.parametrize("n", [0, 1, 8000, 200000])
def test_event_frame_matches_oracle(cuda, n):
    from gsevt.frame import event_frame, image_pyramid
    print(f"[events] n={n}", flush=True)
    W, H = 640, 480
    rng = np.random.default_rng(n)
    x = rng.integers(-3, W + 3, n)
    y = rng.integers(-3, H + 3, n)
    p = rng.integers(0, 2, n)
    sign, absf, count = event_frame(x, y, p, W, H, K_VECTOR, DIST_VECTOR, 9, cuda)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(count.cpu().numpy(), oev.scatter(x, y, p, W, H))
    if n == 0:
        assert torch.all(sign == 0)
        return
    s_ref, a_ref = oev.event_frame(x, y, p, W, H, K_VECTOR, DIST_VECTOR)
    s = sign[0].cpu().numpy()
    assert np.abs(s - s_ref).max() <= 1e-6
    np.testing.assert_array_equal(absf[0].cpu().numpy(), np.abs(s))
    pyr = image_pyramid(sign, 3)
    for lvl in range(3):
        dw, dh = int(W * 0.5 ** lvl), int(H * 0.5 ** lvl)
        np.testing.assert_array_equal(pyr[lvl][0].cpu().numpy(), oev.resize_nearest(s, dw, dh))


def test_pyramid_odd_sizes(cuda):
    from gsevt.frame import image_pyramid
    img = torch.arange(346 * 260, dtype=torch.float32, device=cuda).reshape(1, 260, 346)
    pyr = image_pyramid(img, 3)
    for lvl in range(3):
        dw, dh = int(346 * 0.5 ** lvl), int(260 * 0.5 ** lvl)
        np.testing.assert_array_equal(pyr[lvl][0].cpu().numpy(), oev.resize_nearest(img[0].cpu().numpy(), dw, dh))


def test_event_frame_class_follows_numpy_indexing(cuda):
    """EventFrame (event.py:116-120 indexes a numpy array): coordinates in
    [-W, 0) wrap to the other edge, coordinates beyond raise IndexError."""
    from gsevt.frame import EventArray, EventFrame, event_frame
    W, H = 64, 48
    K = np.array([[50.0, 0, 32.0], [0, 50.0, 24.0], [0, 0, 1.0]])
    ts = np.arange(4, dtype=np.int64)
    ea = EventArray(ts, np.array([-1, 3, -64, 10]), np.array([5, -48, 0, -1]), np.array([1, 0, 1, 1]))
    ef = EventFrame(W, H, K, np.zeros(5), 9, ea, device=cuda)
    _, _, ref = event_frame(np.array([63, 3, 0, 10]), np.array([5, 0, 0, 47]), np.array([1, 0, 1, 1]), W, H, K,
                            np.zeros(5), 9, cuda)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(ef.count.cpu().numpy(), ref.cpu().numpy())
    for xs, ys in (([64], [0]), ([0], [48]), ([-65], [0]), ([0], [-49])):
        bad = EventArray(np.zeros(1, np.int64), np.array(xs), np.array(ys), np.array([1]))
        with pytest.raises(IndexError):
            EventFrame(W, H, K, np.zeros(5), 9, bad, device=cuda)
