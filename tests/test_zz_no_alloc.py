"""gsevt_event_frame_noalloc (static device scratch, no allocation) against
gsevt_event_frame (caller workspace), both called through the C ABI: two
calls in a row, the second on a smaller frame, outputs bit-identical and
within 1e-6 of the numpy restatement."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import events as oev

pytestmark = pytest.mark.gpu

K_VECTOR = np.array([[327.32749, 0, 304.97749], [0, 327.46184, 235.37621], [0, 0, 1.0]])
DIST_VECTOR = np.array([-0.031982, 0.041966, -0.000507, -0.001031, 0.0])


def _count_and_map(L, cuda, W, H, n, seed, K, dist):
    rng = np.random.default_rng(seed)
    x = rng.integers(0, W, n).astype(np.int32)
    y = rng.integers(0, H, n).astype(np.int32)
    p = rng.integers(0, 2, n).astype(np.int32)
    t = lambda a: torch.from_numpy(a).to(cuda)
    xt, yt, pt = t(x), t(y), t(p)
    count = torch.zeros(H, W, dtype=torch.int32, device=cuda)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    P = lambda a: ctypes.c_void_p(a.data_ptr())
    assert L.gsevt_event_scatter(n, P(xt), P(yt), P(pt), W, H, P(count), None, s) == 0
    Kd = torch.tensor(K.reshape(-1), dtype=torch.float64, device=cuda)
    dd = torch.tensor(dist, dtype=torch.float64, device=cuda)
    m = torch.empty(H * W * 2, dtype=torch.int32, device=cuda)
    assert L.gsevt_undistort_map(W, H, P(Kd), P(dd), P(m), s) == 0
    return (x, y, p), count, m


def test_event_frame_noalloc_matches_workspace_entry_point(cuda):
    from gsevt import _lib
    L = _lib.load()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    P = lambda a: ctypes.c_void_p(a.data_ptr())
    # larger frame first, then a smaller one: the second call finds the first
    # call's values in the shared static scratch
    cases = [(640, 480, 200000, 1, K_VECTOR, DIST_VECTOR),
             (346, 260, 8000, 2, np.array([[300.0, 0, 173.0], [0, 300.0, 130.0], [0, 0, 1]]), np.zeros(5))]
    for W, H, n, seed, K, dist in cases:
        print(f"[noalloc] event frame {W}x{H}, {n} events", flush=True)
        ev, count, m = _count_and_map(L, cuda, W, H, n, seed, K, dist)
        ws = torch.empty(int(L.gsevt_event_frame_scratch_bytes(W, H)), dtype=torch.uint8, device=cuda)
        ref_s = torch.full((H, W), float("nan"), device=cuda)
        ref_a = torch.full((H, W), float("nan"), device=cuda)
        got_s = torch.full((H, W), float("nan"), device=cuda)
        got_a = torch.full((H, W), float("nan"), device=cuda)
        assert L.gsevt_event_frame(W, H, P(count), P(m), 9, 0.0, P(ws), P(ref_s), P(ref_a), s) == 0
        rc = L.gsevt_event_frame_noalloc(W, H, P(count), P(m), 9, 0.0, P(got_s), P(got_a), s)
        assert rc == 0, L.gsevt_last_error()
        torch.cuda.synchronize()
        gs, ga = got_s.cpu().numpy(), got_a.cpu().numpy()
        assert np.isfinite(gs).all() and np.isfinite(ga).all()
        np.testing.assert_array_equal(gs.view(np.uint32), ref_s.cpu().numpy().view(np.uint32))
        np.testing.assert_array_equal(ga.view(np.uint32), ref_a.cpu().numpy().view(np.uint32))
        s_ref, _ = oev.event_frame(*ev, W, H, K, dist)
        assert np.abs(gs - s_ref).max() <= 1e-6

