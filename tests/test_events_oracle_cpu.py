"""Event-frame oracle (oracle/events.py) pinned by closed forms and brute force.

OpenCV is absent, so its undistort / GaussianBlur are restated from their
published definitions; these tests pin the restatement so that a dropped
term, a wrong sign, swapped coefficients, transposed weights or another
border mode fails:

* undistortion map: with one distortion coefficient at a time, pixels on the
  principal axes reduce to one-line closed forms of the Brown-Conrady model
  (k1, k2, k3: radial x (1 + k r^2n); p1 on the y axis: y + 3 p1 y^2, on the
  x axis: a vertical shift p1 x^2; p2 the transpose), to the 1/32 pixel grid;
* bilinear remap: reproduces an affine image exactly at any 1/32-pixel
  position inside the image, and reads 0 outside (BORDER_CONSTANT);
* Gaussian kernel: sigma 0 at ksize 9 is 0.3 ((9 - 1) / 2 - 1) + 0.8 = 1.7
  (getGaussianKernel's documented rule) and the weights are
  exp(-x^2 / (2 sigma^2)) normalised, to float32 rounding;
* blur: equals a brute-force 2-D convolution with the outer-product kernel
  over an edge-replicated border (BORDER_REPLICATE), and differs from the
  same convolution over a reflect-101 border (OpenCV's default), so the
  check can tell the two apart;
* nearest resize: every other pixel from the first at a factor of 2.
"""
import numpy as np
import pytest

from oracle import events as oev

W, H = 64, 48
FX, FY, CX, CY = 50.0, 45.0, 32.0, 24.0
K = np.array([[FX, 0, CX], [0, FY, CY], [0, 0, 1.0]])


def _grid(v):
    return np.rint(np.asarray(v, np.float64) * 32).astype(np.int64)


@pytest.mark.parametrize("name,idx,val", [("k1", 0, -0.21), ("k2", 1, 0.07), ("k3", 4, -0.02),
                                          ("p1", 2, 0.003), ("p2", 3, -0.004)])
def test_undistort_map_axis_closed_forms(name, idx, val):
    dist = np.zeros(5)
    dist[idx] = val
    iu, iv = oev.undistort_map(W, H, K, dist)
    j = np.arange(W, dtype=np.float64)
    i = np.arange(H, dtype=np.float64)
    x = (j - CX) / FX            # along the row through the principal point (y = 0)
    y = (i - CY) / FY            # along the column through it (x = 0)
    power = {"k1": 1, "k2": 2, "k3": 3}
    if name in power:
        n = power[name]
        u_row, v_row = CX + FX * x * (1 + val * x ** (2 * n)), np.full(W, CY)
        u_col, v_col = np.full(H, CX), CY + FY * y * (1 + val * y ** (2 * n))
    elif name == "p1":
        u_row, v_row = CX + FX * x, CY + FY * val * x * x
        u_col, v_col = np.full(H, CX), CY + FY * (y + 3 * val * y * y)
    else:   # p2: the transpose of p1
        u_row, v_row = CX + FX * (x + 3 * val * x * x), np.full(W, CY)
        u_col, v_col = CX + FX * val * y * y, CY + FY * y
    r, c = int(CY), int(CX)
    np.testing.assert_array_equal(iu[r], _grid(u_row))
    np.testing.assert_array_equal(iv[r], _grid(v_row))
    np.testing.assert_array_equal(iu[:, c], _grid(u_col))
    np.testing.assert_array_equal(iv[:, c], _grid(v_col))
    # the coefficient moves points: the map is not the identity
    assert np.any(iu != _grid(np.arange(W)[None].repeat(H, 0))) or np.any(iv != _grid(np.arange(H)[:, None]
                                                                                       .repeat(W, 1)))


def test_remap_bilinear_reproduces_affine_images():
    a, b, c = 0.75, -1.25, 3.0
    jj, ii = np.meshgrid(np.arange(W, dtype=np.float64), np.arange(H, dtype=np.float64))
    img = (a * jj + b * ii + c).astype(np.float32)
    rng = np.random.default_rng(0)
    n = 4000
    iu = rng.integers(0, (W - 1) * 32, n)          # inside: the 2x2 neighbourhood is in the image
    iv = rng.integers(0, (H - 1) * 32, n)
    got = oev.remap_bilinear(img, iu, iv)
    want = a * iu / 32.0 + b * iv / 32.0 + c
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-5 * np.abs(want).max())
    # fully outside: BORDER_CONSTANT 0
    out = oev.remap_bilinear(img, np.array([-64, W * 32 + 40, 10 * 32]), np.array([5 * 32, 5 * 32, -80]))
    np.testing.assert_array_equal(out, np.zeros(3, np.float32))


def test_gaussian_kernel_closed_form():
    k = oev.gaussian_kernel(9, 0.0)
    sigma = 0.3 * ((9 - 1) * 0.5 - 1) + 0.8
    assert abs(sigma - 1.7) < 1e-15
    x = np.arange(9) - 4.0
    g = np.exp(-x * x / (2 * sigma * sigma))
    np.testing.assert_allclose(k, g / g.sum(), rtol=2e-7, atol=0)
    np.testing.assert_allclose(oev.gaussian_kernel(9, 2.5), np.exp(-x * x / 12.5) / np.exp(-x * x / 12.5).sum(),
                               rtol=2e-7, atol=0)


def _conv2d(img, k, mode):
    r = k.shape[0] // 2
    p = np.pad(img.astype(np.float64), r, mode=mode)
    k2 = np.outer(k.astype(np.float64), k.astype(np.float64))
    out = np.zeros(img.shape)
    for dy in range(k.shape[0]):
        for dx in range(k.shape[0]):
            out += k2[dy, dx] * p[dy:dy + img.shape[0], dx:dx + img.shape[1]]
    return out


def test_blur_is_replicate_border_convolution():
    rng = np.random.default_rng(1)
    img = rng.normal(size=(H, W)).astype(np.float32)
    img[:, 0] += 5.0            # strong edges make the border mode visible
    img[0, :] -= 4.0
    k = oev.gaussian_kernel(9, 0.0)
    got = oev.blur(img, 9)
    want = _conv2d(img, k, "edge")
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-5)
    reflect101 = _conv2d(img, k, "reflect")      # numpy "reflect" = OpenCV BORDER_REFLECT_101
    assert np.abs(reflect101 - want).max() > 0.1


def test_resize_nearest_factor_two_closed_form():
    """cv2.resize INTER_NEAREST samples src = floor(dst * src_size / dst_size):
    at an exact factor of 2 that is every other pixel from the first one
    (a pixel-centre rule would start at the second)."""
    img = np.arange(48 * 64, dtype=np.float32).reshape(48, 64)
    np.testing.assert_array_equal(oev.resize_nearest(img, 32, 24), img[::2, ::2])
    np.testing.assert_array_equal(oev.resize_nearest(img, 16, 12), img[::4, ::4])


PLANTED = [
    ("undistort_map", "xd = x * kr + p1 * _2xy + p2 * (r2 + 2 * x2)", "xd = x * kr + p2 * _2xy + p1 * (r2 + 2 * x2)",
     "test_undistort_map_axis_closed_forms"),
    ("undistort_map", "yd = y * kr + p1 * (r2 + 2 * y2) + p2 * _2xy", "yd = y * kr - p1 * (r2 + 2 * y2) + p2 * _2xy",
     "test_undistort_map_axis_closed_forms"),
    ("undistort_map", "kr = 1 + ((k3 * r2 + k2) * r2 + k1) * r2", "kr = 1 + ((k2 * r2 + k3) * r2 + k1) * r2",
     "test_undistort_map_axis_closed_forms"),
    ("remap_bilinear", "w00, w01 = (one - ay) * (one - ax), (one - ay) * ax",
     "w00, w01 = (one - ax) * (one - ay), (one - ax) * ay", "test_remap_bilinear_reproduces_affine_images"),
    ("gaussian_kernel", "sigma = ((ksize - 1) * 0.5 - 1) * 0.3 + 0.8", "sigma = ((ksize - 1) * 0.5) * 0.3 + 0.8",
     "test_gaussian_kernel_closed_form"),
    ("blur", "ys = np.clip(np.arange(H)[:, None] + np.arange(-r, r + 1)[None], 0, H - 1)",
     "ys = np.abs(np.arange(H)[:, None] + np.arange(-r, r + 1)[None]); ys = np.where(ys > H - 1, 2 * (H - 1) - ys, ys)",
     "test_blur_is_replicate_border_convolution"),
    ("resize_nearest", "xs = np.minimum(np.floor(np.arange(dw) * ifx)", "xs = np.minimum(np.floor((np.arange(dw) + 0.5) * ifx)",
     "test_resize_nearest_factor_two_closed_form"),
]


@pytest.mark.parametrize("fn,good,bad,test", PLANTED, ids=[f"{p[0]}-{i}" for i, p in enumerate(PLANTED)])
def test_event_oracle_pins_planted_mistakes(monkeypatch, fn, good, bad, test):
    """A plausible mistake planted in the oracle fails its pinning test."""
    import inspect
    import test_events_oracle_cpu as this
    src = inspect.getsource(getattr(oev, fn))
    assert good in src
    ns = dict(vars(oev))
    exec(src.replace(good, bad), ns)
    monkeypatch.setattr(oev, fn, ns[fn])
    check = getattr(this, test)
    with pytest.raises(AssertionError):
        if test == "test_undistort_map_axis_closed_forms":
            for args in [("k1", 0, -0.21), ("k2", 1, 0.07), ("k3", 4, -0.02), ("p1", 2, 0.003), ("p2", 3, -0.004)]:
                check(*args)
        else:
            check()
