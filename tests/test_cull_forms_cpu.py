"""The two forms of the conservative tile cull, pinned in exact arithmetic.

The blends and the emission drop a (splat, rectangle of pixel centres) pair
when the splat's ellipse q(dx, dy) = a dx^2 + 2 b dx dy + c dy^2 <= T misses
the rectangle (T = 2 L', alpha < 1/255 outside it).  The device has two
forms of that test (gs-evt_amd/csrc/gsevt_device.h):

* the rectangle form (cull_rect): the minimum of q over the rectangle, taken
  on the sides facing the mean;
* the band form (band_prep / band_reach, used by the packed emission): the
  ellipse's x-interval over the rectangle's rows, from its two branches
  x = -p dy +- sqrt(u - w dy^2) read at their clamped stationary points.

Both are restated here in float64 without their float32 margins and checked
against an independent minimum of q over the rectangle (its unconstrained
minimum is the mean; otherwise the minimum over the four edges, each a
one-dimensional convex quadratic), on random positive definite conics,
thresholds and rectangles, including the elongated conics (|b| close to
sqrt(a c)) where the band form's square root is ill-conditioned.
"""
import numpy as np


def min_q_rect(mx, my, a, b, c, x0, x1, y0, y1):
    """min of q(x - mx, y - my) over [x0, x1] x [y0, y1] (float64)."""
    def q(dx, dy):
        return a * dx * dx + 2.0 * b * dx * dy + c * dy * dy

    best = np.full(np.shape(mx), np.inf)
    for X in (x0, x1):                      # vertical edges: dy free in [y0, y1]
        dx = X - mx
        dy = np.clip(-b * dx / c, y0 - my, y1 - my)
        best = np.minimum(best, q(dx, dy))
    for Y in (y0, y1):                      # horizontal edges: dx free in [x0, x1]
        dy = Y - my
        dx = np.clip(-b * dy / a, x0 - mx, x1 - mx)
        best = np.minimum(best, q(dx, dy))
    inside = (x0 <= mx) & (mx <= x1) & (y0 <= my) & (my <= y1)
    return np.where(inside, 0.0, best)


def band_reach(mx, my, a, b, c, T, x0, x1, y0, y1):
    """The band form of gsevt_device.h in float64, margins set to zero."""
    det = a * c - b * b
    p, u, w = b / a, T / a, det / (a * a)
    kr = -b * np.sqrt(T / (c * det))
    ey = np.sqrt(T * a / det)
    d0, d1 = y0 - my, y1 - my
    yr = np.minimum(np.maximum(kr, d0), d1)
    yl = np.minimum(np.maximum(-kr, d0), d1)
    sr = np.sqrt(np.maximum(u - w * yr * yr, 0.0))
    sl = np.sqrt(np.maximum(u - w * yl * yl, 0.0))
    xr, xl = sr - p * yr, -sl - p * yl
    return (d1 >= -ey) & (d0 <= ey) & (x0 - mx <= xr) & (x1 - mx >= xl)


def rect_reach(mx, my, a, b, c, T, x0, x1, y0, y1):
    """The rectangle form (cull_rect) in float64, margins set to zero:
    bounding box, then the minimum of q on the sides facing the mean."""
    det = a * c - b * b
    ex, ey = np.sqrt(T * c / det), np.sqrt(T * a / det)
    box = (mx + ex >= x0) & (mx - ex <= x1) & (my + ey >= y0) & (my - ey <= y1)
    outx = (mx < x0) | (mx > x1)
    outy = (my < y0) | (my > y1)
    vx = np.where(mx < x0, x0, x1) - mx
    vy = np.clip(-b * vx / c, y0 - my, y1 - my)
    vq = a * vx * vx + 2.0 * b * vx * vy + c * vy * vy
    hy = np.where(my < y0, y0, y1) - my
    hx = np.clip(-b * hy / a, x0 - mx, x1 - mx)
    hq = a * hx * hx + 2.0 * b * hx * hy + c * hy * hy
    hit = (~outx & ~outy) | (outx & (vq <= T)) | (outy & (hq <= T))
    return box & hit


def _cases(rng, n, rho_max):
    a = np.exp(rng.uniform(np.log(0.01), np.log(3.3), n))
    c = np.exp(rng.uniform(np.log(0.01), np.log(3.3), n))
    b = rng.uniform(-rho_max, rho_max, n) * np.sqrt(a * c)
    T = rng.uniform(0.01, 11.1, n)
    mx, my = rng.uniform(-40, 40, n), rng.uniform(-40, 40, n)
    x0, y0 = rng.uniform(-60, 60, n), rng.uniform(-60, 60, n)
    x1, y1 = x0 + rng.uniform(0, 31, n), y0 + rng.uniform(0, 31, n)
    return mx, my, a, b, c, T, x0, x1, y0, y1


def test_band_and_rectangle_forms_equal_the_minimum_of_q():
    rng = np.random.default_rng(20261203)
    for rho_max, n in ((0.9, 200000), (0.9999, 200000)):
        mx, my, a, b, c, T, x0, x1, y0, y1 = _cases(rng, n, rho_max)
        m = min_q_rect(mx, my, a, b, c, x0, x1, y0, y1)
        truth = m <= T
        # pairs whose minimum lies within rounding of the threshold decide either way
        clear = np.abs(m - T) > 1e-7 * (T + 1.0)
        band = band_reach(mx, my, a, b, c, T, x0, x1, y0, y1)
        rect = rect_reach(mx, my, a, b, c, T, x0, x1, y0, y1)
        assert truth[clear].sum() > n // 20 and (~truth[clear]).sum() > n // 20
        np.testing.assert_array_equal(band[clear], truth[clear])
        np.testing.assert_array_equal(rect[clear], truth[clear])


def test_minimum_of_q_against_dense_sampling():
    """The edge minimum itself, against a 201 x 201 grid over the rectangle."""
    rng = np.random.default_rng(7)
    mx, my, a, b, c, T, x0, x1, y0, y1 = _cases(rng, 200, 0.99)
    m = min_q_rect(mx, my, a, b, c, x0, x1, y0, y1)
    t = np.linspace(0.0, 1.0, 201)
    for i in range(200):
        X = x0[i] + (x1[i] - x0[i]) * t[:, None] - mx[i]
        Y = y0[i] + (y1[i] - y0[i]) * t[None, :] - my[i]
        q = a[i] * X * X + 2.0 * b[i] * X * Y + c[i] * Y * Y
        assert m[i] <= q.min() + 1e-9 * (1.0 + q.min())          # never above a sampled value
        # and close to the grid's minimum (the grid step bounds the gap)
        step = max(x1[i] - x0[i], y1[i] - y0[i]) / 200.0
        grad = 2.0 * (abs(a[i]) + abs(b[i]) + abs(c[i])) * (np.abs(X).max() + np.abs(Y).max() + 1.0)
        assert q.min() - m[i] <= grad * step + 1e-9


def test_band_form_on_a_circle_and_an_axis_aligned_ellipse():
    """Closed forms: the circle x^2 + y^2 <= 4 meets a square above and to
    the right of its centre iff the square's nearest corner lies within 2;
    the ellipse x^2 / 4 + y^2 <= 1 spans |x| <= 2 sqrt(1 - y^2) on the row
    y."""
    z = np.zeros(1)
    one = np.ones(1)
    # circle x^2 + y^2 <= 4: the rectangle [1.5, 3] x [1.5, 3] has corner
    # distance 2.12 > 2, the rectangle [1.4, 3] x [1.4, 3] corner 1.98 < 2
    assert not band_reach(z, z, one, z, one, 4.0 * one, 1.5, 3.0, 1.5, 3.0)[0]
    assert band_reach(z, z, one, z, one, 4.0 * one, 1.4, 3.0, 1.4, 3.0)[0]
    # ellipse a = 1/4, c = 1, T = 1: x-extent 2 on the row y = 0
    a = 0.25 * one
    assert band_reach(z, z, a, z, one, one, 1.99, 5.0, 0.0, 0.0)[0]
    assert not band_reach(z, z, a, z, one, one, 2.01, 5.0, 0.0, 0.0)[0]
    # ... and 2 sqrt(1 - y^2) on the row y = 0.6: 1.6
    assert band_reach(z, z, a, z, one, one, 1.59, 5.0, 0.6, 0.6)[0]
    assert not band_reach(z, z, a, z, one, one, 1.61, 5.0, 0.6, 0.6)[0]
