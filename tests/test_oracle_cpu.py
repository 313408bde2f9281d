"""CPU pinning of the oracle: the FP64 brute-force renderer, closed forms,
binning invariants and FP64 central finite differences.  None of these
retypes the oracle's formulas: each check is an independent computation that
a dropped term, a wrong sign or a transposed operand would fail."""
import math

import numpy as np
import pytest

from helpers import orc, synthetic, oracle_scene, make_cam, se3_exp, scene_astype, cam_astype
from oracle import brute


# --------------------------------------------------------------------------
# deterministic exp
# --------------------------------------------------------------------------
def test_det_expf_within_2ulp_of_libm():
    xs = np.concatenate([np.linspace(-87, 0, 20001), -np.logspace(-8, 1.9, 2000), [0.0, -1e-30]])
    worst = 0.0
    for x in xs.astype(np.float32):
        got = np.float32(orc.det_expf(float(x)))
        ref = np.float32(math.exp(float(x)))
        ulp = np.spacing(ref)
        worst = max(worst, abs(float(got) - float(ref)) / float(ulp))
    assert worst <= 2.0, worst
    assert orc.det_expf(0.0) == 1.0


# --------------------------------------------------------------------------
# oracle FP64 == independent brute force FP64
# --------------------------------------------------------------------------
BRUTE_CASES = [
    dict(seed=1, deg=3, W=48, H=40, f=40.0, tile=(16, 16), pose=[0, 0, 0, 0, 0, 0], bg=(0, 0, 0)),
    dict(seed=2, deg=2, W=40, H=36, f=35.0, tile=(8, 32), pose=[0.05, -0.02, 0.1, 0.02, -0.03, 0.01],
         bg=(0.2, 0.5, 0.1)),
    dict(seed=3, deg=1, W=52, H=30, f=45.0, tile=(32, 8), pose=[-0.03, 0.04, 0.0, -0.05, 0.02, 0.04], bg=(0, 0, 0)),
    dict(seed=4, deg=0, W=33, H=29, f=30.0, tile=(16, 16), pose=[0.0, 0.0, 0.3, 0.1, 0.0, 0.0], bg=(1, 1, 1)),
]


@pytest.mark.parametrize("c", BRUTE_CASES, ids=lambda c: f"seed{c['seed']}_deg{c['deg']}_{c['tile']}")
def test_oracle_fp64_matches_brute_force(c):
    raw = synthetic.small_test_scene(P=120, seed=c["seed"], sh_degree=c["deg"], W=c["W"], H=c["H"])
    s = scene_astype(oracle_scene(raw), np.float64)
    cam = make_cam(se3_exp(c["pose"]), c["W"], c["H"], c["f"], c["f"], dtype=np.float64, tile=c["tile"],
                   bg=np.array(c["bg"], float))
    st = orc.render(s, cam, dtype=np.float64)
    b = brute.render(s, cam, bg=cam.bg)
    np.testing.assert_array_equal(st["pre"]["tiles"], b["proj"]["tiles"])
    for k in ("color", "depth", "opacity", "final_T"):
        assert np.abs(st["fwd"][k] - b[k]).max() < 1e-10, k


def test_oracle_precomputed_colour_and_covariance_paths():
    raw = synthetic.small_test_scene(P=80, seed=9, sh_degree=1)
    s = scene_astype(oracle_scene(raw), np.float64)
    rng = np.random.default_rng(0)
    R = brute.quat_to_rot(s.rots)
    S = s.scales
    M = R * S[:, None, :]
    Sig = M @ np.transpose(M, (0, 2, 1))
    cov = np.stack([Sig[:, 0, 0], Sig[:, 0, 1], Sig[:, 0, 2], Sig[:, 1, 1], Sig[:, 1, 2], Sig[:, 2, 2]], -1)
    cols = rng.uniform(0, 1, (s.P, 3))
    s2 = orc.Scene(means=s.means, scales=None, rots=None, opac=s.opac, shs=None, sh_degree=1, colors_pre=cols,
                   cov_pre=cov)
    cam = make_cam(np.eye(4), 48, 40, 40.0, 40.0, dtype=np.float64)
    st = orc.render(s2, cam, dtype=np.float64)
    b = brute.render(s2, cam)
    assert np.abs(st["fwd"]["color"] - b["color"]).max() < 1e-10
    # same geometry as the scale/rotation path
    st1 = orc.render(s, cam, dtype=np.float64)
    np.testing.assert_array_equal(st["pre"]["radii"], st1["pre"]["radii"])


# --------------------------------------------------------------------------
# closed forms
# --------------------------------------------------------------------------
def _single(mean, scale, opac, sh, deg=0, W=32, H=32, f=30.0, rot=(1, 0, 0, 0)):
    K = (deg + 1) ** 2
    shs = np.zeros((1, K, 3))
    shs[0, :sh.shape[0]] = sh
    return orc.Scene(means=np.array([mean], float), scales=np.array([scale], float), rots=np.array([rot], float),
                     opac=np.array([opac], float), shs=shs, sh_degree=deg), make_cam(np.eye(4), W, H, f, f,
                                                                                       dtype=np.float64)


def test_sh_degree0_colour_closed_form():
    dc = np.array([[0.7, -2.0, 3.0]])
    s, cam = _single([0.01, 0.02, 2.0], [0.05, 0.05, 0.05], 0.8, dc)
    pre = orc.preprocess(s, cam, np.float64)
    expect = np.maximum(0.28209479177387814 * dc[0] + 0.5, 0.0)
    np.testing.assert_allclose(pre["rgb"], expect, rtol=0, atol=1e-15)
    assert list(pre["clamped"]) == [0, 1, 0]


def test_axis_aligned_conic_and_radius_closed_form():
    # Gaussian on the optical axis, axis-aligned: cov2D = diag((f sx/z)^2 + .3, (f sy/z)^2 + .3)
    sx, sy, sz, z, f = 0.03, 0.05, 0.02, 2.5, 30.0
    s, cam = _single([0.0, 0.0, z], [sx, sy, sz], 0.5, np.zeros((1, 3)), f=f, W=31, H=31)
    pre = orc.preprocess(s, cam, np.float64)
    fx = cam.W / (2 * cam.tanfovx)
    fy = cam.H / (2 * cam.tanfovy)
    a = (fx * sx / z) ** 2 + 0.3
    c = (fy * sy / z) ** 2 + 0.3
    co = pre["conic_opac"]
    np.testing.assert_allclose(co[:3], [1 / a, 0.0, 1 / c], rtol=1e-12, atol=1e-15)
    lam = max(a, c)
    assert pre["radii"][0] == math.ceil(3 * math.sqrt(lam))
    # the centre projects to the principal point ((W-1)/2 in pixel units)
    np.testing.assert_allclose(pre["xy"], [(cam.W - 1) / 2, (cam.H - 1) / 2], atol=1e-9)


def test_single_gaussian_on_pixel_centre():
    # mean projecting exactly to pixel (15, 15): alpha = min(.99, o), final T = 1 - alpha
    for o in (0.3, 0.995):
        s, cam = _single([0.0, 0.0, 2.0], [0.02, 0.02, 0.02], o, np.array([[1.0, 0.2, -0.4]]), W=31, H=31)
        st = orc.render(s, cam, dtype=np.float64)
        a = min(0.99, o)
        rgb = np.maximum(0.28209479177387814 * np.array([1.0, 0.2, -0.4]) + 0.5, 0)
        np.testing.assert_allclose(st["fwd"]["color"][:, 15, 15], a * rgb, atol=1e-12)
        np.testing.assert_allclose(st["fwd"]["final_T"][15, 15], 1 - a, atol=1e-12)
        np.testing.assert_allclose(st["fwd"]["opacity"][15, 15], a, atol=1e-12)
        np.testing.assert_allclose(st["fwd"]["depth"][15, 15], a * 2.0, atol=1e-12)


def test_sh_basis_is_orthonormal_on_the_sphere():
    """Pins the SH constants and polynomials of the brute-force evaluator
    (and through oracle == brute, the oracle): integrate b_i b_j over S^2."""
    n = 200
    th = (np.arange(n) + 0.5) / n * np.pi
    ph = (np.arange(2 * n) + 0.5) / (2 * n) * 2 * np.pi
    T, Pp = np.meshgrid(th, ph, indexing="ij")
    dirs = np.stack([np.sin(T) * np.cos(Pp), np.sin(T) * np.sin(Pp), np.cos(T)], -1).reshape(-1, 3)
    w = (np.sin(T) * (np.pi / n) * (np.pi / n)).reshape(-1)
    B = np.zeros((dirs.shape[0], 16))
    for k in range(16):
        sh = np.zeros((dirs.shape[0], 16, 3))
        sh[:, k, 0] = 1.0
        B[:, k] = brute.eval_sh(3, sh, dirs)[:, 0]
    G = (B * w[:, None]).T @ B
    np.testing.assert_allclose(G, np.eye(16), atol=2e-4)


# --------------------------------------------------------------------------
# binning invariants (FP32 oracle)
# --------------------------------------------------------------------------
def test_binning_invariants():
    raw = synthetic.cube_scene(P=3000)
    s = oracle_scene(raw)
    cam = make_cam(np.eye(4), 173, 130, 150.0, 150.0)
    st = orc.render(s, cam)
    pre, b = st["pre"], st["bin"]
    keys = b["keys"]
    assert np.all(keys[1:] >= keys[:-1])
    assert b["n"] == int(pre["tiles"].sum())
    gx, gy = orc.grid(cam)
    rng_ = b["ranges"].reshape(-1, 2)
    nonempty = rng_[rng_[:, 1] > rng_[:, 0]]
    assert nonempty[0, 0] == 0 and nonempty[-1, 1] == b["n"]
    assert np.all(nonempty[1:, 0] == nonempty[:-1, 1])
    tiles = (keys >> np.uint64(32)).astype(np.int64)
    rect = pre["rect"].reshape(-1, 4)[b["vals"]]
    tx, ty = tiles % gx, tiles // gx
    assert np.all((rect[:, 0] <= tx) & (tx < rect[:, 2]) & (rect[:, 1] <= ty) & (ty < rect[:, 3]))
    # equal keys keep Gaussian index order (stable)
    same = keys[1:] == keys[:-1]
    assert np.all(b["vals"][1:][same] > b["vals"][:-1][same])
    appears = np.zeros(s.P, bool)
    appears[b["vals"]] = True
    np.testing.assert_array_equal(appears, pre["radii"] > 0)


def test_render_is_permutation_invariant():
    raw = synthetic.small_test_scene(P=300, seed=21, W=64, H=48)
    s = scene_astype(oracle_scene(raw), np.float64)
    cam = make_cam(np.eye(4), 64, 48, 50.0, 50.0, dtype=np.float64)
    perm = np.random.default_rng(1).permutation(s.P)
    sp = orc.Scene(means=s.means[perm], scales=s.scales[perm], rots=s.rots[perm], opac=s.opac[perm],
                   shs=s.shs[perm], sh_degree=s.sh_degree)
    a = orc.render(s, cam, dtype=np.float64)["fwd"]["color"]
    b = orc.render(sp, cam, dtype=np.float64)["fwd"]["color"]
    assert np.abs(a - b).max() < 1e-12


# --------------------------------------------------------------------------
# FP64 central finite differences
# --------------------------------------------------------------------------
def _fd_setup(seed):
    raw = synthetic.small_test_scene(P=90, seed=seed, W=44, H=36)
    s = scene_astype(oracle_scene(raw), np.float64)
    rng = np.random.default_rng(seed)
    W, H, f = 44, 36, 38.0
    gC = rng.standard_normal((3, H, W))
    gD = 0.1 * rng.standard_normal((H, W))
    gA = rng.standard_normal((H, W))
    return s, (W, H, f), gC, gD, gA, rng


def _loss(s, cam, gC, gD, gA):
    F = orc.render(s, cam, dtype=np.float64, n_touched=False)
    return (gC * F["fwd"]["color"]).sum() + (gD * F["fwd"]["depth"]).sum() + (gA * F["fwd"]["opacity"]).sum(), F


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_pose_and_velocity_gradient_fd(seed):
    s, (W, H, f), gC, gD, gA, rng = _fd_setup(seed)
    Tc = se3_exp(rng.normal(0, [0.03, 0.03, 0.03, 0.03, 0.03, 0.03]))
    lin, ang, dt = rng.normal(0, 0.5, 3), rng.normal(0, 0.5, 3), (-0.005 if seed % 2 else 0.005)

    def L(delta=np.zeros(6), dv=np.zeros(6)):
        E = se3_exp(dt * np.concatenate([lin + dv[:3], ang + dv[3:]]))
        cam = make_cam(E @ se3_exp(delta) @ Tc, W, H, f, f, dtype=np.float64)
        return _loss(s, cam, gC, gD, gA)[0]

    E = se3_exp(dt * np.concatenate([lin, ang]))
    cam = make_cam(E @ Tc, W, H, f, f, dtype=np.float64, vel=E, delta_time=dt, lin=lin, ang=ang)
    _, st = _loss(s, cam, gC, gD, gA)
    g = orc.render_backward(s, cam, st, gC, gD, gA, dtype=np.float64)["pose"]
    an = np.concatenate([g["rho"], g["theta"], g["v"], g["w"]])
    h = 1e-6
    fd = []
    for which in ("delta", "vel"):
        for k in range(6):
            e = np.zeros(6)
            e[k] = h
            if which == "delta":
                fd.append((L(delta=e) - L(delta=-e)) / (2 * h))
            else:
                fd.append((L(dv=e) - L(dv=-e)) / (2 * h))
    fd = np.array(fd)
    err = np.abs(an - fd) / (np.abs(fd) + 1e-3 * np.abs(fd).max())
    assert err.max() < 1e-3, (an, fd)


def test_map_gradients_fd():
    s, (W, H, f), gC, gD, gA, rng = _fd_setup(5)
    cam = make_cam(se3_exp([0.01, 0.02, -0.03, 0.02, 0.01, -0.02]), W, H, f, f, dtype=np.float64)
    _, st = _loss(s, cam, gC, gD, gA)
    g = orc.render_backward(s, cam, st, gC, gD, gA, dtype=np.float64, map_grads=True)["map"]
    vis = np.nonzero(st["pre"]["radii"] > 0)[0]
    idx = rng.choice(vis, 6, replace=False)
    h = 1e-6
    checks = 0
    for i in idx:
        for field, arr, gm in (("means", s.means, g["means3D"].reshape(-1, 3)),
                               ("scales", s.scales, g["scales"].reshape(-1, 3)),
                               ("rots", s.rots, g["rotations"].reshape(-1, 4)),
                               ("opac", s.opac[:, None], g["opacities"].reshape(-1, 1)),
                               ("shs", s.shs.reshape(s.P, -1), g["shs"].reshape(s.P, -1))):
            for k in range(min(arr.shape[1], 5)):
                def L(sign):
                    a2 = arr.copy()
                    a2[i, k] += sign * h
                    kw = dict(means=s.means, scales=s.scales, rots=s.rots, opac=s.opac, shs=s.shs)
                    key = {"means": "means", "scales": "scales", "rots": "rots", "opac": "opac", "shs": "shs"}[field]
                    kw[key] = a2.reshape(kw[key].shape)
                    return _loss(orc.Scene(sh_degree=s.sh_degree, **kw), cam, gC, gD, gA)[0]
                fd = (L(1) - L(-1)) / (2 * h)
                an = gm[i, k]
                assert abs(an - fd) <= 1e-4 * max(1.0, abs(fd)), (field, i, k, an, fd)
                checks += 1
    assert checks > 50


def test_precomputed_input_gradients_fd():
    """Gradients of precomputed colours and precomputed 3D covariances
    (the colors_precomp / cov3D_precomp branches) against FP64 central
    differences.  The screen-space means2D gradient has no forward input to
    difference (the render does not read means2D); it is checked to be zero
    for culled Gaussians and non-zero for visible ones."""
    s, (W, H, f), gC, gD, gA, rng = _fd_setup(6)
    s.colors_pre = rng.uniform(0, 1, (s.P, 3))
    q = s.rots / np.linalg.norm(s.rots, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
                  2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
                  2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], 1).reshape(-1, 3, 3)
    M = R * s.scales[:, None, :]
    S = M @ M.transpose(0, 2, 1)
    s.cov_pre = np.stack([S[:, 0, 0], S[:, 0, 1], S[:, 0, 2], S[:, 1, 1], S[:, 1, 2], S[:, 2, 2]], 1)
    cam = make_cam(se3_exp([0.01, -0.02, 0.02, 0.015, 0.01, -0.01]), W, H, f, f, dtype=np.float64)
    _, st = _loss(s, cam, gC, gD, gA)
    g = orc.render_backward(s, cam, st, gC, gD, gA, dtype=np.float64, map_grads=True)["map"]
    vis = np.nonzero(st["pre"]["radii"] > 0)[0]
    assert np.all(g["means2D"].reshape(-1, 3)[st["pre"]["radii"] == 0] == 0)
    assert np.abs(g["means2D"].reshape(-1, 3)[vis, :2]).max() > 0
    h = 1e-6
    checks = 0
    for i in rng.choice(vis, 6, replace=False):
        for field, gm, n in (("colors_pre", g["colors"].reshape(-1, 3), 3), ("cov_pre", g["cov3D"].reshape(-1, 6), 6)):
            for k in range(n):
                def L(sign):
                    a2 = getattr(s, field).copy()
                    a2[i, k] += sign * h
                    kw = dict(means=s.means, scales=s.scales, rots=s.rots, opac=s.opac, shs=s.shs,
                              colors_pre=s.colors_pre, cov_pre=s.cov_pre)
                    kw[field] = a2
                    return _loss(orc.Scene(sh_degree=s.sh_degree, **kw), cam, gC, gD, gA)[0]
                fd = (L(1) - L(-1)) / (2 * h)
                an = gm[i, k]
                assert abs(an - fd) <= 1e-4 * max(1.0, abs(fd)), (field, i, k, an, fd)
                checks += 1
    assert checks == 54


# --------------------------------------------------------------------------
# FP32 oracle vs FP64 oracle (precision of the FP32 specification)
# --------------------------------------------------------------------------
def test_fp32_oracle_close_to_fp64():
    raw = synthetic.small_test_scene(P=400, seed=8, W=64, H=48)
    s32 = oracle_scene(raw)
    s64 = scene_astype(s32, np.float64)
    cam32 = make_cam(np.eye(4), 64, 48, 50.0, 50.0)
    cam64 = cam_astype(cam32, np.float64)
    a = orc.render(s32, cam32)["fwd"]["color"]
    b = orc.render(s64, cam64, dtype=np.float64)["fwd"]["color"]
    assert np.abs(a - b).max() < 1e-4


# --------------------------------------------------------------------------
# mark_visible: near-plane mask
# --------------------------------------------------------------------------
def test_mark_visible_against_fp64_depth_and_preprocess():
    rng = np.random.default_rng(5)
    T = se3_exp([0.3, -0.2, 0.5, 0.4, -0.7, 0.2])
    cam = make_cam(T, 64, 48, 50.0, 50.0)
    pts = rng.uniform(-3, 3, size=(4000, 3)).astype(np.float32)
    got = orc.mark_visible(pts, cam.view)
    # independent FP64 view-space depth from the pose matrix (row 2 of T)
    z = pts.astype(np.float64) @ T[2, :3] + T[2, 3]
    far = np.abs(z - 0.2) > 1e-4
    assert far.sum() > 3900 and 0.2 < got.mean() < 0.8
    np.testing.assert_array_equal(got[far], z[far] > 0.2)
    # the transposed matrix (rows and columns swapped) gives another mask
    assert (orc.mark_visible(pts, T.astype(np.float32).reshape(-1)) != got).sum() > 100
    # every Gaussian the preprocess keeps is visible
    raw = synthetic.small_test_scene(P=300, seed=3, sh_degree=1)
    s = oracle_scene(raw)
    pre = orc.preprocess(s, cam)
    vis = orc.mark_visible(s.means, cam.view)
    assert (pre["radii"] > 0).any()
    assert np.all(vis[pre["radii"] > 0])


def test_mark_visible_boundary_and_nan():
    view = np.eye(4, dtype=np.float32).reshape(-1)
    z02 = np.float32(0.2)
    zs = np.array([z02, np.nextafter(z02, np.float32(1)), np.nextafter(z02, np.float32(0)), -5.0, np.nan, np.inf,
                   -np.inf], np.float32)
    pts = np.stack([np.zeros_like(zs), np.zeros_like(zs), zs], 1)
    np.testing.assert_array_equal(orc.mark_visible(pts, view), [False, True, False, False, True, True, False])
    # translation along z moves the plane: z_view = z + 1
    view_t = view.copy()
    view_t[14] = 1.0
    np.testing.assert_array_equal(orc.mark_visible(np.array([[0, 0, -0.9], [0, 0, -0.7]], np.float32), view_t),
                                  [False, True])
    assert orc.mark_visible(np.zeros((0, 3), np.float32), view).shape == (0,)


def test_tile_cull_bounds_against_sampled_minimum():
    """qmin of every fragment against a dense sampling of its tile rectangle
    (never below qmin, and within the sampling bound of it), and the
    pixel-centre minimum never below qmin."""
    raw = synthetic.small_test_scene(P=150, seed=21, sh_degree=0, W=64, H=48)
    s = oracle_scene(raw)
    cam = make_cam(se3_exp([0.02, -0.01, 0.05, 0.03, 0.02, -0.01]), 64, 48, 50.0, 50.0)
    pre = orc.preprocess(s, cam)
    b = orc.bin(pre, cam)
    qmin, L, mag = orc.tile_cull_bounds(pre, b, cam)
    assert np.all(mag >= qmin * (1 - 1e-12))
    assert b["n"] > 200 and np.all(np.isfinite(qmin)) and np.all(qmin >= 0)
    gx, _ = orc.grid(cam)
    tile = (b["keys"] >> np.uint64(32)).astype(np.int64)
    g = b["vals"].astype(np.int64)
    xy = pre["xy"].reshape(-1, 2).astype(np.float64)
    co = pre["conic_opac"].reshape(-1, 4).astype(np.float64)
    u = np.linspace(0.0, 1.0, 61)
    pix = np.arange(16, dtype=np.float64)
    for k in range(0, b["n"], 7):
        x0, y0 = (tile[k] % gx) * 16.0, (tile[k] // gx) * 16.0
        a_, b_, c_ = co[g[k], :3]
        for xs, ys in ((x0 + 15 * u, y0 + 15 * u), (x0 + pix, y0 + pix)):
            dx = xs[None, :] - xy[g[k], 0]
            dy = ys[:, None] - xy[g[k], 1]
            qs = a_ * dx * dx + 2 * b_ * dx * dy + c_ * dy * dy
            assert qs.min() >= qmin[k] * (1 - 1e-12) - 1e-12
        # sampling step 0.25 px: the sampled minimum is within the change of q
        # over a quarter pixel of the true minimiser
        dxs = x0 + 15 * u - xy[g[k], 0]
        dys = y0 + 15 * u - xy[g[k], 1]
        qs = a_ * dxs[None, :] ** 2 + 2 * b_ * dxs[None, :] * dys[:, None] + c_ * dys[:, None] ** 2
        grad = 2 * (abs(a_) + abs(b_) + abs(c_)) * (np.sqrt(qs.min() / min(a_, c_) + 1e-12) + 0.25)
        assert qs.min() - qmin[k] <= 0.25 * grad + 1e-9
    rng_mask = (pre["conic_opac"].reshape(-1, 4)[:, 3][g] > 0)
    np.testing.assert_allclose(L[rng_mask], np.log(255.0 * co[g[rng_mask], 3]))


def _culled(b, pre, cam, keep):
    """Kept list and packed-style ranges of a binning filtered by `keep`."""
    gx, gy = orc.grid(cam)
    tile = (b["keys"] >> np.uint64(32)).astype(np.int64)[keep]
    gids = b["vals"][keep]
    ranges = np.zeros((gx * gy, 2), np.int64)
    if gids.shape[0]:
        t, first = np.unique(tile, return_index=True)
        ranges[t, 0] = first
        ranges[t, 1] = np.append(first[1:], gids.shape[0])
    return gids, ranges


def test_culled_binning_checker_accepts_margins_and_rejects_mistakes():
    from helpers import check_culled_binning
    raw = synthetic.small_test_scene(P=200, seed=4, sh_degree=0, W=64, H=48)
    s = oracle_scene(raw)
    cam = make_cam(se3_exp([0.01, 0.02, -0.03, 0.02, -0.01, 0.03]), 64, 48, 50.0, 50.0)
    pre = orc.preprocess(s, cam)
    b = orc.bin(pre, cam)
    qmin, L, _ = orc.tile_cull_bounds(pre, b, cam)
    keep = qmin <= 2.0 * (np.maximum(L, 0) * 1.0001 + 1e-4) + 1e-3
    assert 0.2 * b["n"] < keep.sum() < b["n"]
    kept, slots = check_culled_binning(*_culled(b, pre, cam, keep), b, pre, cam)
    assert kept == keep.sum() and slots == b["n"]
    # keeping every slot fails the tightness bound; dropping a reachable
    # fragment or reordering a tile fails too
    with pytest.raises(AssertionError):
        check_culled_binning(*_culled(b, pre, cam, np.ones_like(keep)), b, pre, cam)
    bad = keep.copy()
    bad[np.nonzero(qmin < 1.0)[0][0]] = False
    with pytest.raises(AssertionError):
        check_culled_binning(*_culled(b, pre, cam, bad), b, pre, cam)
    gids, ranges = _culled(b, pre, cam, keep)
    t = np.nonzero(ranges[:, 1] - ranges[:, 0] >= 2)[0][0]
    gids = gids.copy()
    gids[ranges[t, 0]], gids[ranges[t, 0] + 1] = gids[ranges[t, 0] + 1], gids[ranges[t, 0]]
    with pytest.raises(AssertionError):
        check_culled_binning(gids, ranges, b, pre, cam)


@pytest.mark.parametrize("bad", [("np.clip(-b * dx / c", "np.clip(b * dx / c"),
                                 ("np.clip(-b * dy / a", "np.clip(-b * dy / c"),
                                 ("+ 2.0 * b * dx * dy + c", "+ b * dx * dy + c")],
                         ids=["stationary_sign", "swapped_conic_term", "dropped_factor_2"])
def test_tile_cull_bounds_pins_planted_mistakes(monkeypatch, bad):
    """A plausible mistake planted in tile_cull_bounds fails its pinning test."""
    import inspect
    import test_oracle_cpu as this
    src = inspect.getsource(orc.tile_cull_bounds)
    assert bad[0] in src
    ns = dict(vars(orc))
    exec(src.replace(bad[0], bad[1]), ns)
    monkeypatch.setattr(orc, "tile_cull_bounds", ns["tile_cull_bounds"])
    with pytest.raises(AssertionError):
        this.test_tile_cull_bounds_against_sampled_minimum()
