"""The rest of the drop-in rasterizer API on the GPU (the legacy render of
gaussian_splatting/gaussian_renderer/__init__.py:24-153 and its
colors_precomp / cov3D_precomp branches, __init__.py:117-141):

* renders from precomputed colours and precomputed 3D covariances are
  bit-exact against the FP32 oracle (colour, depth, opacity, radii,
  n_touched);
* per-Gaussian map gradients through GaussianRasterizer with requires_grad
  inputs (means3D, scales, rotations, shs, opacities, means2D; and
  colors_precomp, cov3D_precomp on the precomputed path) are within 1e-3
  (relative to each array's largest entry) of the oracle, whose map
  gradients are pinned by FP64 central differences (test_oracle_cpu.py).
"""
import math

import numpy as np
import pytest
import torch

from helpers import orc, synthetic, make_cam, oracle_scene, se3_exp

pytestmark = pytest.mark.gpu

W, H, F = 160, 120, 140.0


def _settings(cam, dev, sh_degree):
    from diff_gaussian_rasterization import GaussianRasterizationSettings
    t = lambda a: torch.as_tensor(np.asarray(a, np.float32)).to(dev)
    return GaussianRasterizationSettings(
        image_height=cam.H, image_width=cam.W, tanfovx=cam.tanfovx, tanfovy=cam.tanfovy, bg=t(cam.bg),
        scale_modifier=1.0, viewmatrix=t(cam.view).reshape(4, 4), projmatrix=t(cam.proj).reshape(4, 4),
        projmatrix_raw=t(cam.projraw).reshape(4, 4), sh_degree=sh_degree, campos=t(cam.campos), prefiltered=False)


def _cov6(scales, rots):
    """3D covariance (upper triangle xx, xy, xz, yy, yz, zz) of scales and
    (w, x, y, z) quaternions, in float64."""
    q = rots.astype(np.float64)
    q = q / np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
                  2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
                  2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], 1).reshape(-1, 3, 3)
    M = R * scales.astype(np.float64)[:, None, :]
    S = M @ M.transpose(0, 2, 1)
    return np.stack([S[:, 0, 0], S[:, 0, 1], S[:, 0, 2], S[:, 1, 1], S[:, 1, 2], S[:, 2, 2]], 1)


def _case(seed=21, P=3000):
    raw = synthetic.small_test_scene(P=P, seed=seed, sh_degree=3, spread=0.8, log_scale_mean=-3.5)
    s = oracle_scene(raw)
    cam = make_cam(se3_exp([0.02, -0.01, 0.03, 0.01, -0.02, 0.015]), W, H, F, F, bg=[0.1, 0.2, 0.3])
    return s, cam


def _check_render(out, st):
    color, radii, depth, opac, touched = [x.detach().cpu().numpy() for x in out]
    fw = st["fwd"]
    np.testing.assert_array_equal(radii, st["pre"]["radii"])
    for name, got, ref in (("color", color, fw["color"]), ("depth", depth.reshape(-1), fw["depth"].reshape(-1)),
                           ("opacity", opac.reshape(-1), fw["opacity"].reshape(-1))):
        assert np.array_equal(got.astype(np.float32).view(np.uint32), ref.astype(np.float32).view(np.uint32)), name
    np.testing.assert_array_equal(touched, fw["n_touched"])


@pytest.mark.parametrize("precomp", ["colors", "cov3D", "both"])
def test_precomputed_inputs_render_bitexact(cuda, precomp):
    from diff_gaussian_rasterization import GaussianRasterizer
    s, cam = _case()
    rng = np.random.default_rng(5)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(cuda)
    if precomp in ("colors", "both"):
        s.colors_pre = rng.uniform(0, 1, (s.P, 3)).astype(np.float32)
    if precomp in ("cov3D", "both"):
        s.cov_pre = _cov6(s.scales, s.rots).astype(np.float32)
    st = orc.render(s, cam)
    kw = dict(means3D=T(s.means), means2D=torch.zeros(s.P, 3, device=cuda), opacities=T(s.opac).reshape(-1, 1))
    if s.colors_pre is not None:
        kw["colors_precomp"] = T(s.colors_pre)
    else:
        kw["shs"] = T(s.shs)
    if s.cov_pre is not None:
        kw["cov3D_precomp"] = T(s.cov_pre)
    else:
        kw["scales"], kw["rotations"] = T(s.scales), T(s.rots)
    out = GaussianRasterizer(_settings(cam, cuda, s.sh_degree))(**kw)
    torch.cuda.synchronize()
    print(f"[api] {precomp} precomputed: {int((st['pre']['radii'] > 0).sum())} visible, render bit-exact",
          flush=True)
    _check_render(out, st)


@pytest.mark.parametrize("precomp", [False, True])
def test_map_gradients_through_dropin(cuda, precomp):
    from diff_gaussian_rasterization import GaussianRasterizer
    s, cam = _case(seed=22)
    rng = np.random.default_rng(9)
    if precomp:
        s.colors_pre = rng.uniform(0, 1, (s.P, 3)).astype(np.float32)
        s.cov_pre = _cov6(s.scales, s.rots).astype(np.float32)
    gC = rng.standard_normal((3, H, W)).astype(np.float32)
    gD = rng.standard_normal((H, W)).astype(np.float32)
    gA = rng.standard_normal((H, W)).astype(np.float32)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(cuda).requires_grad_(True)
    ins = dict(means3D=T(s.means), means2D=torch.zeros(s.P, 3, device=cuda, requires_grad=True),
               opacities=T(s.opac.reshape(-1, 1)))
    if precomp:
        ins["colors_precomp"], ins["cov3D_precomp"] = T(s.colors_pre), T(s.cov_pre)
    else:
        ins["shs"], ins["scales"], ins["rotations"] = T(s.shs), T(s.scales), T(s.rots)
    color, radii, depth, opac, _ = GaussianRasterizer(_settings(cam, cuda, s.sh_degree))(**ins)
    g = lambda a: torch.from_numpy(a).to(cuda)
    loss = (color * g(gC)).sum() + (depth[0] * g(gD)).sum() + (opac[0] * g(gA)).sum()
    loss.backward()
    torch.cuda.synchronize()
    st = orc.render(s, cam)
    ref = orc.render_backward(s, cam, st, gC, gD, gA, map_grads=True)["map"]
    names = dict(means3D="means3D", means2D="means2D", opacities="opacities", shs="shs", scales="scales",
                 rotations="rotations", colors_precomp="colors", cov3D_precomp="cov3D")
    for k, t in ins.items():
        got = t.grad.detach().cpu().numpy().reshape(-1)
        want = ref[names[k]].reshape(-1)
        scale = np.abs(want).max()
        err = np.abs(got - want).max() / scale
        print(f"[api] d/d{k}: max|grad| {scale:.3e}, rel err {err:.2e}", flush=True)
        assert scale > 0 and err < 1e-3, k
