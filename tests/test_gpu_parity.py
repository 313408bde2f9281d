"""GPU parity of the HIP rasterizer against the FP32 oracle (bit-exact for
integer and key work; forward render within 1e-5 absolute, observed exact)."""
import numpy as np
import pytest
import torch

from helpers import orc, gpu_cam, gpu_map, splat_fields, workload_cases, make_cam

pytestmark = pytest.mark.gpu

CASES = workload_cases()
IDS = [c[0] for c in CASES]


@pytest.fixture(scope="module", params=range(len(CASES)), ids=IDS)
def case(request, cuda):
    name, scene, cam = CASES[request.param]
    print(f"\n[case] {name} P={scene.P} {cam.W}x{cam.H} tile={cam.tile_w}x{cam.tile_h}", flush=True)
    from gsevt import raster
    st = orc.render(scene, cam)
    out, state = raster.render_forward(gpu_map(scene, cuda), gpu_cam(cam, cuda), channels=3, n_touched=True)
    torch.cuda.synchronize()
    return name, scene, cam, st, out, state


def test_preprocess_bitexact(case):
    name, scene, cam, st, out, state = case
    pre = st["pre"]
    radii = out["radii"].cpu().numpy()
    np.testing.assert_array_equal(radii, pre["radii"])
    np.testing.assert_array_equal(state["tiles"].cpu().numpy(), pre["tiles"])
    vis = pre["radii"] > 0
    f = splat_fields(state["splats"].cpu().numpy(), scene.P)
    xy = f["xy"].reshape(-1, 2)[vis]
    np.testing.assert_array_equal(xy.view(np.uint32), pre["xy"].reshape(-1, 2)[vis].view(np.uint32))
    co = pre["conic_opac"].reshape(-1, 4)[vis]
    np.testing.assert_array_equal(f["conic"][vis].view(np.uint32), co[:, :2].view(np.uint32))
    np.testing.assert_array_equal(f["conic_c"][vis].view(np.uint32), co[:, 2].view(np.uint32))
    np.testing.assert_array_equal(f["depth"][vis].view(np.uint32), pre["depth"][vis].view(np.uint32))
    np.testing.assert_array_equal(f["rgb"].reshape(-1, 3)[vis].view(np.uint32),
                                  pre["rgb"].reshape(-1, 3)[vis].view(np.uint32))


def test_binning_bitexact(case):
    name, scene, cam, st, out, state = case
    b = st["bin"]
    assert state["n"] == b["n"]
    keys = state["keys"][:b["n"]].cpu().numpy().view(np.uint64)
    vals = state["vals"][:b["n"]].cpu().numpy().view(np.uint32)
    np.testing.assert_array_equal(keys, b["keys"])
    np.testing.assert_array_equal(vals, b["vals"])
    np.testing.assert_array_equal(state["ranges"].cpu().numpy().view(np.uint32), b["ranges"])


def test_blend_forward(case):
    name, scene, cam, st, out, state = case
    f = st["fwd"]
    tol = 1e-5
    col = out["color"].cpu().numpy()
    assert np.abs(col - f["color"]).max() <= tol
    assert np.abs(out["depth"].cpu().numpy()[0] - f["depth"]).max() <= tol
    assert np.abs(out["opacity"].cpu().numpy()[0] - f["opacity"]).max() <= tol
    assert np.abs(state["final_T"].cpu().numpy() - f["final_T"]).max() <= tol
    np.testing.assert_array_equal(state["n_contrib"].cpu().numpy(), f["n_contrib"])
    np.testing.assert_array_equal(out["n_touched"].cpu().numpy(), f["n_touched"])
    exact = np.array_equal(col.view(np.uint32), f["color"].view(np.uint32))
    print(f"[blend] {name}: max|d color| = {np.abs(col - f['color']).max():.3e} bit-exact={exact}")


def test_backward_vs_oracle(case):
    name, scene, cam, st, out, state = case
    from gsevt import raster
    rng = np.random.default_rng(5)
    H, W = cam.H, cam.W
    gC = rng.standard_normal((3, H, W)).astype(np.float32)
    gD = (0.1 * rng.standard_normal((H, W))).astype(np.float32)
    gA = rng.standard_normal((H, W)).astype(np.float32)
    dev = out["color"].device
    pose12, _ = raster.render_backward(gpu_map(scene, dev), gpu_cam(cam, dev), state, torch.from_numpy(gC).to(dev),
                                       torch.from_numpy(gD).to(dev), torch.from_numpy(gA).to(dev))
    torch.cuda.synchronize()
    ob = orc.render_backward(scene, cam, st, gC, gD, gA)
    g10 = state["grad10"].cpu().numpy().reshape(-1, 10)[:scene.P]
    ref = ob["grad10"]
    scale = np.abs(ref).max(axis=0) + 1e-12
    err = (np.abs(g10 - ref).max(axis=0) / scale)
    assert err.max() < 1e-3, err
    p = pose12.cpu().numpy().astype(np.float64)
    op = ob["pose"]
    ref12 = np.concatenate([op["theta"], op["rho"], op["w"], op["v"]])
    rel = np.abs(p - ref12).max() / np.abs(ref12).max()
    print(f"[backward] {name}: grad10 rel err {err.max():.2e}, pose rel err {rel:.2e}")
    assert rel < 1e-3


def test_sort_pairs_random(cuda):
    from gsevt import _lib
    from gsevt._lib import call, ptr
    rng = np.random.default_rng(1)
    for n, bits in [(1, 33), (4095, 40), (4097, 43), (100000, 43), (257, 37)]:
        print(f"[sort] n={n} bits={bits}", flush=True)
        keys = rng.integers(0, 1 << min(bits, 20), n, dtype=np.uint64) << np.uint64(max(0, bits - 20))
        keys |= rng.integers(0, 4, n, dtype=np.uint64)     # many duplicates
        vals = np.arange(n, dtype=np.uint32)
        cap = n
        k = torch.zeros(2 * cap, dtype=torch.int64, device=cuda)
        v = torch.zeros(2 * cap, dtype=torch.int32, device=cuda)
        k[:n] = torch.from_numpy(keys.view(np.int64)).to(cuda)
        v[:n] = torch.from_numpy(vals.view(np.int32)).to(cuda)
        cnt = torch.tensor([n], dtype=torch.int32, device=cuda)
        scr = torch.empty(int(_lib.load().gsevt_sort_scratch_elems(cap)), dtype=torch.int32, device=cuda)
        call("gsevt_sort_pairs", ptr(k), ptr(v), ptr(k[cap:]), ptr(v[cap:]), ptr(cnt), cap, bits, ptr(scr),
             _lib.stream_of(cuda))
        order = np.argsort(keys, kind="stable")
        np.testing.assert_array_equal(k[:n].cpu().numpy().view(np.uint64), keys[order])
        np.testing.assert_array_equal(v[:n].cpu().numpy().view(np.uint32), vals[order])


@pytest.mark.parametrize("bits,vals,drop", [(32, True, False), (11, False, False), (11, False, True),
                                             (16, True, False), (13, False, False)])
def test_seg_sort_random(cuda, bits, vals, drop):
    """gsevt_seg_sort_u32_bits against numpy's stable argsort per segment:
    segments with full, partial and empty 4096-element blocks, 8-bit digits
    (32, 16 key bits), 6-bit digits (11) and a mix (13: 6 then 7 bits), with
    values or keys only, and the dropping first pass (0xFFFFFFFF keys removed,
    kept counts returned)."""
    from gsevt import fused
    from gsevt._lib import call, ptr, stream_of
    fused._lib_ext()
    rng = np.random.default_rng(bits * 7 + vals + 2 * drop)
    R, stride = 4, 9000
    counts = np.array([9000, 4097, 0, 2049], np.int64)
    keys = rng.integers(0, 1 << bits, (R, stride), dtype=np.uint64).astype(np.uint32)
    keys[:, ::3] &= np.uint32(0xF0F)          # many equal keys
    if drop:
        keys[rng.random((R, stride)) < 0.4] = np.uint32(0xFFFFFFFF)
    v = np.tile(np.arange(stride, dtype=np.uint32), (R, 1))
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a).view(np.int32)).to(cuda)
    k0, k1 = t(keys), torch.zeros(R * stride, dtype=torch.int32, device=cuda)
    v0 = t(v) if vals else None
    v1 = torch.zeros(R * stride, dtype=torch.int32, device=cuda) if vals else None
    cnt = torch.from_numpy(counts.astype(np.int32)).to(cuda)
    kept = torch.zeros(R, dtype=torch.int32, device=cuda)
    L = fused._lib_ext()
    scr = torch.empty(int(L.gsevt_seg_sort_scratch_elems(R, stride)), dtype=torch.int32, device=cuda)
    print(f"[seg_sort] bits={bits} vals={vals} drop={drop}", flush=True)
    call("gsevt_seg_sort_u32_bits", R, stride, ptr(cnt), 0, ptr(k0), ptr(v0), ptr(k1), ptr(v1), 0, bits, ptr(scr),
         ptr(kept) if drop else None, 0, stream_of(cuda))
    got_k = k0.cpu().numpy().view(np.uint32).reshape(R, stride)
    got_v = v0.cpu().numpy().view(np.uint32).reshape(R, stride) if vals else None
    got_kept = kept.cpu().numpy()
    for r in range(R):
        kk, vv = keys[r, :counts[r]], v[r, :counts[r]]
        if drop:
            sel = kk != np.uint32(0xFFFFFFFF)
            kk, vv = kk[sel], vv[sel]
            assert got_kept[r] == kk.shape[0]
        order = np.argsort(kk & np.uint32((1 << bits) - 1), kind="stable")
        np.testing.assert_array_equal(got_k[r, :kk.shape[0]], kk[order])
        if vals:
            np.testing.assert_array_equal(got_v[r, :kk.shape[0]], vv[order])


@pytest.mark.parametrize("offset", [0, 1])
def test_scan_random(cuda, offset):
    """Inclusive scan against numpy: partial and whole 2048-element tiles,
    whole tiles ending the input (the total comes from the vector path), and
    arrays that start off a 16-byte boundary (offset 1: the scalar path)."""
    from gsevt import _lib
    from gsevt._lib import call, ptr
    rng = np.random.default_rng(2)
    for n in [1, 2047, 2048, 2049, 4096, 1000000, 5000001]:
        a = rng.integers(0, 20, n).astype(np.uint32)
        xb = torch.zeros(n + offset, dtype=torch.int32, device=cuda)
        yb = torch.zeros(n + offset, dtype=torch.int32, device=cuda)
        x, y = xb[offset:], yb[offset:]
        x.copy_(torch.from_numpy(a.view(np.int32)))
        tot = torch.zeros(1, dtype=torch.int32, device=cuda)
        scr = torch.empty(int(_lib.load().gsevt_scan_scratch_elems(n)), dtype=torch.int32, device=cuda)
        call("gsevt_scan_inclusive_u32", ptr(x), ptr(y), n, ptr(scr), ptr(tot), _lib.stream_of(cuda))
        ref = np.cumsum(a.astype(np.uint64)).astype(np.uint32)
        np.testing.assert_array_equal(y.cpu().numpy().view(np.uint32), ref)
        assert int(tot.item()) == int(ref[-1])


def test_dropin_transposed_settings_matrices(cuda):
    """Regression: settings carrying transposed (non-contiguous) matrices, as
    Camera properties produce, must give the oracle's render and gradient."""
    import __graft_entry__
    __graft_entry__.smoke()


def test_mark_visible_bitexact(cuda):
    """GaussianRasterizer.markVisible through gsevt_mark_visible against the
    oracle's near-plane mask: random points, the exact 0.2 boundary, NaN and
    infinities, a ragged P and the empty case."""
    from helpers import se3_exp
    from diff_gaussian_rasterization import GaussianRasterizationSettings, GaussianRasterizer
    rng = np.random.default_rng(11)
    cam = make_cam(se3_exp([0.3, -0.2, 0.5, 0.4, -0.7, 0.2]), 64, 48, 50.0, 50.0)
    pts = rng.uniform(-3, 3, size=(100003, 3)).astype(np.float32)
    z02 = np.float32(0.2)
    special = np.array([z02, np.nextafter(z02, np.float32(1)), np.nextafter(z02, np.float32(0)), np.nan, np.inf,
                        -np.inf, -1.0], np.float32)
    eye = np.eye(4, dtype=np.float32).reshape(-1)
    sp = np.stack([np.zeros_like(special), np.zeros_like(special), special], 1)
    t = lambda a: torch.as_tensor(np.asarray(a, np.float32)).to(cuda)
    for view, p in ((cam.view, pts), (eye, sp), (cam.view, pts[:0])):
        print(f"[mark_visible] P={p.shape[0]}", flush=True)
        st = GaussianRasterizationSettings(
            image_height=cam.H, image_width=cam.W, tanfovx=cam.tanfovx, tanfovy=cam.tanfovy, bg=t(cam.bg),
            scale_modifier=1.0, viewmatrix=t(view).reshape(4, 4), projmatrix=t(cam.proj).reshape(4, 4),
            projmatrix_raw=t(cam.projraw).reshape(4, 4), sh_degree=0, campos=t(cam.campos), prefiltered=False, debug=False)
        got = GaussianRasterizer(raster_settings=st).markVisible(t(p))
        torch.cuda.synchronize()
        assert got.dtype == torch.bool and got.shape == (p.shape[0],)
        np.testing.assert_array_equal(got.cpu().numpy(), orc.mark_visible(p, view))
