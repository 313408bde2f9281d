// CPU oracle for the Gaussian-splatting rasterizer on GS-EVT's tracking path.
//
// TEST INFRASTRUCTURE ONLY. Nothing in the product links, loads or calls this
// file; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// use it, always as the checker.
//
// The reference rasterizer (submodule diff-gaussian-rasterization-vwRT,
// .gitmodules:1-3) is absent from the reference tree, so this file restates
// the published 3D Gaussian Splatting rasterizer (Kerbl et al. 2023) as it is
// called from gaussian_splatting/gaussian_renderer/__init__.py:313-384:
//   preprocess  : cull (view z <= 0.2), EWA covariance, conic, radius, tile
//                 rect, SH -> RGB (+0.5, clamp >= 0), depth
//   binning     : key = tile << 32 | float_bits(depth), stable sort, ranges
//   blend fwd   : front-to-back alpha blend (alpha <= 0.99, skip < 1/255,
//                 stop when T < 1e-4) of colour, depth and opacity
//   blend bwd   : reverse replay with T recovered by division
//   preprocess bwd + pose/velocity adjoint (exact, see pose_finalize)
//
// Every floating-point expression is written in a fixed evaluation order and
// the file is compiled with -ffp-contract=off, so the FP32 instantiation is the
// bit-level specification the HIP kernels are checked against.  The FP64
// instantiation is the finite-difference engine for gradient tests.
//
// Parity status: the algorithm is pinned by tests/ against an independent
// FP64 numpy brute-force renderer (no tiles), closed forms (SH degree 0,
// axis-aligned EWA conic, single-Gaussian alpha), invariants (sorted keys,
// partitioning ranges, sum of tiles_touched) and FP64 central finite
// differences for every gradient.  Parity against the reference's own
// rasterizer is unpinned: its source is not in the reference tree.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include <omp.h>

namespace {

// ---------------------------------------------------------------------------
// Deterministic exp.  exp(x) = 2^k * p(r), r = x - k ln2, p = degree-7 Taylor
// polynomial evaluated with fmaf.  Only IEEE basic operations and fma are
// used, so host and device produce identical bits.  Max error < 2 ulp on
// [-87, 0] (checked in tests against libm).
// ---------------------------------------------------------------------------
inline float det_exp(float x) {
  if (x < -87.0f) x = -87.0f;
  const float k = rintf(x * 1.44269504f);
  float r = fmaf(k, -0.693145751953125f, x);
  r = fmaf(k, -1.428606765330187e-06f, r);
  float p = 1.98412698e-4f;
  p = fmaf(p, r, 1.38888889e-3f);
  p = fmaf(p, r, 8.33333333e-3f);
  p = fmaf(p, r, 4.16666667e-2f);
  p = fmaf(p, r, 1.66666667e-1f);
  p = fmaf(p, r, 0.5f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  return ldexpf(p, (int)k);
}
inline double det_exp(double x) { return std::exp(x); }

inline float det_sqrt(float x) { return sqrtf(x); }
inline double det_sqrt(double x) { return std::sqrt(x); }
inline float det_ceil(float x) { return ceilf(x); }
inline double det_ceil(double x) { return std::ceil(x); }
template <typename T> inline T tmin(T a, T b) { return a < b ? a : b; }
template <typename T> inline T tmax(T a, T b) { return a > b ? a : b; }
// fminf/fmaxf semantics (NaN loses) for the rect clamp
inline float fclamp(float v, float lo, float hi) { return fminf(hi, fmaxf(lo, v)); }
inline double fclamp(double v, double lo, double hi) { return std::fmin(hi, std::fmax(lo, v)); }

// Real spherical-harmonic constants (degree 0..3), sign convention of
// gaussian_splatting/utils/sh_utils.py:24-52.
const double SH_C0 = 0.28209479177387814;
const double SH_C1 = 0.4886025119029199;
const double SH_C2[5] = {1.0925484305920792, -1.0925484305920792, 0.31539156525252005,
                         -1.0925484305920792, 0.5462742152960396};
const double SH_C3[7] = {-0.5900435899266435, 2.890611442640554, -0.4570457994644658,
                         0.3731763325901154, -0.4570457994644658, 1.445305721320277,
                         -0.5900435899266435};

// SH basis values at a unit direction; writes K = (deg+1)^2 values.
template <typename T>
void sh_basis(int deg, T x, T y, T z, T* b) {
  b[0] = T(SH_C0);
  if (deg < 1) return;
  b[1] = -T(SH_C1) * y;
  b[2] = T(SH_C1) * z;
  b[3] = -T(SH_C1) * x;
  if (deg < 2) return;
  const T xx = x * x, yy = y * y, zz = z * z;
  const T xy = x * y, yz = y * z, xz = x * z;
  b[4] = T(SH_C2[0]) * xy;
  b[5] = T(SH_C2[1]) * yz;
  b[6] = T(SH_C2[2]) * (T(2) * zz - xx - yy);
  b[7] = T(SH_C2[3]) * xz;
  b[8] = T(SH_C2[4]) * (xx - yy);
  if (deg < 3) return;
  b[9] = T(SH_C3[0]) * y * (T(3) * xx - yy);
  b[10] = T(SH_C3[1]) * xy * z;
  b[11] = T(SH_C3[2]) * y * (T(4) * zz - xx - yy);
  b[12] = T(SH_C3[3]) * z * (T(2) * zz - T(3) * xx - T(3) * yy);
  b[13] = T(SH_C3[4]) * x * (T(4) * zz - xx - yy);
  b[14] = T(SH_C3[5]) * z * (xx - yy);
  b[15] = T(SH_C3[6]) * x * (xx - T(3) * yy);
}

// Gradient of each basis function w.r.t. (x, y, z), accumulated as
// g += gb[k] * d b_k / d dir.
template <typename T>
void sh_basis_grad(int deg, T x, T y, T z, const T* gb, T* g) {
  T gx = 0, gy = 0, gz = 0;
  if (deg >= 1) {
    gy += gb[1] * (-T(SH_C1));
    gz += gb[2] * T(SH_C1);
    gx += gb[3] * (-T(SH_C1));
  }
  if (deg >= 2) {
    const T c0 = T(SH_C2[0]), c1 = T(SH_C2[1]), c2 = T(SH_C2[2]), c3 = T(SH_C2[3]), c4 = T(SH_C2[4]);
    gx += gb[4] * c0 * y;            gy += gb[4] * c0 * x;
    gy += gb[5] * c1 * z;            gz += gb[5] * c1 * y;
    gx += gb[6] * (-T(2) * c2 * x);  gy += gb[6] * (-T(2) * c2 * y);  gz += gb[6] * (T(4) * c2 * z);
    gx += gb[7] * c3 * z;            gz += gb[7] * c3 * x;
    gx += gb[8] * (T(2) * c4 * x);   gy += gb[8] * (-T(2) * c4 * y);
  }
  if (deg >= 3) {
    const T xx = x * x, yy = y * y, zz = z * z;
    const T d0 = T(SH_C3[0]), d1 = T(SH_C3[1]), d2 = T(SH_C3[2]), d3 = T(SH_C3[3]);
    const T d4 = T(SH_C3[4]), d5 = T(SH_C3[5]), d6 = T(SH_C3[6]);
    gx += gb[9] * (d0 * T(6) * x * y);           gy += gb[9] * (d0 * (T(3) * xx - T(3) * yy));
    gx += gb[10] * (d1 * y * z);                 gy += gb[10] * (d1 * x * z);      gz += gb[10] * (d1 * x * y);
    gx += gb[11] * (-d2 * T(2) * x * y);         gy += gb[11] * (d2 * (T(4) * zz - xx - T(3) * yy));
    gz += gb[11] * (d2 * T(8) * y * z);
    gx += gb[12] * (-d3 * T(6) * x * z);         gy += gb[12] * (-d3 * T(6) * y * z);
    gz += gb[12] * (d3 * (T(6) * zz - T(3) * xx - T(3) * yy));
    gx += gb[13] * (d4 * (T(4) * zz - T(3) * xx - yy));  gy += gb[13] * (-d4 * T(2) * x * y);
    gz += gb[13] * (d4 * T(8) * x * z);
    gx += gb[14] * (d5 * T(2) * x * z);          gy += gb[14] * (-d5 * T(2) * y * z);  gz += gb[14] * (d5 * (xx - yy));
    gx += gb[15] * (d6 * (T(3) * xx - T(3) * yy));  gy += gb[15] * (-d6 * T(6) * x * y);
  }
  g[0] += gx; g[1] += gy; g[2] += gz;
}

// Column-major 4x4 (the transposed matrices the settings carry): M[r][c] = m[4c + r].
template <typename T> inline T mat(const T* m, int r, int c) { return m[4 * c + r]; }

template <typename T>
struct Camera {
  const T* view;   // world -> camera, transposed storage
  const T* proj;   // full projection (view * proj), transposed storage
  const T* projraw;
  const T* campos;
  int W, H;
  T tanfovx, tanfovy;
  int tile_w, tile_h;
};

// Forward intermediates of one Gaussian, recomputed by the backward.
template <typename T>
struct Fwd {
  bool visible;
  T pv[3];        // view-space mean
  T hx, hy, hw, pw;
  T ndc[2];
  T sig[6];       // 3D covariance (00,01,02,11,12,22)
  T J00, J02, J11, J12;
  bool clampx, clampy;
  T u, v;         // clamped x/z, y/z
  T A[2][3];      // J W
  T B[2][3];      // A Sigma
  T cov[3];       // 2D covariance incl. low-pass
  T det;
  T conic[3];
  int radius;
  T px, py;
  int rect[4];    // xmin, ymin, xmax, ymax (tiles)
  T dir[3];
  T dlen;
  T basis[16];
  T rgb[3];
  bool clamped[3];
};

template <typename T>
void cov3d_from_scale_rot(const T* s, T mod, const T* q, T* sig, T R[3][3]) {
  const T r = q[0], x = q[1], y = q[2], z = q[3];
  R[0][0] = T(1) - T(2) * (y * y + z * z);
  R[0][1] = T(2) * (x * y - r * z);
  R[0][2] = T(2) * (x * z + r * y);
  R[1][0] = T(2) * (x * y + r * z);
  R[1][1] = T(1) - T(2) * (x * x + z * z);
  R[1][2] = T(2) * (y * z - r * x);
  R[2][0] = T(2) * (x * z - r * y);
  R[2][1] = T(2) * (y * z + r * x);
  R[2][2] = T(1) - T(2) * (x * x + y * y);
  const T s0 = mod * s[0], s1 = mod * s[1], s2 = mod * s[2];
  T Mm[3][3];
  for (int i = 0; i < 3; ++i) {
    Mm[i][0] = R[i][0] * s0;
    Mm[i][1] = R[i][1] * s1;
    Mm[i][2] = R[i][2] * s2;
  }
  auto dotr = [&](int i, int j) { return Mm[i][0] * Mm[j][0] + Mm[i][1] * Mm[j][1] + Mm[i][2] * Mm[j][2]; };
  sig[0] = dotr(0, 0); sig[1] = dotr(0, 1); sig[2] = dotr(0, 2);
  sig[3] = dotr(1, 1); sig[4] = dotr(1, 2); sig[5] = dotr(2, 2);
}

template <typename T>
void forward_one(int i, const Camera<T>& cam, int D, int M, const T* means, const T* scales, T mod,
                 const T* rots, const T* shs, const T* colors_pre, const T* cov_pre, Fwd<T>& f) {
  f.visible = false;
  f.radius = 0;
  const T* m = means + 3 * i;
  const T* V = cam.view;
  f.pv[0] = V[0] * m[0] + V[4] * m[1] + V[8] * m[2] + V[12];
  f.pv[1] = V[1] * m[0] + V[5] * m[1] + V[9] * m[2] + V[13];
  f.pv[2] = V[2] * m[0] + V[6] * m[1] + V[10] * m[2] + V[14];
  if (!(f.pv[2] > T(0.2))) return;
  const T* Pm = cam.proj;
  f.hx = Pm[0] * m[0] + Pm[4] * m[1] + Pm[8] * m[2] + Pm[12];
  f.hy = Pm[1] * m[0] + Pm[5] * m[1] + Pm[9] * m[2] + Pm[13];
  f.hw = Pm[3] * m[0] + Pm[7] * m[1] + Pm[11] * m[2] + Pm[15];
  f.pw = T(1) / (f.hw + T(1e-7));
  f.ndc[0] = f.hx * f.pw;
  f.ndc[1] = f.hy * f.pw;

  if (cov_pre) {
    for (int k = 0; k < 6; ++k) f.sig[k] = cov_pre[6 * i + k];
  } else {
    T R[3][3];
    cov3d_from_scale_rot(scales + 3 * i, mod, rots + 4 * i, f.sig, R);
  }
  // EWA: Jacobian of the perspective map at the (clamped) view-space mean
  const T fx = T(cam.W) / (T(2) * cam.tanfovx);
  const T fy = T(cam.H) / (T(2) * cam.tanfovy);
  const T limx = T(1.3) * cam.tanfovx, limy = T(1.3) * cam.tanfovy;
  const T tz = f.pv[2];
  const T txtz = f.pv[0] / tz, tytz = f.pv[1] / tz;
  f.clampx = (txtz < -limx) || (txtz > limx);
  f.clampy = (tytz < -limy) || (tytz > limy);
  f.u = tmin(limx, tmax(-limx, txtz));
  f.v = tmin(limy, tmax(-limy, tytz));
  const T tx = f.u * tz, ty = f.v * tz;
  const T tz2 = tz * tz;
  f.J00 = fx / tz;
  f.J02 = -(fx * tx) / tz2;
  f.J11 = fy / tz;
  f.J12 = -(fy * ty) / tz2;
  for (int c = 0; c < 3; ++c) {
    f.A[0][c] = f.J00 * mat(V, 0, c) + f.J02 * mat(V, 2, c);
    f.A[1][c] = f.J11 * mat(V, 1, c) + f.J12 * mat(V, 2, c);
  }
  const T S[3][3] = {{f.sig[0], f.sig[1], f.sig[2]}, {f.sig[1], f.sig[3], f.sig[4]}, {f.sig[2], f.sig[4], f.sig[5]}};
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 3; ++c) f.B[r][c] = f.A[r][0] * S[0][c] + f.A[r][1] * S[1][c] + f.A[r][2] * S[2][c];
  f.cov[0] = f.B[0][0] * f.A[0][0] + f.B[0][1] * f.A[0][1] + f.B[0][2] * f.A[0][2] + T(0.3);
  f.cov[1] = f.B[0][0] * f.A[1][0] + f.B[0][1] * f.A[1][1] + f.B[0][2] * f.A[1][2];
  f.cov[2] = f.B[1][0] * f.A[1][0] + f.B[1][1] * f.A[1][1] + f.B[1][2] * f.A[1][2] + T(0.3);
  f.det = f.cov[0] * f.cov[2] - f.cov[1] * f.cov[1];
  if (f.det == T(0)) return;
  const T inv = T(1) / f.det;
  f.conic[0] = f.cov[2] * inv;
  f.conic[1] = -f.cov[1] * inv;
  f.conic[2] = f.cov[0] * inv;
  const T mid = T(0.5) * (f.cov[0] + f.cov[2]);
  const T sq = det_sqrt(tmax(T(0.1), mid * mid - f.det));
  const T l1 = mid + sq, l2 = mid - sq;
  T rad = det_ceil(T(3) * det_sqrt(tmax(l1, l2)));
  if (!(rad < T(1073741824))) rad = T(1073741824);  // keeps the int conversion defined
  f.radius = (int)rad;
  f.px = ((f.ndc[0] + T(1)) * T(cam.W) - T(1)) * T(0.5);
  f.py = ((f.ndc[1] + T(1)) * T(cam.H) - T(1)) * T(0.5);
  const int gx = (cam.W + cam.tile_w - 1) / cam.tile_w;
  const int gy = (cam.H + cam.tile_h - 1) / cam.tile_h;
  const T r = T(f.radius);
  const T tw = T(cam.tile_w), th = T(cam.tile_h);
  f.rect[0] = (int)fclamp((f.px - r) / tw, T(0), T(gx));
  f.rect[1] = (int)fclamp((f.py - r) / th, T(0), T(gy));
  f.rect[2] = (int)fclamp((((f.px + r) + tw) - T(1)) / tw, T(0), T(gx));
  f.rect[3] = (int)fclamp((((f.py + r) + th) - T(1)) / th, T(0), T(gy));
  if ((f.rect[2] - f.rect[0]) * (f.rect[3] - f.rect[1]) == 0) return;

  // colour
  if (colors_pre) {
    for (int c = 0; c < 3; ++c) { f.rgb[c] = colors_pre[3 * i + c]; f.clamped[c] = false; }
  } else {
    const T* cp = cam.campos;
    T d0 = m[0] - cp[0], d1 = m[1] - cp[1], d2 = m[2] - cp[2];
    f.dlen = det_sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    f.dir[0] = d0 / f.dlen; f.dir[1] = d1 / f.dlen; f.dir[2] = d2 / f.dlen;
    sh_basis(D, f.dir[0], f.dir[1], f.dir[2], f.basis);
    const int K = (D + 1) * (D + 1);
    const T* sh = shs + (size_t)i * M * 3;
    for (int c = 0; c < 3; ++c) {
      T acc = f.basis[0] * sh[c];
      for (int k = 1; k < K; ++k) acc = acc + f.basis[k] * sh[3 * k + c];
      acc = acc + T(0.5);
      f.clamped[c] = acc < T(0);
      f.rgb[c] = f.clamped[c] ? T(0) : acc;
    }
  }
  f.visible = true;
}

template <typename T>
Camera<T> make_cam(const T* view, const T* proj, const T* projraw, const T* campos, int W, int H, T tfx, T tfy,
                   int tw, int th) {
  Camera<T> c;
  c.view = view; c.proj = proj; c.projraw = projraw; c.campos = campos;
  c.W = W; c.H = H; c.tanfovx = tfx; c.tanfovy = tfy; c.tile_w = tw; c.tile_h = th;
  return c;
}

// ---------------------------------------------------------------------------
// Blend forward for one pixel over a sorted fragment list.
// ---------------------------------------------------------------------------
template <typename T>
struct PixOut { T col[4]; T depth; T opac; T finalT; int ncontrib; };

template <typename T>
inline bool frag_alpha(const T* xy, const T* co, int g, T pxf, T pyf, T& alpha, T& G, T& dx, T& dy) {
  dx = xy[2 * g] - pxf;
  dy = xy[2 * g + 1] - pyf;
  const T a = co[4 * g], b = co[4 * g + 1], c = co[4 * g + 2], o = co[4 * g + 3];
  const T power = T(-0.5) * ((a * dx) * dx + (c * dy) * dy) - (b * dx) * dy;
  if (power > T(0)) return false;
  G = det_exp(power);
  alpha = tmin(T(0.99), o * G);
  if (alpha < T(1.0f / 255.0f)) return false;
  return true;
}

}  // namespace

// ===========================================================================
// extern "C" surface (ctypes).  Host pointers only.
// ===========================================================================
namespace {

template <typename T>
int preprocess_impl(int P, int D, int M, const T* means, const T* scales, T mod, const T* rots, const T* opac,
                    const T* shs, const T* colors_pre, const T* cov_pre, const T* view, const T* proj,
                    const T* campos, int W, int H, T tfx, T tfy, int tw, int th, int* radii, int* tiles,
                    T* xy, T* depth, T* conic_opac, T* rgb, uint8_t* clamped, int* rect) {
  Camera<T> cam = make_cam<T>(view, proj, nullptr, campos, W, H, tfx, tfy, tw, th);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < P; ++i) {
    Fwd<T> f;
    forward_one(i, cam, D, M, means, scales, mod, rots, shs, colors_pre, cov_pre, f);
    if (!f.visible) {
      radii[i] = 0; tiles[i] = 0;
      xy[2 * i] = xy[2 * i + 1] = T(0);
      depth[i] = T(0);
      for (int k = 0; k < 4; ++k) conic_opac[4 * i + k] = T(0);
      for (int k = 0; k < 3; ++k) { rgb[3 * i + k] = T(0); clamped[3 * i + k] = 0; }
      for (int k = 0; k < 4; ++k) rect[4 * i + k] = 0;
      continue;
    }
    radii[i] = f.radius;
    tiles[i] = (f.rect[2] - f.rect[0]) * (f.rect[3] - f.rect[1]);
    xy[2 * i] = f.px; xy[2 * i + 1] = f.py;
    depth[i] = f.pv[2];
    conic_opac[4 * i] = f.conic[0]; conic_opac[4 * i + 1] = f.conic[1]; conic_opac[4 * i + 2] = f.conic[2];
    conic_opac[4 * i + 3] = opac[i];
    for (int k = 0; k < 3; ++k) { rgb[3 * i + k] = f.rgb[k]; clamped[3 * i + k] = f.clamped[k] ? 1 : 0; }
    for (int k = 0; k < 4; ++k) rect[4 * i + k] = f.rect[k];
  }
  return 0;
}

inline uint32_t float_bits(float v) { uint32_t u; std::memcpy(&u, &v, 4); return u; }

template <typename T>
int64_t bin_impl(int P, int W, int H, int tw, int th, const int* tiles, const int* rect, const T* depth,
                 int64_t cap, uint64_t* keys, uint32_t* vals, uint32_t* ranges) {
  const int gx = (W + tw - 1) / tw, gy = (H + th - 1) / th;
  const int ntiles = gx * gy;
  int64_t n = 0;
  for (int i = 0; i < P; ++i) n += tiles[i];
  if (n > cap) return -n;
  struct Frag { uint64_t tile; T depth; uint32_t gid; };
  std::vector<Frag> fr;
  fr.reserve((size_t)n);
  for (int i = 0; i < P; ++i) {
    if (tiles[i] == 0) continue;
    for (int y = rect[4 * i + 1]; y < rect[4 * i + 3]; ++y)
      for (int x = rect[4 * i]; x < rect[4 * i + 2]; ++x) fr.push_back({(uint64_t)(y * gx + x), depth[i], (uint32_t)i});
  }
  std::stable_sort(fr.begin(), fr.end(), [](const Frag& a, const Frag& b) {
    if (a.tile != b.tile) return a.tile < b.tile;
    return a.depth < b.depth;
  });
  for (int t = 0; t < ntiles; ++t) ranges[2 * t] = ranges[2 * t + 1] = 0;
  for (int64_t k = 0; k < n; ++k) {
    keys[k] = (fr[k].tile << 32) | (uint64_t)float_bits((float)fr[k].depth);
    vals[k] = fr[k].gid;
    const uint32_t t = (uint32_t)fr[k].tile;
    if (k == 0 || fr[k - 1].tile != fr[k].tile) ranges[2 * t] = (uint32_t)k;
    if (k == n - 1 || fr[k + 1].tile != fr[k].tile) ranges[2 * t + 1] = (uint32_t)(k + 1);
  }
  return n;
}

// C colour channels (1..4).  colors: [P, C].
template <typename T>
int blend_fwd_impl(int W, int H, int tw, int th, const uint32_t* ranges, const uint32_t* vals, const T* xy,
                   const T* co, const T* colors, int C, const T* depth, const T* bg, T* out_col, T* out_depth,
                   T* out_opac, T* finalT, int* ncontrib, int* ntouched, int P) {
  const int gx = (W + tw - 1) / tw;
  if (ntouched) for (int i = 0; i < P; ++i) ntouched[i] = 0;
#pragma omp parallel for schedule(dynamic, 64)
  for (int pix = 0; pix < W * H; ++pix) {
    const int px = pix % W, py = pix / W;
    const int tile = (py / th) * gx + (px / tw);
    const uint32_t s = ranges[2 * tile], e = ranges[2 * tile + 1];
    const T pxf = T(px), pyf = T(py);
    T Tr = T(1);
    T col[4] = {0, 0, 0, 0};
    T D = 0, A = 0;
    int last = 0;
    for (uint32_t k = s; k < e; ++k) {
      const int g = (int)vals[k];
      T alpha, G, dx, dy;
      if (!frag_alpha(xy, co, g, pxf, pyf, alpha, G, dx, dy)) continue;
      const T testT = Tr * (T(1) - alpha);
      if (testT < T(0.0001)) break;
      const T w = alpha * Tr;
      for (int c = 0; c < C; ++c) col[c] = col[c] + colors[C * g + c] * w;
      D = D + depth[g] * w;
      A = A + w;
      Tr = testT;
      last = (int)(k - s) + 1;
      if (ntouched) {
#pragma omp atomic
        ntouched[g] += 1;
      }
    }
    for (int c = 0; c < C; ++c) out_col[(size_t)c * W * H + pix] = col[c] + Tr * bg[c];
    out_depth[pix] = D;
    out_opac[pix] = A;
    finalT[pix] = Tr;
    ncontrib[pix] = last;
  }
  return 0;
}

// Per-Gaussian gradient of the 2D attributes, 10 floats per Gaussian:
// [d mean_x, d mean_y (pixel units), d conic a, b, c, d opacity, d col 0..2, d depth]
// For C == 1 the single channel gradient goes to slot 6.
template <typename T>
int blend_bwd_impl(int W, int H, int tw, int th, const uint32_t* ranges, const uint32_t* vals, const T* xy,
                   const T* co, const T* colors, int C, const T* depth, const T* bg, const T* finalT,
                   const int* ncontrib, const T* g_col, const T* g_depth, const T* g_opac, T* grad10, int P) {
  const int gx = (W + tw - 1) / tw;
  // one dense accumulator per OpenMP thread, summed in thread order at the end
  const int nthr = omp_get_max_threads();
  std::vector<double> accs((size_t)P * 10 * nthr, 0.0);
#pragma omp parallel num_threads(nthr)
  {
  double* acc = accs.data() + (size_t)omp_get_thread_num() * P * 10;
#pragma omp for schedule(dynamic, 64)
  for (int pix = 0; pix < W * H; ++pix) {
    const int px = pix % W, py = pix / W;
    const int tile = (py / th) * gx + (px / tw);
    const uint32_t s = ranges[2 * tile];
    const int last = ncontrib[pix];
    const T pxf = T(px), pyf = T(py);
    T gC[4] = {0, 0, 0, 0};
    for (int c = 0; c < C; ++c) gC[c] = g_col[(size_t)c * W * H + pix];
    const T gD = g_depth ? g_depth[pix] : T(0);
    const T gA = g_opac ? g_opac[pix] : T(0);
    T bgdot = 0;
    for (int c = 0; c < C; ++c) bgdot = bgdot + bg[c] * gC[c];
    const T TF = finalT[pix];
    T Tr = TF;
    T accC[4] = {0, 0, 0, 0};
    T accD = 0, accA = 0;
    for (int k = last - 1; k >= 0; --k) {
      const int g = (int)vals[s + k];
      T alpha, G, dx, dy;
      if (!frag_alpha(xy, co, g, pxf, pyf, alpha, G, dx, dy)) continue;
      const T om = T(1) - alpha;
      Tr = Tr / om;
      const T w = alpha * Tr;
      T dLda = 0;
      for (int c = 0; c < C; ++c) {
        acc[(size_t)g * 10 + 6 + c] += (double)(w * gC[c]);
        dLda = dLda + (colors[C * g + c] - accC[c]) * gC[c];
      }
      acc[(size_t)g * 10 + 9] += (double)(w * gD);
      dLda = dLda + (depth[g] - accD) * gD;
      dLda = dLda + (T(1) - accA) * gA;
      dLda = dLda * Tr;
      dLda = dLda - (TF / om) * bgdot;
      for (int c = 0; c < C; ++c) accC[c] = alpha * colors[C * g + c] + om * accC[c];
      accD = alpha * depth[g] + om * accD;
      accA = alpha + om * accA;
      const T o = co[4 * g + 3];
      if (o * G > T(0.99)) continue;  // alpha clamped: no dependence on G or o
      acc[(size_t)g * 10 + 5] += (double)(G * dLda);
      const T gpow = (o * dLda) * G;
      const T a = co[4 * g], b = co[4 * g + 1], c2 = co[4 * g + 2];
      acc[(size_t)g * 10 + 0] += (double)(gpow * (-a * dx - b * dy));
      acc[(size_t)g * 10 + 1] += (double)(gpow * (-b * dx - c2 * dy));
      acc[(size_t)g * 10 + 2] += (double)(gpow * (T(-0.5) * dx * dx));
      acc[(size_t)g * 10 + 3] += (double)(gpow * (-dx * dy));
      acc[(size_t)g * 10 + 4] += (double)(gpow * (T(-0.5) * dy * dy));
    }
  }
  }  // omp parallel
#pragma omp parallel for schedule(static)
  for (long long k = 0; k < (long long)P * 10; ++k) {
    double v = 0.0;
    for (int t = 0; t < nthr; ++t) v += accs[(size_t)t * P * 10 + k];
    grad10[k] = (T)v;
  }
  return 0;
}

// Preprocess backward.  Accumulates the 18 pose partial sums
//   [0..2]  sum dL/d p_view                       (translation part)
//   [3..5]  sum p_view x dL/d p_view              (rotation part)
//   [6..14] sum dL/dW (row-major 3x3, covariance path)
//   [15..17] sum dL/d campos
// and optionally per-Gaussian map gradients.
template <typename T>
int preprocess_bwd_impl(int P, int D, int M, const T* means, const T* scales, T mod, const T* rots,
                        const T* shs, const T* colors_pre, const T* cov_pre, const T* view, const T* proj,
                        const T* projraw, const T* campos, int W, int H, T tfx, T tfy, int tw, int th,
                        const T* grad10, int C, double* pose18, T* g_means, T* g_scales, T* g_rots, T* g_shs,
                        T* g_opac, T* g_means2d, T* g_colors, T* g_cov) {
  const double iw[3] = {0.2989, 0.5870, 0.1140};
  Camera<T> cam = make_cam<T>(view, proj, projraw, campos, W, H, tfx, tfy, tw, th);
  double acc[18] = {0};
  const int K = (D + 1) * (D + 1);
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : acc[:18])
  for (int i = 0; i < P; ++i) {
    Fwd<T> f;
    forward_one(i, cam, D, M, means, scales, mod, rots, shs, colors_pre, cov_pre, f);
    if (!f.visible) continue;
    const T* g = grad10 + (size_t)i * 10;
    // conic -> 2D covariance
    const T A = f.cov[0], B = f.cov[1], Cc = f.cov[2];
    const T d2 = f.det * f.det;
    const T ga = g[2], gb = g[3], gc = g[4];
    const T gA = (ga * (-Cc * Cc) + gb * (B * Cc) + gc * (-B * B)) / d2;
    const T gB = (ga * (T(2) * B * Cc) + gb * (-(A * Cc + B * B)) + gc * (T(2) * A * B)) / d2;
    const T gC = (ga * (-B * B) + gb * (A * B) + gc * (-A * A)) / d2;
    const T G2[2][2] = {{gA, T(0.5) * gB}, {T(0.5) * gB, gC}};
    T gM2[2][3];
    for (int r = 0; r < 2; ++r)
      for (int c = 0; c < 3; ++c) gM2[r][c] = T(2) * (G2[r][0] * f.B[0][c] + G2[r][1] * f.B[1][c]);
    // M2 = J W
    T Wm[3][3];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) Wm[r][c] = mat(view, r, c);
    const T gJ00 = gM2[0][0] * Wm[0][0] + gM2[0][1] * Wm[0][1] + gM2[0][2] * Wm[0][2];
    const T gJ02 = gM2[0][0] * Wm[2][0] + gM2[0][1] * Wm[2][1] + gM2[0][2] * Wm[2][2];
    const T gJ11 = gM2[1][0] * Wm[1][0] + gM2[1][1] * Wm[1][1] + gM2[1][2] * Wm[1][2];
    const T gJ12 = gM2[1][0] * Wm[2][0] + gM2[1][1] * Wm[2][1] + gM2[1][2] * Wm[2][2];
    T gW[3][3];
    // J (2x3) has J00, J02 in row 0 and J11, J12 in row 1
    for (int c = 0; c < 3; ++c) {
      gW[0][c] = f.J00 * gM2[0][c];
      gW[1][c] = f.J11 * gM2[1][c];
      gW[2][c] = f.J02 * gM2[0][c] + f.J12 * gM2[1][c];
    }
    const T fx = T(cam.W) / (T(2) * cam.tanfovx);
    const T fy = T(cam.H) / (T(2) * cam.tanfovy);
    const T z = f.pv[2], z2 = z * z;
    T gpv[3] = {0, 0, 0};
    // J00 = fx/z, J11 = fy/z, J02 = -fx u/z, J12 = -fy v/z
    gpv[2] += gJ00 * (-fx / z2) + gJ11 * (-fy / z2);
    if (!f.clampx) {
      gpv[0] += gJ02 * (-fx / z2);
      gpv[2] += gJ02 * (T(2) * fx * f.pv[0] / (z2 * z));
    } else {
      gpv[2] += gJ02 * (fx * f.u / z2);
    }
    if (!f.clampy) {
      gpv[1] += gJ12 * (-fy / z2);
      gpv[2] += gJ12 * (T(2) * fy * f.pv[1] / (z2 * z));
    } else {
      gpv[2] += gJ12 * (fy * f.v / z2);
    }
    // mean2D (pixel units) -> homogeneous projection -> view space
    const T ghx = g[0] * (T(0.5) * T(cam.W)) * f.pw;
    const T ghy = g[1] * (T(0.5) * T(cam.H)) * f.pw;
    const T ghw = -(g[0] * (T(0.5) * T(cam.W)) * f.hx + g[1] * (T(0.5) * T(cam.H)) * f.hy) * f.pw * f.pw;
    for (int c = 0; c < 3; ++c)
      gpv[c] += ghx * mat(projraw, 0, c) + ghy * mat(projraw, 1, c) + ghw * mat(projraw, 3, c);
    // depth
    gpv[2] += g[9];
    // colour -> SH direction -> camera centre
    T gcol[3] = {0, 0, 0};
    if (C == 3) {
      gcol[0] = g[6]; gcol[1] = g[7]; gcol[2] = g[8];
    } else {  // one grey channel I = sum_c iw_c rgb_c
      gcol[0] = T(iw[0]) * g[6]; gcol[1] = T(iw[1]) * g[6]; gcol[2] = T(iw[2]) * g[6];
    }
    T gd[3] = {0, 0, 0};
    if (!colors_pre) {
      for (int c = 0; c < 3; ++c) if (f.clamped[c]) gcol[c] = T(0);
      T gbasis[16];
      const T* sh = shs + (size_t)i * M * 3;
      for (int k = 0; k < K; ++k) gbasis[k] = gcol[0] * sh[3 * k] + gcol[1] * sh[3 * k + 1] + gcol[2] * sh[3 * k + 2];
      T gdir[3] = {0, 0, 0};
      sh_basis_grad(D, f.dir[0], f.dir[1], f.dir[2], gbasis, gdir);
      const T dd = f.dir[0] * gdir[0] + f.dir[1] * gdir[1] + f.dir[2] * gdir[2];
      for (int c = 0; c < 3; ++c) gd[c] = (gdir[c] - f.dir[c] * dd) / f.dlen;
      if (g_shs) {
        for (int k = 0; k < K; ++k)
          for (int c = 0; c < 3; ++c) g_shs[(size_t)i * M * 3 + 3 * k + c] = f.basis[k] * gcol[c];
      }
    }
    // pose partials
    acc[0] += gpv[0]; acc[1] += gpv[1]; acc[2] += gpv[2];
    acc[3] += (double)(f.pv[1] * gpv[2] - f.pv[2] * gpv[1]);
    acc[4] += (double)(f.pv[2] * gpv[0] - f.pv[0] * gpv[2]);
    acc[5] += (double)(f.pv[0] * gpv[1] - f.pv[1] * gpv[0]);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) acc[6 + 3 * r + c] += gW[r][c];
    acc[15] -= gd[0]; acc[16] -= gd[1]; acc[17] -= gd[2];

    // per-Gaussian map gradients
    if (g_opac) g_opac[i] = g[5];
    if (g_colors && colors_pre) for (int c = 0; c < 3; ++c) g_colors[3 * i + c] = gcol[c];
    if (g_means2d) {
      g_means2d[3 * i] = g[0] * (T(0.5) * T(cam.W));
      g_means2d[3 * i + 1] = g[1] * (T(0.5) * T(cam.H));
      g_means2d[3 * i + 2] = T(0);
    }
    if (g_means) {
      for (int c = 0; c < 3; ++c)
        g_means[3 * i + c] = Wm[0][c] * gpv[0] + Wm[1][c] * gpv[1] + Wm[2][c] * gpv[2] + gd[c];
    }
    if (g_scales || g_rots || g_cov) {
      // dL/dSigma = M2^T G2 M2 (3x3 symmetric)
      T gS[3][3];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          T s = 0;
          for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) s += f.A[a][r] * G2[a][b] * f.A[b][c];
          gS[r][c] = s;
        }
      if (cov_pre) {
        if (g_cov) {
          g_cov[6 * i + 0] = gS[0][0]; g_cov[6 * i + 1] = T(2) * gS[0][1]; g_cov[6 * i + 2] = T(2) * gS[0][2];
          g_cov[6 * i + 3] = gS[1][1]; g_cov[6 * i + 4] = T(2) * gS[1][2]; g_cov[6 * i + 5] = gS[2][2];
        }
      } else {
        T sig[6], R[3][3];
        const T* q = rots + 4 * i;
        const T* s = scales + 3 * i;
        cov3d_from_scale_rot(s, mod, q, sig, R);
        const T sc[3] = {mod * s[0], mod * s[1], mod * s[2]};
        // Sigma = Mm Mm^T, Mm = R diag(sc): dL/dMm = 2 gS Mm
        T gMm[3][3];
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c)
            gMm[r][c] = T(2) * (gS[r][0] * R[0][c] * sc[c] + gS[r][1] * R[1][c] * sc[c] + gS[r][2] * R[2][c] * sc[c]);
        if (g_scales)
          for (int c = 0; c < 3; ++c)
            g_scales[3 * i + c] = (gMm[0][c] * R[0][c] + gMm[1][c] * R[1][c] + gMm[2][c] * R[2][c]) * mod;
        if (g_rots) {
          T gR[3][3];
          for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) gR[r][c] = gMm[r][c] * sc[c];
          const T qr = q[0], qx = q[1], qy = q[2], qz = q[3];
          const T dR_r[3][3] = {{0, -2 * qz, 2 * qy}, {2 * qz, 0, -2 * qx}, {-2 * qy, 2 * qx, 0}};
          const T dR_x[3][3] = {{0, 2 * qy, 2 * qz}, {2 * qy, -4 * qx, -2 * qr}, {2 * qz, 2 * qr, -4 * qx}};
          const T dR_y[3][3] = {{-4 * qy, 2 * qx, 2 * qr}, {2 * qx, 0, 2 * qz}, {-2 * qr, 2 * qz, -4 * qy}};
          const T dR_z[3][3] = {{-4 * qz, -2 * qr, 2 * qx}, {2 * qr, -4 * qz, 2 * qy}, {2 * qx, 2 * qy, 0}};
          T out[4] = {0, 0, 0, 0};
          for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
              out[0] += gR[r][c] * dR_r[r][c];
              out[1] += gR[r][c] * dR_x[r][c];
              out[2] += gR[r][c] * dR_y[r][c];
              out[3] += gR[r][c] * dR_z[r][c];
            }
          for (int k = 0; k < 4; ++k) g_rots[4 * i + k] = out[k];
        }
      }
    }
  }
  for (int k = 0; k < 18; ++k) pose18[k] = acc[k];
  return 0;
}

// SE(3) left Jacobian in [rho; theta] order via the series sum_n ad^n/(n+1)!.
void se3_left_jacobian(const double* xi, double J[6][6]) {
  double ad[6][6] = {};
  auto skew_into = [](double m[6][6], int r0, int c0, const double* v) {
    m[r0 + 0][c0 + 1] = -v[2]; m[r0 + 0][c0 + 2] = v[1];
    m[r0 + 1][c0 + 0] = v[2];  m[r0 + 1][c0 + 2] = -v[0];
    m[r0 + 2][c0 + 0] = -v[1]; m[r0 + 2][c0 + 1] = v[0];
  };
  skew_into(ad, 0, 0, xi + 3);
  skew_into(ad, 0, 3, xi);
  skew_into(ad, 3, 3, xi + 3);
  double term[6][6], tmp[6][6];
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 6; ++c) { term[r][c] = (r == c) ? 1.0 : 0.0; J[r][c] = term[r][c]; }
  for (int n = 1; n < 30; ++n) {
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c < 6; ++c) {
        double s = 0;
        for (int k = 0; k < 6; ++k) s += term[r][k] * ad[k][c];
        tmp[r][c] = s / (double)(n + 1);
      }
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c < 6; ++c) { term[r][c] = tmp[r][c]; J[r][c] += term[r][c]; }
  }
}

}  // namespace

extern "C" {

int orc_preprocess_f32(int P, int D, int M, const float* means, const float* scales, float mod, const float* rots,
                       const float* opac, const float* shs, const float* colors_pre, const float* cov_pre,
                       const float* view, const float* proj, const float* campos, int W, int H, float tfx,
                       float tfy, int tw, int th, int* radii, int* tiles, float* xy, float* depth,
                       float* conic_opac, float* rgb, uint8_t* clamped, int* rect) {
  return preprocess_impl<float>(P, D, M, means, scales, mod, rots, opac, shs, colors_pre, cov_pre, view, proj,
                                campos, W, H, tfx, tfy, tw, th, radii, tiles, xy, depth, conic_opac, rgb, clamped,
                                rect);
}
int orc_preprocess_f64(int P, int D, int M, const double* means, const double* scales, double mod,
                       const double* rots, const double* opac, const double* shs, const double* colors_pre,
                       const double* cov_pre, const double* view, const double* proj, const double* campos, int W,
                       int H, double tfx, double tfy, int tw, int th, int* radii, int* tiles, double* xy,
                       double* depth, double* conic_opac, double* rgb, uint8_t* clamped, int* rect) {
  return preprocess_impl<double>(P, D, M, means, scales, mod, rots, opac, shs, colors_pre, cov_pre, view, proj,
                                 campos, W, H, tfx, tfy, tw, th, radii, tiles, xy, depth, conic_opac, rgb, clamped,
                                 rect);
}

int64_t orc_bin_f32(int P, int W, int H, int tw, int th, const int* tiles, const int* rect, const float* depth,
                    int64_t cap, uint64_t* keys, uint32_t* vals, uint32_t* ranges) {
  return bin_impl<float>(P, W, H, tw, th, tiles, rect, depth, cap, keys, vals, ranges);
}
int64_t orc_bin_f64(int P, int W, int H, int tw, int th, const int* tiles, const int* rect, const double* depth,
                    int64_t cap, uint64_t* keys, uint32_t* vals, uint32_t* ranges) {
  return bin_impl<double>(P, W, H, tw, th, tiles, rect, depth, cap, keys, vals, ranges);
}

int orc_blend_fwd_f32(int W, int H, int tw, int th, const uint32_t* ranges, const uint32_t* vals, const float* xy,
                      const float* co, const float* colors, int C, const float* depth, const float* bg,
                      float* out_col, float* out_depth, float* out_opac, float* finalT, int* ncontrib,
                      int* ntouched, int P) {
  return blend_fwd_impl<float>(W, H, tw, th, ranges, vals, xy, co, colors, C, depth, bg, out_col, out_depth,
                               out_opac, finalT, ncontrib, ntouched, P);
}
int orc_blend_fwd_f64(int W, int H, int tw, int th, const uint32_t* ranges, const uint32_t* vals, const double* xy,
                      const double* co, const double* colors, int C, const double* depth, const double* bg,
                      double* out_col, double* out_depth, double* out_opac, double* finalT, int* ncontrib,
                      int* ntouched, int P) {
  return blend_fwd_impl<double>(W, H, tw, th, ranges, vals, xy, co, colors, C, depth, bg, out_col, out_depth,
                                out_opac, finalT, ncontrib, ntouched, P);
}

int orc_blend_bwd_f32(int W, int H, int tw, int th, const uint32_t* ranges, const uint32_t* vals, const float* xy,
                      const float* co, const float* colors, int C, const float* depth, const float* bg,
                      const float* finalT, const int* ncontrib, const float* g_col, const float* g_depth,
                      const float* g_opac, float* grad10, int P) {
  return blend_bwd_impl<float>(W, H, tw, th, ranges, vals, xy, co, colors, C, depth, bg, finalT, ncontrib, g_col,
                               g_depth, g_opac, grad10, P);
}
int orc_blend_bwd_f64(int W, int H, int tw, int th, const uint32_t* ranges, const uint32_t* vals,
                      const double* xy, const double* co, const double* colors, int C, const double* depth,
                      const double* bg, const double* finalT, const int* ncontrib, const double* g_col,
                      const double* g_depth, const double* g_opac, double* grad10, int P) {
  return blend_bwd_impl<double>(W, H, tw, th, ranges, vals, xy, co, colors, C, depth, bg, finalT, ncontrib, g_col,
                                g_depth, g_opac, grad10, P);
}

int orc_preprocess_bwd_f32(int P, int D, int M, const float* means, const float* scales, float mod,
                           const float* rots, const float* shs, const float* colors_pre, const float* cov_pre,
                           const float* view, const float* proj, const float* projraw, const float* campos, int W,
                           int H, float tfx, float tfy, int tw, int th, const float* grad10, int C, double* pose18,
                           float* g_means, float* g_scales, float* g_rots, float* g_shs, float* g_opac,
                           float* g_means2d, float* g_colors, float* g_cov) {
  return preprocess_bwd_impl<float>(P, D, M, means, scales, mod, rots, shs, colors_pre, cov_pre, view, proj,
                                    projraw, campos, W, H, tfx, tfy, tw, th, grad10, C, pose18, g_means, g_scales,
                                    g_rots, g_shs, g_opac, g_means2d, g_colors, g_cov);
}
int orc_preprocess_bwd_f64(int P, int D, int M, const double* means, const double* scales, double mod,
                           const double* rots, const double* shs, const double* colors_pre, const double* cov_pre,
                           const double* view, const double* proj, const double* projraw, const double* campos,
                           int W, int H, double tfx, double tfy, int tw, int th, const double* grad10, int C,
                           double* pose18, double* g_means, double* g_scales, double* g_rots, double* g_shs,
                           double* g_opac, double* g_means2d, double* g_colors, double* g_cov) {
  return preprocess_bwd_impl<double>(P, D, M, means, scales, mod, rots, shs, colors_pre, cov_pre, view, proj,
                                     projraw, campos, W, H, tfx, tfy, tw, th, grad10, C, pose18, g_means, g_scales,
                                     g_rots, g_shs, g_opac, g_means2d, g_colors, g_cov);
}

// Pose finalize: 18 partial sums -> 12 parameter gradients
//   out[0..2] = dL/d theta (cam_rot_delta), out[3..5] = dL/d rho (cam_trans_delta),
//   out[6..8] = dL/d w (cam_w_delta),       out[9..11] = dL/d v (cam_v_delta).
// view: render camera (transposed storage), vel: vel_transofrm (transposed storage),
// dt: delta_time, lin/ang: linear_vel / angular_vel.
// Render pose T_r = E * Exp(delta) * T_c with E = Exp(dt * [v; w]):
//   dL/d delta = Ad(E)^T g_r,   dL/d [v; w] = dt * J_l(dt [v; w])^T g_r.
void orc_pose_finalize(const double* pose18, const double* view, const double* vel, double dt, const double* lin,
                       const double* ang, double* out) {
  double Wm[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) Wm[r][c] = view[4 * c + r];
  const double* S = pose18;
  double g[6];
  // translation: sum dL/dp_view - W * sum dL/dcampos
  for (int r = 0; r < 3; ++r) g[r] = S[r] - (Wm[r][0] * S[15] + Wm[r][1] * S[16] + Wm[r][2] * S[17]);
  // rotation: sum p x g + fold(dL/dW)
  const double* GW = S + 6;
  auto gw = [&](int r, int c) { return GW[3 * r + c]; };
  double f0 = 0, f1 = 0, f2 = 0;
  for (int j = 0; j < 3; ++j) {
    f0 += -gw(1, j) * Wm[2][j] + gw(2, j) * Wm[1][j];
    f1 += gw(0, j) * Wm[2][j] - gw(2, j) * Wm[0][j];
    f2 += -gw(0, j) * Wm[1][j] + gw(1, j) * Wm[0][j];
  }
  g[3] = S[3] + f0; g[4] = S[4] + f1; g[5] = S[5] + f2;
  // adjoint of E
  double RE[3][3], tE[3];
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) RE[r][c] = vel[4 * c + r];
    tE[r] = vel[4 * 3 + r];
  }
  const double tg[3] = {tE[1] * g[2] - tE[2] * g[1], tE[2] * g[0] - tE[0] * g[2], tE[0] * g[1] - tE[1] * g[0]};
  for (int c = 0; c < 3; ++c) {
    double rho = 0, th = 0;
    for (int r = 0; r < 3; ++r) {
      rho += RE[r][c] * g[r];
      th += RE[r][c] * (g[3 + r] - tg[r]);
    }
    out[3 + c] = rho;
    out[c] = th;
  }
  // velocity
  double xi[6];
  for (int k = 0; k < 3; ++k) { xi[k] = dt * lin[k]; xi[3 + k] = dt * ang[k]; }
  double J[6][6];
  se3_left_jacobian(xi, J);
  double gv[6];
  for (int c = 0; c < 6; ++c) {
    double s = 0;
    for (int r = 0; r < 6; ++r) s += J[r][c] * g[r];
    gv[c] = dt * s;
  }
  for (int k = 0; k < 3; ++k) { out[9 + k] = gv[k]; out[6 + k] = gv[3 + k]; }
}

// The deterministic exp, exported so tests can bound its error against libm.
float orc_det_expf(float x) { return det_exp(x); }

// OpenMP threads of the calls that follow on this thread (the CPU baseline
// times 1 thread and all threads); returns the previous maximum.
int orc_set_num_threads(int n) {
  const int prev = omp_get_max_threads();
  if (n > 0) omp_set_num_threads(n);
  return prev;
}

}  // extern "C"
