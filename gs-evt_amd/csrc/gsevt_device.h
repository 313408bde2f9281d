// Device-side building blocks shared by the gfx950 kernels.
//
// Every floating-point expression here is written in one fixed evaluation
// order and the library is compiled with -ffp-contract=off and correctly
// rounded division/sqrt, so preprocess outputs, tile keys and the forward
// blend are reproducible bit for bit by a CPU implementation that uses the
// same order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gsevt {

constexpr int WAVE = 64;

// Deterministic exp for x in [-87, 0]: exp(x) = 2^k p(r), r = x - k ln2, p a
// degree-7 Taylor polynomial in fma form.  Uses only IEEE add/mul/fma/rint/
// ldexp, so its bits do not depend on the hardware exp unit (max error < 2 ulp
// on [-87, 0]); it equals the oracle's det_expf there (which also clamps
// inputs below -87).  The blends call it on every lane but keep the result
// only where the power lies in [-L', 0] with L' < 6; elsewhere their selects
// discard the value, so no clamp is applied.
__device__ __forceinline__ float det_exp_unclamped(float x) {
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

__device__ __forceinline__ float tmin(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float tmax(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float fclamp(float v, float lo, float hi) { return fminf(hi, fmaxf(lo, v)); }

// Real SH constants, degree 0..3 (sign convention of the 3DGS SH evaluator).
#define GS_SH_C0 0.28209479177387814f
#define GS_SH_C1 0.4886025119029199f
__constant__ static const float SH_C2[5] = {1.0925484305920792f, -1.0925484305920792f, 0.31539156525252005f,
                                           -1.0925484305920792f, 0.5462742152960396f};
__constant__ static const float SH_C3[7] = {-0.5900435899266435f, 2.890611442640554f, -0.4570457994644658f,
                                           0.3731763325901154f, -0.4570457994644658f, 1.445305721320277f,
                                           -0.5900435899266435f};

__device__ __forceinline__ void sh_basis(int deg, float x, float y, float z, float* b) {
  b[0] = GS_SH_C0;
  if (deg < 1) return;
  b[1] = -GS_SH_C1 * y;
  b[2] = GS_SH_C1 * z;
  b[3] = -GS_SH_C1 * x;
  if (deg < 2) return;
  const float xx = x * x, yy = y * y, zz = z * z;
  const float xy = x * y, yz = y * z, xz = x * z;
  b[4] = SH_C2[0] * xy;
  b[5] = SH_C2[1] * yz;
  b[6] = SH_C2[2] * (2.0f * zz - xx - yy);
  b[7] = SH_C2[3] * xz;
  b[8] = SH_C2[4] * (xx - yy);
  if (deg < 3) return;
  b[9] = SH_C3[0] * y * (3.0f * xx - yy);
  b[10] = SH_C3[1] * xy * z;
  b[11] = SH_C3[2] * y * (4.0f * zz - xx - yy);
  b[12] = SH_C3[3] * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
  b[13] = SH_C3[4] * x * (4.0f * zz - xx - yy);
  b[14] = SH_C3[5] * z * (xx - yy);
  b[15] = SH_C3[6] * x * (xx - 3.0f * yy);
}

// g += sum_k gb[k] d b_k / d dir
__device__ __forceinline__ void sh_basis_grad(int deg, float x, float y, float z, const float* gb, float* g) {
  float gx = 0.f, gy = 0.f, gz = 0.f;
  if (deg >= 1) {
    gy += gb[1] * (-GS_SH_C1);
    gz += gb[2] * GS_SH_C1;
    gx += gb[3] * (-GS_SH_C1);
  }
  if (deg >= 2) {
    const float c0 = SH_C2[0], c1 = SH_C2[1], c2 = SH_C2[2], c3 = SH_C2[3], c4 = SH_C2[4];
    gx += gb[4] * c0 * y;            gy += gb[4] * c0 * x;
    gy += gb[5] * c1 * z;            gz += gb[5] * c1 * y;
    gx += gb[6] * (-2.0f * c2 * x);  gy += gb[6] * (-2.0f * c2 * y);  gz += gb[6] * (4.0f * c2 * z);
    gx += gb[7] * c3 * z;            gz += gb[7] * c3 * x;
    gx += gb[8] * (2.0f * c4 * x);   gy += gb[8] * (-2.0f * c4 * y);
  }
  if (deg >= 3) {
    const float xx = x * x, yy = y * y, zz = z * z;
    const float d0 = SH_C3[0], d1 = SH_C3[1], d2 = SH_C3[2], d3 = SH_C3[3];
    const float d4 = SH_C3[4], d5 = SH_C3[5], d6 = SH_C3[6];
    gx += gb[9] * (d0 * 6.0f * x * y);           gy += gb[9] * (d0 * (3.0f * xx - 3.0f * yy));
    gx += gb[10] * (d1 * y * z);                 gy += gb[10] * (d1 * x * z);      gz += gb[10] * (d1 * x * y);
    gx += gb[11] * (-d2 * 2.0f * x * y);         gy += gb[11] * (d2 * (4.0f * zz - xx - 3.0f * yy));
    gz += gb[11] * (d2 * 8.0f * y * z);
    gx += gb[12] * (-d3 * 6.0f * x * z);         gy += gb[12] * (-d3 * 6.0f * y * z);
    gz += gb[12] * (d3 * (6.0f * zz - 3.0f * xx - 3.0f * yy));
    gx += gb[13] * (d4 * (4.0f * zz - 3.0f * xx - yy));  gy += gb[13] * (-d4 * 2.0f * x * y);
    gz += gb[13] * (d4 * 8.0f * x * z);
    gx += gb[14] * (d5 * 2.0f * x * z);          gy += gb[14] * (-d5 * 2.0f * y * z);  gz += gb[14] * (d5 * (xx - yy));
    gx += gb[15] * (d6 * (3.0f * xx - 3.0f * yy));  gy += gb[15] * (-d6 * 6.0f * x * y);
  }
  g[0] += gx; g[1] += gy; g[2] += gz;
}

// One render's camera as the kernels read it.  Matrices use the transposed
// storage of the rasterizer settings: M[r][c] = m[4c + r].
struct CamDev {
  float view[16];
  float proj[16];
  float projraw[16];
  float campos[3];
  float tanfovx, tanfovy;
  int W, H;
  int tile_w, tile_h;
  int grid_x, grid_y;
};

__device__ __forceinline__ float mat(const float* m, int r, int c) { return m[4 * c + r]; }

// Per-Gaussian forward state (recomputed by the backward instead of stored).
struct Fwd {
  float pv[3];
  float hx, hy, hw, pw;
  float ndc0, ndc1;
  float sig[6];
  float J00, J02, J11, J12;
  bool clampx, clampy;
  float u, v;
  float A[2][3];
  float B[2][3];
  float cov[3];
  float det;
  float conic[3];
  int radius;
  float px, py;
  int rect[4];
};

__device__ __forceinline__ void cov3d(const float* s, float mod, const float* q, float* sig, float R[3][3]) {
  const float r = q[0], x = q[1], y = q[2], z = q[3];
  R[0][0] = 1.0f - 2.0f * (y * y + z * z);
  R[0][1] = 2.0f * (x * y - r * z);
  R[0][2] = 2.0f * (x * z + r * y);
  R[1][0] = 2.0f * (x * y + r * z);
  R[1][1] = 1.0f - 2.0f * (x * x + z * z);
  R[1][2] = 2.0f * (y * z - r * x);
  R[2][0] = 2.0f * (x * z - r * y);
  R[2][1] = 2.0f * (y * z + r * x);
  R[2][2] = 1.0f - 2.0f * (x * x + y * y);
  const float s0 = mod * s[0], s1 = mod * s[1], s2 = mod * s[2];
  float Mm[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    Mm[i][0] = R[i][0] * s0;
    Mm[i][1] = R[i][1] * s1;
    Mm[i][2] = R[i][2] * s2;
  }
#define GS_DOTR(i, j) (Mm[i][0] * Mm[j][0] + Mm[i][1] * Mm[j][1] + Mm[i][2] * Mm[j][2])
  sig[0] = GS_DOTR(0, 0); sig[1] = GS_DOTR(0, 1); sig[2] = GS_DOTR(0, 2);
  sig[3] = GS_DOTR(1, 1); sig[4] = GS_DOTR(1, 2); sig[5] = GS_DOTR(2, 2);
#undef GS_DOTR
}

// Tile rectangle of a splat; fminf/fmaxf clamp before the integer conversion
// keeps out-of-range centres well defined.
__device__ __forceinline__ void tile_rect(float px, float py, int radius, const CamDev& c, int* rect) {
  const float r = (float)radius;
  const float tw = (float)c.tile_w, th = (float)c.tile_h;
  rect[0] = (int)fclamp((px - r) / tw, 0.0f, (float)c.grid_x);
  rect[1] = (int)fclamp((py - r) / th, 0.0f, (float)c.grid_y);
  rect[2] = (int)fclamp((((px + r) + tw) - 1.0f) / tw, 0.0f, (float)c.grid_x);
  rect[3] = (int)fclamp((((py + r) + th) - 1.0f) / th, 0.0f, (float)c.grid_y);
}

// Geometry of one Gaussian (everything except colour).  Returns false when
// the Gaussian is culled.  `m` is the world mean, `sig` the 3D covariance.
__device__ __forceinline__ bool project_one(const float* m, const float* sig_in, const CamDev& c, Fwd& f) {
  const float* V = c.view;
  f.pv[0] = V[0] * m[0] + V[4] * m[1] + V[8] * m[2] + V[12];
  f.pv[1] = V[1] * m[0] + V[5] * m[1] + V[9] * m[2] + V[13];
  f.pv[2] = V[2] * m[0] + V[6] * m[1] + V[10] * m[2] + V[14];
  if (!(f.pv[2] > 0.2f)) return false;
  const float* Pm = c.proj;
  f.hx = Pm[0] * m[0] + Pm[4] * m[1] + Pm[8] * m[2] + Pm[12];
  f.hy = Pm[1] * m[0] + Pm[5] * m[1] + Pm[9] * m[2] + Pm[13];
  f.hw = Pm[3] * m[0] + Pm[7] * m[1] + Pm[11] * m[2] + Pm[15];
  f.pw = 1.0f / (f.hw + 1e-7f);
  f.ndc0 = f.hx * f.pw;
  f.ndc1 = f.hy * f.pw;
#pragma unroll
  for (int k = 0; k < 6; ++k) f.sig[k] = sig_in[k];
  const float fx = (float)c.W / (2.0f * c.tanfovx);
  const float fy = (float)c.H / (2.0f * c.tanfovy);
  const float limx = 1.3f * c.tanfovx, limy = 1.3f * c.tanfovy;
  const float tz = f.pv[2];
  const float txtz = f.pv[0] / tz, tytz = f.pv[1] / tz;
  f.clampx = (txtz < -limx) || (txtz > limx);
  f.clampy = (tytz < -limy) || (tytz > limy);
  f.u = tmin(limx, tmax(-limx, txtz));
  f.v = tmin(limy, tmax(-limy, tytz));
  const float tx = f.u * tz, ty = f.v * tz;
  const float tz2 = tz * tz;
  f.J00 = fx / tz;
  f.J02 = -(fx * tx) / tz2;
  f.J11 = fy / tz;
  f.J12 = -(fy * ty) / tz2;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    f.A[0][k] = f.J00 * mat(V, 0, k) + f.J02 * mat(V, 2, k);
    f.A[1][k] = f.J11 * mat(V, 1, k) + f.J12 * mat(V, 2, k);
  }
  const float S[3][3] = {{f.sig[0], f.sig[1], f.sig[2]}, {f.sig[1], f.sig[3], f.sig[4]}, {f.sig[2], f.sig[4], f.sig[5]}};
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int k = 0; k < 3; ++k) f.B[r][k] = f.A[r][0] * S[0][k] + f.A[r][1] * S[1][k] + f.A[r][2] * S[2][k];
  f.cov[0] = f.B[0][0] * f.A[0][0] + f.B[0][1] * f.A[0][1] + f.B[0][2] * f.A[0][2] + 0.3f;
  f.cov[1] = f.B[0][0] * f.A[1][0] + f.B[0][1] * f.A[1][1] + f.B[0][2] * f.A[1][2];
  f.cov[2] = f.B[1][0] * f.A[1][0] + f.B[1][1] * f.A[1][1] + f.B[1][2] * f.A[1][2] + 0.3f;
  f.det = f.cov[0] * f.cov[2] - f.cov[1] * f.cov[1];
  if (f.det == 0.0f) return false;
  const float inv = 1.0f / f.det;
  f.conic[0] = f.cov[2] * inv;
  f.conic[1] = -f.cov[1] * inv;
  f.conic[2] = f.cov[0] * inv;
  const float mid = 0.5f * (f.cov[0] + f.cov[2]);
  const float sq = sqrtf(tmax(0.1f, mid * mid - f.det));
  const float l1 = mid + sq, l2 = mid - sq;
  float rad = ceilf(3.0f * sqrtf(tmax(l1, l2)));
  if (!(rad < 1073741824.0f)) rad = 1073741824.0f;
  f.radius = (int)rad;
  f.px = ((f.ndc0 + 1.0f) * (float)c.W - 1.0f) * 0.5f;
  f.py = ((f.ndc1 + 1.0f) * (float)c.H - 1.0f) * 0.5f;
  tile_rect(f.px, f.py, f.radius, c, f.rect);
  if ((f.rect[2] - f.rect[0]) * (f.rect[3] - f.rect[1]) == 0) return false;
  return true;
}

// SH colour at the view direction; writes rgb (clamped at 0) and clamp flags.
__device__ __forceinline__ void sh_color(int deg, int M, const float* m, const float* sh, const CamDev& c,
                                         float* dir, float& dlen, float* rgb, bool* clamped) {
  const float d0 = m[0] - c.campos[0], d1 = m[1] - c.campos[1], d2 = m[2] - c.campos[2];
  dlen = sqrtf(d0 * d0 + d1 * d1 + d2 * d2);
  dir[0] = d0 / dlen; dir[1] = d1 / dlen; dir[2] = d2 / dlen;
  float b[16];
  sh_basis(deg, dir[0], dir[1], dir[2], b);
  const int K = (deg + 1) * (deg + 1);
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) {
    float acc = b[0] * sh[ch];
    for (int k = 1; k < K; ++k) acc = acc + b[k] * sh[3 * k + ch];
    acc = acc + 0.5f;
    clamped[ch] = acc < 0.0f;
    rgb[ch] = clamped[ch] ? 0.0f : acc;
  }
}

// SH colour with a compile-time degree and the coefficients in registers
// (sh[3k + c]); same operation order as sh_color.
template <int D>
__device__ __forceinline__ void sh_color_reg(const float* m, const float* sh, const CamDev& c, float* dir,
                                             float& dlen, float* rgb, bool* clamped) {
  constexpr int K = (D + 1) * (D + 1);
  const float d0 = m[0] - c.campos[0], d1 = m[1] - c.campos[1], d2 = m[2] - c.campos[2];
  dlen = sqrtf(d0 * d0 + d1 * d1 + d2 * d2);
  dir[0] = d0 / dlen; dir[1] = d1 / dlen; dir[2] = d2 / dlen;
  float b[16];
  sh_basis(D, dir[0], dir[1], dir[2], b);
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) {
    float acc = b[0] * sh[ch];
#pragma unroll
    for (int k = 1; k < K; ++k) acc = acc + b[k] * sh[3 * k + ch];
    acc = acc + 0.5f;
    clamped[ch] = acc < 0.0f;
    rgb[ch] = clamped[ch] ? 0.0f : acc;
  }
}

// ---------------------------------------------------------------------------
// Wave helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}

__device__ __forceinline__ int wave_max_i(int v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    const int w = __shfl_xor(v, o, WAVE);
    v = v > w ? v : w;
  }
  return v;
}

// Register-only cross-lane steps (gfx950 permlane swaps and DPP row modes).
// swap_add32(a, b): lanes 0-31 receive a[l] + a[l + 32], lanes 32-63 receive
// b[l - 32] + b[l] (v_permlane32_swap exchanges the upper half of a with the
// lower half of b).  swap_add16: the same within each 32-lane half at offset 16.
__device__ __forceinline__ float swap_add32(float a, float b) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_int(a), __float_as_int(b), false, false);
  return __int_as_float(r[0]) + __int_as_float(r[1]);
}
__device__ __forceinline__ float swap_add16(float a, float b) {
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_int(a), __float_as_int(b), false, false);
  return __int_as_float(r[0]) + __int_as_float(r[1]);
}
// DPP row modes: value of lane 15 - (l % 16) of the row, of lane 7 - (l % 8)
// of the 8-lane group, of lane l ^ 2 and of lane l ^ 1 within the quad.
__device__ __forceinline__ float dpp_mirror16(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x140, 0xf, 0xf, false));
}
__device__ __forceinline__ float dpp_mirror8(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x141, 0xf, 0xf, false));
}
__device__ __forceinline__ float dpp_quad_xor2(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x4E, 0xf, 0xf, false));
}
__device__ __forceinline__ float dpp_quad_xor1(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0xB1, 0xf, 0xf, false));
}

// Reduce-scatter of 16 per-lane values over the wave.  Each step halves the
// live slots: the lane keeps one slot of a pair and adds its partner's copy
// of it (partners differ in lane bit 5, 4, 3, 2).  Lane L ends with the total
// of slot ((L>>5)&1)*8 + ((L>>4)&1)*4 + ((L>>3)&1)*2 + ((L>>2)&1); lanes with
// (L & 3) == 0 hold each slot once.
__device__ __forceinline__ float wave_reduce16(float* v, int lane) {
#pragma unroll
  for (int i = 0; i < 8; ++i) v[i] = swap_add32(v[i], v[8 + i]);
#pragma unroll
  for (int i = 0; i < 4; ++i) v[i] = swap_add16(v[i], v[4 + i]);
  {
    const bool hi = (lane >> 3) & 1;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const float mine = hi ? v[2 + i] : v[i];
      const float give = hi ? v[i] : v[2 + i];
      v[i] = mine + dpp_mirror16(give);
    }
  }
  {
    const bool hi = (lane >> 2) & 1;
    const float mine = hi ? v[1] : v[0];
    const float give = hi ? v[0] : v[1];
    v[0] = mine + dpp_mirror8(give);
  }
  float s = v[0];
  s += dpp_quad_xor2(s);
  s += dpp_quad_xor1(s);
  return s;
}

// 6-slot variant (no zero padding).  Halves 0-31 / 32-63 first take slots
// 0-2 / 3-5; then lane bit 4 splits the first two of those and the third is
// summed over bit 4 as well; lane bit 3 picks between them.  Lane L ends with
// the total of slot 3 * ((L>>5)&1) + ((L>>3)&1 ? 2 : (L>>4)&1); lanes with
// (L & 7) == 0, except those with bits 3 and 4 both set, hold each slot once.
__device__ __forceinline__ float wave_reduce6(const float* v, int lane) {
  const float a0 = swap_add32(v[0], v[3]);
  const float a1 = swap_add32(v[1], v[4]);
  const float a2 = swap_add32(v[2], v[5]);
  const float b0 = swap_add16(a0, a1);
  const float b1 = swap_add16(a2, a2);
  const bool hi = (lane >> 3) & 1;
  const float mine = hi ? b1 : b0;
  const float give = hi ? b0 : b1;
  float s = mine + dpp_mirror16(give);
  s += dpp_mirror8(s);
  s += dpp_quad_xor2(s);
  s += dpp_quad_xor1(s);
  return s;
}

__device__ __forceinline__ int reduce6_slot(int lane) {
  return 3 * ((lane >> 5) & 1) + (((lane >> 3) & 1) ? 2 : ((lane >> 4) & 1));
}

// Lanes that write wave_reduce6's slots, each slot once.
__device__ __forceinline__ bool reduce6_writer(int lane) { return (lane & 7) == 0 && (lane & 24) != 24; }

__device__ __forceinline__ int reduce16_slot(int lane) {
  return ((lane >> 5) & 1) * 8 + ((lane >> 4) & 1) * 4 + ((lane >> 3) & 1) * 2 + ((lane >> 2) & 1);
}

// ---------------------------------------------------------------------------
// Staging and culling helpers shared by the single-render and batched kernels
// ---------------------------------------------------------------------------
// SH coefficients are [P][M][3] floats (192 B per Gaussian at degree 3).  A
// lane reading its own Gaussian touches a different cache line per dword, so
// each wave first copies its 64 Gaussians' block with lane-contiguous loads
// into LDS (odd row stride: conflict-free column reads).
constexpr int SH_MAX_FLOATS = 48;
constexpr int SH_LDS_STRIDE = 49;

__device__ __forceinline__ const float* stage_sh(const float* __restrict__ shs, int P, int M, int wave_first,
                                                 int lane, float* s_wave) {
  const int rowf = M * 3;
  int nvalid = P - wave_first;
  nvalid = nvalid < 0 ? 0 : (nvalid > WAVE ? WAVE : nvalid);
  const int nfl = nvalid * rowf;
  const float* src = shs + (size_t)wave_first * rowf;
  int g = lane / rowf, r = lane - g * rowf;
  for (int f = lane; f < nfl; f += WAVE) {
    s_wave[g * SH_LDS_STRIDE + r] = src[f];
    r += WAVE;
    while (r >= rowf) { r -= rowf; ++g; }
  }
  wave_sync();
  return s_wave + lane * SH_LDS_STRIDE;
}

// Pixel bounds of the 64 pixels a wave owns inside its tile (tile_w divides 64).
struct Strip {
  float x0, x1, y0, y1;
};

// A 256-pixel tile (tile_w, tile_h multiples of 8) is split into four 8x8
// blocks, one per wave; lane l of a wave owns pixel (l % 8, l / 8) of its
// block.  Square wave footprints meet fewer splat ellipses than 16x4 rows.
__device__ __forceinline__ void tile_pixel(int tile, int tid, int grid_x, int tile_w, int tile_h, int& px,
                                           int& py) {
  const int wave = tid >> 6, lane = tid & 63, bw = tile_w >> 3;
  px = (tile % grid_x) * tile_w + (wave % bw) * 8 + (lane & 7);
  py = (tile / grid_x) * tile_h + (wave / bw) * 8 + (lane >> 3);
}

__device__ __forceinline__ Strip wave_strip(int tile, int wave, int grid_x, int tile_w, int tile_h) {
  const int bw = tile_w >> 3;
  Strip sb;
  sb.x0 = (float)((tile % grid_x) * tile_w + (wave % bw) * 8);
  sb.x1 = sb.x0 + 7.0f;
  sb.y0 = (float)((tile / grid_x) * tile_h + (wave / bw) * 8);
  sb.y1 = sb.y0 + 7.0f;
  return sb;
}

// Can this splat reach alpha >= 1/255 at any pixel of a rectangle of pixel
// centres?  Conservative: the set {power >= -L'} with L' slightly above
// ln(255 o) contains every pixel where the blend's own float test passes
// (det_exp_unclamped is within 2 ulp of exp and never exceeds 1 for
// power <= 0), and its bounding box is |dx| <= sqrt(2 L' Sxx),
// |dy| <= sqrt(2 L' Syy) with
// S = inverse conic.  The test is split into the per-splat part (cull_prep)
// and the per-rectangle part (cull_rect).
struct CullPrep {
  float L;        // L' (the set is q <= 2 L')
  float ex, ey;   // bounding-box half extents
  bool any;       // false: the splat reaches no pixel (opacity below 1/255)
  bool open;      // conic not positive definite: every rectangle is accepted
};

// Per-splat half of the test.  Every field is set on every path (a value, not
// an out-parameter, so the struct stays in registers).
__device__ __forceinline__ CullPrep cull_prep(float4 r0, float4 r1) {
  const float a = r0.z, b = r0.w, c = r1.x, o = r1.y;
  CullPrep p;
  p.any = o * 255.0f >= 0.999f;   // else alpha <= o < 1/255 everywhere
  const float L = __logf(255.0f * o);
  p.L = (L > 0.0f ? L : 0.0f) * 1.0001f + 1e-4f;
  const float det = a * c - b * b;
  p.open = !(det > 0.0f) || !(a > 0.0f) || !(c > 0.0f);
  // hardware reciprocal and square root (about 1 ulp each): far inside the
  // 1e-4 relative and 1e-3 absolute margins of the extents
  const float inv_det = __builtin_amdgcn_rcpf(det);
  p.ex = __builtin_amdgcn_sqrtf(2.0f * p.L * c * inv_det) * 1.0001f + 1e-3f;
  p.ey = __builtin_amdgcn_sqrtf(2.0f * p.L * a * inv_det) * 1.0001f + 1e-3f;
  return p;
}

// Rectangle test after cull_prep: bounding box, then the exact test.  q = a dx^2
// + 2 b dx dy + c dy^2 grows along every ray from the mean, so when the mean
// lies outside the rectangle the minimum of q over it lies on a side facing the
// mean (at most one vertical and one horizontal); on a side the minimum is at
// the clamped stationary point.  Accept when min q <= 2 L' plus a relative
// margin far above float32 rounding of the terms (the blend evaluates the same
// terms).  A splat that reaches no pixel is rejected; one whose conic is not
// positive definite is accepted.
// Both sides are evaluated and the outcome selected, without branches: the
// lanes of a wave test different splats and would take every branch anyway.
__device__ __forceinline__ bool cull_rect(float4 r0, float4 r1, const CullPrep& p, const Strip& sb) {
  const float a = r0.z, b = r0.w, c = r1.x;
  const float mx = r0.x, my = r0.y;
  // bitwise operators: no short-circuit branches
  const bool box = (mx + p.ex >= sb.x0) & (mx - p.ex <= sb.x1) & (my + p.ey >= sb.y0) & (my - p.ey <= sb.y1);
  const bool outx = (mx < sb.x0) | (mx > sb.x1), outy = (my < sb.y0) | (my > sb.y1);
  // vertical side facing the mean
  const float vx = (mx < sb.x0 ? sb.x0 : sb.x1) - mx;
  const float vy = fclamp(-b * vx * __builtin_amdgcn_rcpf(c), sb.y0 - my, sb.y1 - my);
  const float vq = a * vx * vx + 2.0f * b * vx * vy + c * vy * vy;
  const float vm = a * vx * vx + fabsf(2.0f * b * vx * vy) + c * vy * vy;
  // horizontal side facing the mean
  const float hy = (my < sb.y0 ? sb.y0 : sb.y1) - my;
  const float hx = fclamp(-b * hy * __builtin_amdgcn_rcpf(a), sb.x0 - mx, sb.x1 - mx);
  const float hq = a * hx * hx + 2.0f * b * hx * hy + c * hy * hy;
  const float hm = a * hx * hx + fabsf(2.0f * b * hx * hy) + c * hy * hy;
  const bool hit = (!outx & !outy) | (outx & (vq <= 2.0f * p.L + 1e-4f * vm + 1e-3f)) |
                   (outy & (hq <= 2.0f * p.L + 1e-4f * hm + 1e-3f));
  return p.any & (p.open | (box & hit));
}

// Band form of the same test, cheaper per rectangle once its per-splat
// parameters are set (the packed emission sets them once per Gaussian and
// tests every tile of its rect).  Over a band dy in [d0, d1] of rows, the
// ellipse q(dx, dy) <= 2 L' spans exactly dx in [xl, xr] with
//   xr = max of -p dy + sqrt(u - w dy^2),  xl = min of -p dy - sqrt(u - w dy^2)
// over the band's rows inside the ellipse, p = b / a, u = 2 L' / a,
// w = det / a^2.  The first is concave with its peak at dy = kr (the ellipse's
// rightmost point, kr = -b sqrt(2 L' / (c det))), the second convex with its
// minimum at -kr, so each is read at its stationary point clamped to the
// band.  A rectangle [x0, x1] x [y0, y1] of pixel centres is reached iff the
// band meets the ellipse's rows (|dy| <= ey, the bounding-box extent with its
// margins) and [x0, x1] meets [xl, xr] widened by mg = 1e-3 + 1e-3 sqrt(u):
// above the float32 error of the square root near the ellipse's top and
// bottom (at most about 4.5e-4 sqrt(u)) and, in q, below 0.04 (|dq/dx| =
// 2 a sqrt(u - w dy^2) on the boundary, a <= 1 / 0.3 after the blur).
// q2 = {p, u, w, kr}, q3 = {ey, mg, kind}; kind 0: reaches no pixel, 1: test,
// 2: conic not positive definite (always reached).
struct BandRec {
  float4 q2, q3;
};

__device__ __forceinline__ BandRec band_prep(float4 r0, float4 r1) {
  const CullPrep cp = cull_prep(r0, r1);
  const float a = r0.z, b = r0.w, c = r1.x;
  const float T = 2.0f * cp.L;
  // det with the product b^2 split exactly (fma gives its rounding error)
  const float bb = b * b;
  const float det = __builtin_fmaf(a, c, -bb) - __builtin_fmaf(b, b, -bb);
  float kind = !cp.any ? 0.0f : ((cp.open || !(det > 0.0f)) ? 2.0f : 1.0f);
  BandRec o;
  if (kind == 1.0f) {
    // hardware reciprocal and square root (about 1 ulp): their errors move
    // the bounds by far less than mg
    const float ia = __builtin_amdgcn_rcpf(a);
    const float u = T * ia;
    float kr = -b * __builtin_amdgcn_sqrtf(T * __builtin_amdgcn_rcpf(c * det));
    if (!(kr == kr)) kr = 0.0f;   // b = 0 with an overflowing root: the peak is at dy = 0
    o.q2 = make_float4(b * ia, u, (det * ia) * ia, kr);
    o.q3 = make_float4(cp.ey, 1e-3f + 1e-3f * __builtin_amdgcn_sqrtf(u), kind, 0.0f);
  } else {
    o.q2 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    o.q3 = make_float4(0.0f, 0.0f, kind, 0.0f);
  }
  return o;
}

// The band test on a rectangle of pixel centres (mx, my: the splat's mean).
__device__ __forceinline__ bool band_reach(float mx, float my, float4 q2, float ey, float mg, const Strip& sb) {
  const float d0 = sb.y0 - my, d1 = sb.y1 - my;
  const float yr = fminf(fmaxf(q2.w, d0), d1), yl = fminf(fmaxf(-q2.w, d0), d1);
  const float sr = __builtin_amdgcn_sqrtf(fmaxf(q2.y - q2.z * (yr * yr), 0.0f));
  const float sl = __builtin_amdgcn_sqrtf(fmaxf(q2.y - q2.z * (yl * yl), 0.0f));
  const float xr = sr - q2.x * yr, xl = -sl - q2.x * yl;
  return (d1 >= -ey) & (d0 <= ey) & (sb.x0 - mx <= xr + mg) & (sb.x1 - mx >= xl - mg);
}

// The blend's per-wave test on its 8x8 block; neg_l receives -L', below which
// a pixel can skip the exp (alpha < 1/255).
__device__ __forceinline__ bool splat_may_touch(float4 r0, float4 r1, const Strip& sb, float& neg_l) {
  const CullPrep p = cull_prep(r0, r1);
  neg_l = p.any ? -p.L : 3.0e38f;
  return cull_rect(r0, r1, p, sb);
}

}  // namespace gsevt
