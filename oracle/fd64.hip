// FP64 finite-difference engine for the rasterizer's pose and velocity
// gradients, on the GPU.
//
// TEST INFRASTRUCTURE ONLY. Nothing in the product links, loads or calls this
// file; tests/ use it as the checker of the hand-written adjoint at sizes the
// CPU oracle (gsevt_oracle.cpp) cannot difference in reasonable time.
//
// It restates, in double precision, the forward of the published 3D Gaussian
// Splatting rasterizer as gaussian_splatting/gaussian_renderer/__init__.py:
// 313-384 calls it (the rasterizer source itself, submodule
// diff-gaussian-rasterization-vwRT at .gitmodules:1-3, is absent), rendering
// the grey intensity I = 0.2989 R + 0.5870 G + 0.1140 B of
// utils/render_camera/frame.py:39-44:
//   preprocess : cull (view z <= 0.2), EWA covariance with the 1.3 tan(fov/2)
//                clamp of the Jacobian point, +0.3 px^2, conic, radius
//                ceil(3 sqrt(lambda_max)), tile rect, SH (degree <= 3) -> RGB
//                + 0.5 clamped at 0
//   blend      : front to back, alpha = min(0.99, o exp(power)), skip
//                power > 0 or alpha < 1/255, stop before T < 1e-4
// A central difference of a piecewise function straddles its thresholds, so
// the engine evaluates it in two modes:
//   record : at the base point, every discrete decision is taken and stored -
//            per Gaussian the cull, the Jacobian clamps (and their side) and
//            the three SH clamps; per (fragment, pixel) whether the fragment
//            is blended into the pixel and whether its alpha is clamped at
//            0.99 (two bits, written by wave ballots);
//   frozen : the same arithmetic with every decision read back, so the
//            function is smooth in the pose around the base point and its
//            central difference converges to the derivative the adjoint
//            computes (which differentiates the branches taken).
// At the base point both modes give the same bits (tests check it), and the
// record mode agrees with the CPU FP64 oracle (tests check that too).

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

namespace {

constexpr double SH_C0 = 0.28209479177387814;
constexpr double SH_C1 = 0.4886025119029199;
constexpr double SH_C2_0 = 1.0925484305920792, SH_C2_1 = -1.0925484305920792, SH_C2_2 = 0.31539156525252005,
                 SH_C2_3 = -1.0925484305920792, SH_C2_4 = 0.5462742152960396;
constexpr double SH_C3_0 = -0.5900435899266435, SH_C3_1 = 2.890611442640554, SH_C3_2 = -0.4570457994644658,
                 SH_C3_3 = 0.3731763325901154, SH_C3_4 = -0.4570457994644658, SH_C3_5 = 1.445305721320277,
                 SH_C3_6 = -0.5900435899266435;

// Per-Gaussian decision bits.
constexpr uint32_t F_VISIBLE = 1u, F_CLAMPX = 2u, F_CLAMPX_POS = 4u, F_CLAMPY = 8u, F_CLAMPY_POS = 16u;
constexpr uint32_t F_SHCLAMP0 = 32u;   // + c for channel c

struct Cam64 {
  double V[3][4];    // world -> camera [R | t], row-major
  double Pf[4][4];   // full projection (projection * view), row-major, acting on column vectors
  double campos[3];
  double tanfovx, tanfovy;
  int W, H, tile_w, tile_h;
};

struct Pre64 {
  const double* means;   // [P][3]
  const double* cov6;    // [P][6] 3D covariance (00, 01, 02, 11, 12, 22)
  const double* opac;    // [P]
  const double* shs;     // [P][M][3]
  int P, D, M;
};

struct Out64 {
  double* xy;        // [P][2] pixel-space mean
  double* conic;     // [P][4] a, b, c, opacity
  double* inten;     // [P] grey intensity
  double* depth;     // [P]
  int* rect;         // [P][4] tile rect (record mode)
  uint32_t* flags;   // [P]
};

__device__ void sh_eval(int D, double x, double y, double z, const double* sh, double rgb[3]) {
  double b[16];
  b[0] = SH_C0;
  int K = 1;
  if (D >= 1) {
    b[1] = -SH_C1 * y; b[2] = SH_C1 * z; b[3] = -SH_C1 * x;
    K = 4;
  }
  if (D >= 2) {
    const double xx = x * x, yy = y * y, zz = z * z;
    b[4] = SH_C2_0 * x * y;
    b[5] = SH_C2_1 * y * z;
    b[6] = SH_C2_2 * (2.0 * zz - xx - yy);
    b[7] = SH_C2_3 * x * z;
    b[8] = SH_C2_4 * (xx - yy);
    K = 9;
    if (D >= 3) {
      b[9] = SH_C3_0 * y * (3.0 * xx - yy);
      b[10] = SH_C3_1 * x * y * z;
      b[11] = SH_C3_2 * y * (4.0 * zz - xx - yy);
      b[12] = SH_C3_3 * z * (2.0 * zz - 3.0 * xx - 3.0 * yy);
      b[13] = SH_C3_4 * x * (4.0 * zz - xx - yy);
      b[14] = SH_C3_5 * z * (xx - yy);
      b[15] = SH_C3_6 * x * (xx - 3.0 * yy);
      K = 16;
    }
  }
  for (int c = 0; c < 3; ++c) {
    double acc = 0.0;
    for (int k = 0; k < K; ++k) acc += b[k] * sh[3 * k + c];
    rgb[c] = acc + 0.5;
  }
}

__global__ void k_fd_preprocess(Pre64 m, Cam64 cam, int record, Out64 o) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.P) return;
  uint32_t fl = record ? 0u : o.flags[i];
  o.xy[2 * i] = o.xy[2 * i + 1] = 0.0;
  for (int k = 0; k < 4; ++k) o.conic[4 * i + k] = 0.0;
  o.inten[i] = 0.0;
  o.depth[i] = 0.0;
  if (record)
    for (int k = 0; k < 4; ++k) o.rect[4 * i + k] = 0;
  const double* p = m.means + 3 * i;
  double pv[3];
  for (int r = 0; r < 3; ++r) pv[r] = cam.V[r][0] * p[0] + cam.V[r][1] * p[1] + cam.V[r][2] * p[2] + cam.V[r][3];
  if (record) {
    if (!(pv[2] > 0.2)) { o.flags[i] = 0u; return; }
  } else if (!(fl & F_VISIBLE)) {
    return;
  }
  const double hx = cam.Pf[0][0] * p[0] + cam.Pf[0][1] * p[1] + cam.Pf[0][2] * p[2] + cam.Pf[0][3];
  const double hy = cam.Pf[1][0] * p[0] + cam.Pf[1][1] * p[1] + cam.Pf[1][2] * p[2] + cam.Pf[1][3];
  const double hw = cam.Pf[3][0] * p[0] + cam.Pf[3][1] * p[1] + cam.Pf[3][2] * p[2] + cam.Pf[3][3];
  const double pw = 1.0 / (hw + 1e-7);
  const double nx = hx * pw, ny = hy * pw;
  // Jacobian of the perspective map at the clamped view-space point
  const double fx = cam.W / (2.0 * cam.tanfovx), fy = cam.H / (2.0 * cam.tanfovy);
  const double limx = 1.3 * cam.tanfovx, limy = 1.3 * cam.tanfovy;
  const double tz = pv[2];
  double u = pv[0] / tz, v = pv[1] / tz;
  if (record) {
    if (u < -limx || u > limx) fl |= F_CLAMPX | (u > limx ? F_CLAMPX_POS : 0u);
    if (v < -limy || v > limy) fl |= F_CLAMPY | (v > limy ? F_CLAMPY_POS : 0u);
  }
  if (fl & F_CLAMPX) u = (fl & F_CLAMPX_POS) ? limx : -limx;
  if (fl & F_CLAMPY) v = (fl & F_CLAMPY_POS) ? limy : -limy;
  const double tx = u * tz, ty = v * tz;
  const double J00 = fx / tz, J02 = -(fx * tx) / (tz * tz);
  const double J11 = fy / tz, J12 = -(fy * ty) / (tz * tz);
  double A[2][3];
  for (int c = 0; c < 3; ++c) {
    A[0][c] = J00 * cam.V[0][c] + J02 * cam.V[2][c];
    A[1][c] = J11 * cam.V[1][c] + J12 * cam.V[2][c];
  }
  const double* s = m.cov6 + 6 * i;
  const double S[3][3] = {{s[0], s[1], s[2]}, {s[1], s[3], s[4]}, {s[2], s[4], s[5]}};
  double B[2][3];
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 3; ++c) B[r][c] = A[r][0] * S[0][c] + A[r][1] * S[1][c] + A[r][2] * S[2][c];
  const double c00 = B[0][0] * A[0][0] + B[0][1] * A[0][1] + B[0][2] * A[0][2] + 0.3;
  const double c01 = B[0][0] * A[1][0] + B[0][1] * A[1][1] + B[0][2] * A[1][2];
  const double c11 = B[1][0] * A[1][0] + B[1][1] * A[1][1] + B[1][2] * A[1][2] + 0.3;
  const double det = c00 * c11 - c01 * c01;
  if (record && det == 0.0) { o.flags[i] = 0u; return; }
  const double px = ((nx + 1.0) * cam.W - 1.0) * 0.5;
  const double py = ((ny + 1.0) * cam.H - 1.0) * 0.5;
  if (record) {
    const double mid = 0.5 * (c00 + c11);
    const double sq = sqrt(fmax(0.1, mid * mid - det));
    double rad = ceil(3.0 * sqrt(fmax(mid + sq, mid - sq)));
    if (!(rad < 1073741824.0)) rad = 1073741824.0;
    const int gx = (cam.W + cam.tile_w - 1) / cam.tile_w, gy = (cam.H + cam.tile_h - 1) / cam.tile_h;
    const double tw = cam.tile_w, th = cam.tile_h;
    const int x0 = (int)fmin((double)gx, fmax(0.0, (px - rad) / tw));
    const int y0 = (int)fmin((double)gy, fmax(0.0, (py - rad) / th));
    const int x1 = (int)fmin((double)gx, fmax(0.0, (px + rad + tw - 1.0) / tw));
    const int y1 = (int)fmin((double)gy, fmax(0.0, (py + rad + th - 1.0) / th));
    if ((x1 - x0) * (y1 - y0) == 0) { o.flags[i] = 0u; return; }
    o.rect[4 * i] = x0; o.rect[4 * i + 1] = y0; o.rect[4 * i + 2] = x1; o.rect[4 * i + 3] = y1;
  }
  const double inv = 1.0 / det;
  o.xy[2 * i] = px;
  o.xy[2 * i + 1] = py;
  o.conic[4 * i] = c11 * inv;
  o.conic[4 * i + 1] = -c01 * inv;
  o.conic[4 * i + 2] = c00 * inv;
  o.conic[4 * i + 3] = m.opac[i];
  o.depth[i] = tz;
  // colour from the view direction (mean - camera centre)
  const double d0 = p[0] - cam.campos[0], d1 = p[1] - cam.campos[1], d2 = p[2] - cam.campos[2];
  const double dl = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
  double rgb[3];
  sh_eval(m.D, d0 / dl, d1 / dl, d2 / dl, m.shs + (size_t)i * m.M * 3, rgb);
  for (int c = 0; c < 3; ++c) {
    if (record && rgb[c] < 0.0) fl |= F_SHCLAMP0 << c;
    if (fl & (F_SHCLAMP0 << c)) rgb[c] = 0.0;
  }
  o.inten[i] = 0.2989 * rgb[0] + 0.5870 * rgb[1] + 0.1140 * rgb[2];
  if (record) o.flags[i] = fl | F_VISIBLE;
}

// One workgroup of 256 threads per tile, one pixel per thread (four 64-lane
// waves).  bits: per fragment of the sorted list, 4 words (one per wave) of
// "blended" and then 4 words of "alpha clamped at 0.99".
__global__ __launch_bounds__(256) void k_fd_blend(int W, int H, int tile_w, int tile_h, const int* ranges,
                                                  const int* vals, const double* xy, const double* conic,
                                                  const double* inten, double bg, int record,
                                                  unsigned long long* bits, double* out) {
  const int gx = (W + tile_w - 1) / tile_w;
  const int tile = blockIdx.x;
  const int t = threadIdx.x;
  const int wave = t >> 6, lane = t & 63;
  const int px = (tile % gx) * tile_w + t % tile_w;
  const int py = (tile / gx) * tile_h + t / tile_w;
  const bool inside = px < W && py < H;
  const int s = ranges[2 * tile], e = ranges[2 * tile + 1];
  double T = 1.0, C = 0.0;
  bool done = !inside;
  for (int j = s; j < e; ++j) {
    const int g = vals[j];
    unsigned long long* wb = bits + (size_t)j * 8;
    bool blend = false, clamp = false;
    if (record) {
      double alpha = 0.0;
      if (!done) {
        const double dx = xy[2 * g] - px, dy = xy[2 * g + 1] - py;
        const double* q = conic + 4 * g;
        const double power = -0.5 * (q[0] * dx * dx + q[2] * dy * dy) - q[1] * dx * dy;
        if (power <= 0.0) {
          const double a = q[3] * exp(power);
          clamp = a > 0.99;
          alpha = clamp ? 0.99 : a;
          if (alpha >= 1.0 / 255.0) {
            if (T * (1.0 - alpha) < 1e-4) done = true;
            else blend = true;
          }
        }
      }
      clamp = clamp && blend;
      const unsigned long long bb = __ballot(blend), bc = __ballot(clamp);
      if (lane == 0) { wb[wave] = bb; wb[4 + wave] = bc; }
      if (blend) {
        C += inten[g] * alpha * T;
        T *= 1.0 - alpha;
      }
      if (__all(done)) break;
    } else {
      const unsigned long long bb = wb[wave];
      if (bb == 0ull) continue;
      blend = (bb >> lane) & 1ull;
      clamp = (wb[4 + wave] >> lane) & 1ull;
      if (blend) {
        double alpha;
        if (clamp) {
          alpha = 0.99;
        } else {
          const double dx = xy[2 * g] - px, dy = xy[2 * g + 1] - py;
          const double* q = conic + 4 * g;
          const double power = -0.5 * (q[0] * dx * dx + q[2] * dy * dy) - q[1] * dx * dy;
          alpha = q[3] * exp(power);
        }
        C += inten[g] * alpha * T;
        T *= 1.0 - alpha;
      }
    }
  }
  if (inside) out[(size_t)py * W + px] = C + T * bg;
}

// Deterministic float exp of the rasterizer (Cody-Waite reduction by ln 2,
// degree-7 Taylor polynomial with fmaf), bit for bit the function the
// FP32 forward blends with, so that decisions taken here match its decisions.
__device__ float det_expf(float x) {
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

// Decisions of an FP32 forward blend, from its own binning (ranges, vals),
// splat records ({x, y, conic a, b}, {conic c, opacity, ...}, S float4 per
// Gaussian) and per-pixel count of fragments up to the last blended one: a
// fragment before that count is blended iff power <= 0 and
// min(0.99, o exp(power)) >= 1/255, evaluated with the forward's float
// operations; its alpha is clamped iff o exp(power) > 0.99.  Same bit layout
// as k_fd_blend's record mode.
__global__ __launch_bounds__(256) void k_fd_record32(int W, int H, int tile_w, int tile_h, const uint32_t* ranges,
                                                     const uint32_t* vals, uint32_t gid_mask, const float4* splats,
                                                     int S, const int* n_contrib, unsigned long long* bits) {
  const int gx = (W + tile_w - 1) / tile_w;
  const int tile = blockIdx.x;
  const int t = threadIdx.x;
  const int wave = t >> 6, lane = t & 63;
  const int px = (tile % gx) * tile_w + t % tile_w;
  const int py = (tile / gx) * tile_h + t / tile_w;
  const bool inside = px < W && py < H;
  const uint32_t s = ranges[2 * tile], e = ranges[2 * tile + 1];
  const int last = inside ? n_contrib[(size_t)py * W + px] : 0;
  int wmax = last;
  for (int o = 32; o >= 1; o >>= 1) { const int y = __shfl_xor(wmax, o, 64); wmax = y > wmax ? y : wmax; }
  const float pxf = (float)px, pyf = (float)py;
  for (uint32_t j = s; j < e && (int)(j - s) < wmax; ++j) {
    const uint32_t g = vals[j] & gid_mask;
    const float4 r0 = splats[(size_t)g * S], r1 = splats[(size_t)g * S + 1];
    bool blend = false, clamp = false;
    if ((int)(j - s) < last) {
      const float dx = r0.x - pxf, dy = r0.y - pyf;
      const float power = -0.5f * ((r0.z * dx) * dx + (r1.x * dy) * dy) - (r0.w * dx) * dy;
      if (!(power > 0.0f)) {
        const float oG = r1.y * det_expf(power);
        const float alpha = oG < 0.99f ? oG : 0.99f;
        blend = !(alpha < 1.0f / 255.0f);
        clamp = blend && oG > 0.99f;
      }
    }
    const unsigned long long bb = __ballot(blend), bc = __ballot(clamp);
    if (lane == 0) { bits[(size_t)j * 8 + wave] = bb; bits[(size_t)j * 8 + 4 + wave] = bc; }
  }
}

thread_local char g_err[256] = "";

int check(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    snprintf(g_err, sizeof(g_err), "%s: %s", what, hipGetErrorString(e));
    return (int)e;
  }
  return 0;
}

}  // namespace

extern "C" {

const char* fd64_last_error(void) { return g_err; }

// view12: world -> camera [R | t] row-major (12 doubles, host); proj16: the
// full projection (projection * view) row-major (16 doubles, host); campos
// (3 doubles, host).  Device arrays: means [P][3], cov6 [P][6], opac [P],
// shs [P][M][3] (double); xy [P][2], conic [P][4], inten [P], depth [P],
// rect [P][4] (written in record mode), flags [P] (written in record mode,
// read in frozen mode).
int fd64_preprocess(int P, int D, int M, const double* means, const double* cov6, const double* opac,
                    const double* shs, const double* view12, const double* proj16, const double* campos, int W,
                    int H, double tanfovx, double tanfovy, int tile_w, int tile_h, int record, double* xy,
                    double* conic, double* inten, double* depth, int* rect, uint32_t* flags, void* stream) {
  if (P <= 0 || D < 0 || D > 3 || M < (D + 1) * (D + 1) || W <= 0 || H <= 0 || tile_w * tile_h != 256) {
    snprintf(g_err, sizeof(g_err), "fd64_preprocess: bad sizes (P %d, D %d, M %d, %dx%d, tile %dx%d)", P, D, M, W,
             H, tile_w, tile_h);
    return -1;
  }
  Cam64 c;
  for (int r = 0; r < 3; ++r)
    for (int k = 0; k < 4; ++k) c.V[r][k] = view12[4 * r + k];
  for (int r = 0; r < 4; ++r)
    for (int k = 0; k < 4; ++k) c.Pf[r][k] = proj16[4 * r + k];
  for (int k = 0; k < 3; ++k) c.campos[k] = campos[k];
  c.tanfovx = tanfovx; c.tanfovy = tanfovy; c.W = W; c.H = H; c.tile_w = tile_w; c.tile_h = tile_h;
  Pre64 m{means, cov6, opac, shs, P, D, M};
  Out64 o{xy, conic, inten, depth, rect, flags};
  hipLaunchKernelGGL(k_fd_preprocess, dim3((P + 255) / 256), dim3(256), 0, (hipStream_t)stream, m, c, record, o);
  return check("fd64_preprocess");
}

// ranges [n_tiles][2] (first, one past last) and vals [N_f] (Gaussian index
// per fragment, tile-major, front to back) on the device; bits [N_f][8]
// uint64 (zeroed by the caller before a record pass); out [H][W].
int fd64_blend(int W, int H, int tile_w, int tile_h, const int* ranges, const int* vals, const double* xy,
               const double* conic, const double* inten, double bg, int record, unsigned long long* bits,
               double* out, void* stream) {
  if (W <= 0 || H <= 0 || tile_w * tile_h != 256) {
    snprintf(g_err, sizeof(g_err), "fd64_blend: bad sizes");
    return -1;
  }
  const int gx = (W + tile_w - 1) / tile_w, gy = (H + tile_h - 1) / tile_h;
  hipLaunchKernelGGL(k_fd_blend, dim3(gx * gy), dim3(256), 0, (hipStream_t)stream, W, H, tile_w, tile_h, ranges,
                     vals, xy, conic, inten, bg, record, bits, out);
  return check("fd64_blend");
}

// Frozen-decision bits from an FP32 forward (k_fd_record32): ranges
// [n_tiles][2] and vals [N_f] of its binning (the Gaussian index is
// vals & gid_mask), its splat records (S float4 per Gaussian) and n_contrib
// [H][W]; bits [N_f][8] (zeroed by the caller).
int fd64_record32(int W, int H, int tile_w, int tile_h, const uint32_t* ranges, const uint32_t* vals,
                  uint32_t gid_mask, const float* splats, int S, const int* n_contrib, unsigned long long* bits,
                  void* stream) {
  if (W <= 0 || H <= 0 || tile_w * tile_h != 256 || S < 2) {
    snprintf(g_err, sizeof(g_err), "fd64_record32: bad sizes");
    return -1;
  }
  const int gx = (W + tile_w - 1) / tile_w, gy = (H + tile_h - 1) / tile_h;
  hipLaunchKernelGGL(k_fd_record32, dim3(gx * gy), dim3(256), 0, (hipStream_t)stream, W, H, tile_w, tile_h, ranges,
                     vals, gid_mask, (const float4*)splats, S, n_contrib, bits);
  return check("fd64_record32");
}

}  // extern "C"
