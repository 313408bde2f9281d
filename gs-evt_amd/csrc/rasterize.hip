// gfx950 kernels of the differentiable Gaussian rasterizer.
//
//   preprocess        one thread per Gaussian: cull, EWA covariance, conic,
//                     radius, tile rect, SH colour (or grey intensity)
//   duplicate         one thread per visible Gaussian: tile|depth keys
//   tile_ranges       one thread per fragment
//   blend_forward     one 256-thread workgroup per tile; each wave owns 64
//                     pixels and streams the tile's fragments through its own
//                     LDS slice in 64-entry batches, leaving the loop as soon
//                     as a ballot shows every lane saturated
//   blend_backward    same decomposition in reverse, transmittance recovered
//                     by division; per-fragment gradients are reduced across
//                     the wave with a butterfly reduce-scatter and added once
//                     per (wave, fragment)
//   preprocess_bwd    one thread per Gaussian: chain to view space, reduce
//                     the 18 pose partial sums per workgroup
//   pose_finalize     one wave: deterministic FP64 sum of the partials and
//                     the se(3) adjoint / left-Jacobian maps to the 12
//                     tracked parameters
#include "gsevt_api.h"
#include "gsevt_device.h"
#include "blend_impl.h"

using namespace gsevt;

namespace {

struct CamArgs {
  const float* view;
  const float* proj;
  const float* projraw;
  const float* campos;
  float tanfovx, tanfovy;
  int W, H, tile_w, tile_h, grid_x, grid_y;
};

__device__ __forceinline__ CamDev load_cam(const CamArgs& a) {
  CamDev c;
#pragma unroll
  for (int k = 0; k < 16; ++k) c.view[k] = a.view[k];
#pragma unroll
  for (int k = 0; k < 16; ++k) c.proj[k] = a.proj[k];
  if (a.projraw) {
#pragma unroll
    for (int k = 0; k < 16; ++k) c.projraw[k] = a.projraw[k];
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) c.campos[k] = a.campos[k];
  c.tanfovx = a.tanfovx; c.tanfovy = a.tanfovy;
  c.W = a.W; c.H = a.H; c.tile_w = a.tile_w; c.tile_h = a.tile_h; c.grid_x = a.grid_x; c.grid_y = a.grid_y;
  return c;
}

struct MapArgs {
  int P, D, M;
  const float* means;
  const float* scales;
  float mod;
  const float* rots;
  const float* opac;
  const float* shs;
  const float* colors_pre;
  const float* cov_pre;
};

__device__ __forceinline__ void gaussian_sigma(const MapArgs& m, int i, float* sig) {
  if (m.cov_pre) {
#pragma unroll
    for (int k = 0; k < 6; ++k) sig[k] = m.cov_pre[6 * i + k];
  } else {
    float R[3][3];
    cov3d(m.scales + 3 * i, m.mod, m.rots + 4 * i, sig, R);
  }
}

// ---------------------------------------------------------------------------
// preprocess
// ---------------------------------------------------------------------------
// Splat record: S float4 per Gaussian (S = 3 for RGB, 2 for grey intensity)
//   rec0 = {x, y, conic_a, conic_b}
//   rec1 = {conic_c, opacity, depth, c0}
//   rec2 = {c1, c2, 0, 0}                     (RGB only)
template <int C>
__global__ __launch_bounds__(256) void k_preprocess(MapArgs m, CamArgs ca, float iw0, float iw1, float iw2,
                                                    int* __restrict__ radii, uint32_t* __restrict__ tiles,
                                                    float4* __restrict__ splats) {
  __shared__ float s_sh[4][WAVE * SH_LDS_STRIDE];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const bool stage = m.shs && m.M * 3 <= SH_MAX_FLOATS;
  const float* sh_lds = nullptr;
  if (stage) sh_lds = stage_sh(m.shs, m.P, m.M, blockIdx.x * blockDim.x + wave * WAVE, lane, s_sh[wave]);
  if (i >= m.P) return;
  constexpr int S = (C == 1) ? 2 : 3;
  const CamDev c = load_cam(ca);
  Fwd f;
  float sig[6];
  const float* mu = m.means + 3 * i;
  gaussian_sigma(m, i, sig);
  if (!project_one(mu, sig, c, f)) {
    radii[i] = 0;
    tiles[i] = 0;
    return;
  }
  float rgb[3];
  if (m.colors_pre) {
    rgb[0] = m.colors_pre[3 * i]; rgb[1] = m.colors_pre[3 * i + 1]; rgb[2] = m.colors_pre[3 * i + 2];
  } else {
    float dir[3], dlen;
    bool cl[3];
    sh_color(m.D, m.M, mu, stage ? sh_lds : m.shs + (size_t)i * m.M * 3, c, dir, dlen, rgb, cl);
  }
  radii[i] = f.radius;
  tiles[i] = (uint32_t)((f.rect[2] - f.rect[0]) * (f.rect[3] - f.rect[1]));
  float4* rec = splats + (size_t)i * S;
  rec[0] = make_float4(f.px, f.py, f.conic[0], f.conic[1]);
  if (C == 1) {
    const float inten = iw0 * rgb[0] + iw1 * rgb[1] + iw2 * rgb[2];
    rec[1] = make_float4(f.conic[2], m.opac[i], f.pv[2], inten);
  } else {
    rec[1] = make_float4(f.conic[2], m.opac[i], f.pv[2], rgb[0]);
    rec[S - 1] = make_float4(rgb[1], rgb[2], 0.0f, 0.0f);
  }
}

// ---------------------------------------------------------------------------
// duplicate + ranges
// ---------------------------------------------------------------------------
template <int S>
__global__ __launch_bounds__(256) void k_duplicate(int P, const uint32_t* __restrict__ offsets,
                                                   const int* __restrict__ radii,
                                                   const float4* __restrict__ splats, CamArgs ca,
                                                   uint64_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                                   long long cap) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  const int rad = radii[i];
  if (rad <= 0) return;
  CamDev c;
  c.tile_w = ca.tile_w; c.tile_h = ca.tile_h; c.grid_x = ca.grid_x; c.grid_y = ca.grid_y;
  const float4 r0 = splats[(size_t)i * S];
  const float4 r1 = splats[(size_t)i * S + 1];
  int rect[4];
  tile_rect(r0.x, r0.y, rad, c, rect);
  long long off = (i == 0) ? 0 : (long long)offsets[i - 1];
  const uint64_t dbits = (uint64_t)__float_as_uint(r1.z);
  for (int y = rect[1]; y < rect[3]; ++y)
    for (int x = rect[0]; x < rect[2]; ++x) {
      if (off < cap) {
        keys[off] = ((uint64_t)(y * c.grid_x + x) << 32) | dbits;
        vals[off] = (uint32_t)i;
      }
      ++off;
    }
}

// Grid-stride fill (any n; launches size the grid with fill_blocks).
__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* p, long long n, uint32_t v) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// View-space depth test of markVisible: a point is present unless its view z
// is <= 0.2 (so a NaN depth counts as present, as the comparison is false).
__global__ __launch_bounds__(256) void k_mark_visible(int P, const float* __restrict__ means,
                                                      const float* __restrict__ V, uint8_t* __restrict__ present) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  const float x = means[3 * i], y = means[3 * i + 1], z = means[3 * i + 2];
  const float vz = V[2] * x + V[6] * y + V[10] * z + V[14];
  present[i] = (vz <= 0.2f) ? 0 : 1;
}

__global__ __launch_bounds__(256) void k_tile_ranges(const uint32_t* __restrict__ d_n, long long cap,
                                                     const uint64_t* __restrict__ keys, uint2* __restrict__ ranges) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long n = (long long)*d_n;
  if (n > cap) n = cap;
  if (i >= n) return;
  const uint32_t t = (uint32_t)(keys[i] >> 32);
  if (i == 0 || (uint32_t)(keys[i - 1] >> 32) != t) ranges[t].x = (uint32_t)i;
  if (i == n - 1 || (uint32_t)(keys[i + 1] >> 32) != t) ranges[t].y = (uint32_t)(i + 1);
}

// ---------------------------------------------------------------------------
// blend forward
// ---------------------------------------------------------------------------
template <int C>
__global__ __launch_bounds__(256) void k_blend_fwd(BlendArgs a) {
  blend_fwd_tile<C, true>(a, blockIdx.x);
}

template <int C>
__global__ __launch_bounds__(256) void k_blend_bwd(BlendArgs a) {
  const int tile = blockIdx.x;
  const int tid = threadIdx.x;
  int px, py;
  tile_pixel(tile, tid, a.grid_x, a.tile_w, a.tile_h, px, py);
  const bool inside = px < a.W && py < a.H;
  const int pix = py * a.W + px;
  const size_t HW = (size_t)a.W * a.H;
  float gC[C];
#pragma unroll
  for (int c = 0; c < C; ++c) gC[c] = inside ? a.g_color[c * HW + pix] : 0.0f;
  const float gD = (inside && a.g_depth) ? a.g_depth[pix] : 0.0f;
  const float gA = (inside && a.g_opac) ? a.g_opac[pix] : 0.0f;
  blend_bwd_tile<C>(a, tile, gC, gD, gA);
}

// ---------------------------------------------------------------------------
// preprocess backward + pose partials
// ---------------------------------------------------------------------------
struct BwdArgs {
  const int* radii;
  const float* grad10;
  int C;
  float iw[3];
  float* partials;     // [gridDim.x][18]
  float* g_means;      // [P][3]   optional map gradients
  float* g_scales;     // [P][3]
  float* g_rots;       // [P][4]
  float* g_shs;        // [P][M][3]
  float* g_opac;       // [P]
  float* g_means2d;    // [P][3]
  float* g_colors;     // [P][3]
  float* g_cov;        // [P][6]
};

__global__ __launch_bounds__(256) void k_preprocess_bwd(MapArgs m, CamArgs ca, BwdArgs b) {
  __shared__ float s_part[4][18];
  __shared__ float s_sh[4][WAVE * SH_LDS_STRIDE];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const bool stage = m.shs && m.M * 3 <= SH_MAX_FLOATS;
  const float* sh_lds = nullptr;
  if (stage) sh_lds = stage_sh(m.shs, m.P, m.M, blockIdx.x * blockDim.x + wave * WAVE, lane, s_sh[wave]);
  float part[18];
#pragma unroll
  for (int k = 0; k < 18; ++k) part[k] = 0.0f;
  if (i < m.P && b.radii[i] > 0) {
    const CamDev c = load_cam(ca);
    Fwd f;
    float sig[6];
    const float* mu = m.means + 3 * i;
    gaussian_sigma(m, i, sig);
    project_one(mu, sig, c, f);
    const float* g = b.grad10 + (size_t)i * 10;
    float gq[10];
#pragma unroll
    for (int k = 0; k < 10; ++k) gq[k] = g[k];
    const float A = f.cov[0], B = f.cov[1], Cc = f.cov[2];
    const float d2 = f.det * f.det;
    const float ga = gq[2], gb = gq[3], gc = gq[4];
    const float gA = (ga * (-Cc * Cc) + gb * (B * Cc) + gc * (-B * B)) / d2;
    const float gB = (ga * (2.0f * B * Cc) + gb * (-(A * Cc + B * B)) + gc * (2.0f * A * B)) / d2;
    const float gCc = (ga * (-B * B) + gb * (A * B) + gc * (-A * A)) / d2;
    const float G2[2][2] = {{gA, 0.5f * gB}, {0.5f * gB, gCc}};
    float gM2[2][3];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int k = 0; k < 3; ++k) gM2[r][k] = 2.0f * (G2[r][0] * f.B[0][k] + G2[r][1] * f.B[1][k]);
    float Wm[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int k = 0; k < 3; ++k) Wm[r][k] = mat(c.view, r, k);
    const float gJ00 = gM2[0][0] * Wm[0][0] + gM2[0][1] * Wm[0][1] + gM2[0][2] * Wm[0][2];
    const float gJ02 = gM2[0][0] * Wm[2][0] + gM2[0][1] * Wm[2][1] + gM2[0][2] * Wm[2][2];
    const float gJ11 = gM2[1][0] * Wm[1][0] + gM2[1][1] * Wm[1][1] + gM2[1][2] * Wm[1][2];
    const float gJ12 = gM2[1][0] * Wm[2][0] + gM2[1][1] * Wm[2][1] + gM2[1][2] * Wm[2][2];
    float gW[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      gW[0][k] = f.J00 * gM2[0][k];
      gW[1][k] = f.J11 * gM2[1][k];
      gW[2][k] = f.J02 * gM2[0][k] + f.J12 * gM2[1][k];
    }
    const float fx = (float)c.W / (2.0f * c.tanfovx);
    const float fy = (float)c.H / (2.0f * c.tanfovy);
    const float z = f.pv[2], z2 = z * z;
    float gpv[3] = {0.0f, 0.0f, 0.0f};
    gpv[2] += gJ00 * (-fx / z2) + gJ11 * (-fy / z2);
    if (!f.clampx) {
      gpv[0] += gJ02 * (-fx / z2);
      gpv[2] += gJ02 * (2.0f * fx * f.pv[0] / (z2 * z));
    } else {
      gpv[2] += gJ02 * (fx * f.u / z2);
    }
    if (!f.clampy) {
      gpv[1] += gJ12 * (-fy / z2);
      gpv[2] += gJ12 * (2.0f * fy * f.pv[1] / (z2 * z));
    } else {
      gpv[2] += gJ12 * (fy * f.v / z2);
    }
    const float ghx = gq[0] * (0.5f * (float)c.W) * f.pw;
    const float ghy = gq[1] * (0.5f * (float)c.H) * f.pw;
    const float ghw = -(gq[0] * (0.5f * (float)c.W) * f.hx + gq[1] * (0.5f * (float)c.H) * f.hy) * f.pw * f.pw;
#pragma unroll
    for (int k = 0; k < 3; ++k)
      gpv[k] += ghx * mat(c.projraw, 0, k) + ghy * mat(c.projraw, 1, k) + ghw * mat(c.projraw, 3, k);
    gpv[2] += gq[9];
    float gcol[3] = {0.0f, 0.0f, 0.0f};
    if (b.C == 3) {
      gcol[0] = gq[6]; gcol[1] = gq[7]; gcol[2] = gq[8];
    } else {
      gcol[0] = b.iw[0] * gq[6]; gcol[1] = b.iw[1] * gq[6]; gcol[2] = b.iw[2] * gq[6];
    }
    float gd[3] = {0.0f, 0.0f, 0.0f};
    if (!m.colors_pre) {
      float dir[3], dlen, rgb[3];
      bool cl[3];
      const float* sh = stage ? sh_lds : m.shs + (size_t)i * m.M * 3;
      sh_color(m.D, m.M, mu, sh, c, dir, dlen, rgb, cl);
#pragma unroll
      for (int ch = 0; ch < 3; ++ch)
        if (cl[ch]) gcol[ch] = 0.0f;
      const int K = (m.D + 1) * (m.D + 1);
      float gbasis[16];
      for (int k = 0; k < K; ++k) gbasis[k] = gcol[0] * sh[3 * k] + gcol[1] * sh[3 * k + 1] + gcol[2] * sh[3 * k + 2];
      float gdir[3] = {0.0f, 0.0f, 0.0f};
      sh_basis_grad(m.D, dir[0], dir[1], dir[2], gbasis, gdir);
      const float dd = dir[0] * gdir[0] + dir[1] * gdir[1] + dir[2] * gdir[2];
#pragma unroll
      for (int k = 0; k < 3; ++k) gd[k] = (gdir[k] - dir[k] * dd) / dlen;
      if (b.g_shs) {
        float bas[16];
        sh_basis(m.D, dir[0], dir[1], dir[2], bas);
        for (int k = 0; k < K; ++k)
#pragma unroll
          for (int ch = 0; ch < 3; ++ch) b.g_shs[(size_t)i * m.M * 3 + 3 * k + ch] = bas[k] * gcol[ch];
      }
    }
    part[0] = gpv[0]; part[1] = gpv[1]; part[2] = gpv[2];
    part[3] = f.pv[1] * gpv[2] - f.pv[2] * gpv[1];
    part[4] = f.pv[2] * gpv[0] - f.pv[0] * gpv[2];
    part[5] = f.pv[0] * gpv[1] - f.pv[1] * gpv[0];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int k = 0; k < 3; ++k) part[6 + 3 * r + k] = gW[r][k];
    part[15] = -gd[0]; part[16] = -gd[1]; part[17] = -gd[2];

    if (b.g_opac) b.g_opac[i] = gq[5];
    if (b.g_colors && m.colors_pre)
#pragma unroll
      for (int ch = 0; ch < 3; ++ch) b.g_colors[3 * i + ch] = gcol[ch];
    if (b.g_means2d) {
      b.g_means2d[3 * i] = gq[0] * (0.5f * (float)c.W);
      b.g_means2d[3 * i + 1] = gq[1] * (0.5f * (float)c.H);
      b.g_means2d[3 * i + 2] = 0.0f;
    }
    if (b.g_means)
#pragma unroll
      for (int k = 0; k < 3; ++k)
        b.g_means[3 * i + k] = Wm[0][k] * gpv[0] + Wm[1][k] * gpv[1] + Wm[2][k] * gpv[2] + gd[k];
    if (b.g_scales || b.g_rots || b.g_cov) {
      float gS[3][3];
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          float s = 0.0f;
#pragma unroll
          for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int q = 0; q < 2; ++q) s += f.A[p][r] * G2[p][q] * f.A[q][k];
          gS[r][k] = s;
        }
      if (m.cov_pre) {
        if (b.g_cov) {
          float* o = b.g_cov + 6 * i;
          o[0] = gS[0][0]; o[1] = 2.0f * gS[0][1]; o[2] = 2.0f * gS[0][2];
          o[3] = gS[1][1]; o[4] = 2.0f * gS[1][2]; o[5] = gS[2][2];
        }
      } else {
        float sg[6], R[3][3];
        const float* q = m.rots + 4 * i;
        const float* s = m.scales + 3 * i;
        cov3d(s, m.mod, q, sg, R);
        const float sc[3] = {m.mod * s[0], m.mod * s[1], m.mod * s[2]};
        float gMm[3][3];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int k = 0; k < 3; ++k)
            gMm[r][k] = 2.0f * (gS[r][0] * R[0][k] * sc[k] + gS[r][1] * R[1][k] * sc[k] + gS[r][2] * R[2][k] * sc[k]);
        if (b.g_scales)
#pragma unroll
          for (int k = 0; k < 3; ++k)
            b.g_scales[3 * i + k] = (gMm[0][k] * R[0][k] + gMm[1][k] * R[1][k] + gMm[2][k] * R[2][k]) * m.mod;
        if (b.g_rots) {
          const float qr = q[0], qx = q[1], qy = q[2], qz = q[3];
          const float dRr[3][3] = {{0, -2 * qz, 2 * qy}, {2 * qz, 0, -2 * qx}, {-2 * qy, 2 * qx, 0}};
          const float dRx[3][3] = {{0, 2 * qy, 2 * qz}, {2 * qy, -4 * qx, -2 * qr}, {2 * qz, 2 * qr, -4 * qx}};
          const float dRy[3][3] = {{-4 * qy, 2 * qx, 2 * qr}, {2 * qx, 0, 2 * qz}, {-2 * qr, 2 * qz, -4 * qy}};
          const float dRz[3][3] = {{-4 * qz, -2 * qr, 2 * qx}, {2 * qr, -4 * qz, 2 * qy}, {2 * qx, 2 * qy, 0}};
          float out[4] = {0, 0, 0, 0};
#pragma unroll
          for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
              const float gR = gMm[r][k] * sc[k];
              out[0] += gR * dRr[r][k];
              out[1] += gR * dRx[r][k];
              out[2] += gR * dRy[r][k];
              out[3] += gR * dRz[r][k];
            }
#pragma unroll
          for (int k = 0; k < 4; ++k) b.g_rots[4 * i + k] = out[k];
        }
      }
    }
  }
  // workgroup reduction of the 18 pose partials
  float v16[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) v16[k] = part[k];
  const float s16 = wave_reduce16(v16, lane);
  const float s16b = wave_sum(part[16]);
  const float s17 = wave_sum(part[17]);
  if ((lane & 3) == 0) s_part[wave][reduce16_slot(lane)] = s16;
  if (lane == 0) { s_part[wave][16] = s16b; s_part[wave][17] = s17; }
  __syncthreads();
  if (tid < 18) {
    const float t = s_part[0][tid] + s_part[1][tid] + s_part[2][tid] + s_part[3][tid];
    b.partials[(size_t)blockIdx.x * 18 + tid] = t;
  }
}

// ---------------------------------------------------------------------------
// pose finalize (one wave, FP64)
// ---------------------------------------------------------------------------
__device__ void se3_left_jacobian_T_times(const double* xi, const double* g, double* out) {
  // out = J_l(xi)^T g with J_l = sum_n ad(xi)^n / (n+1)!, twist order [rho; theta]
  double ad[6][6];
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 6; ++c) ad[r][c] = 0.0;
  const double* rho = xi;
  const double* th = xi + 3;
  // [th]x in (0,0) and (3,3) blocks, [rho]x in (0,3)
  for (int blk = 0; blk < 3; ++blk) {
    const int r0 = (blk == 2) ? 3 : 0, c0 = (blk == 0) ? 0 : 3;
    const double* v = (blk == 1) ? rho : th;
    ad[r0 + 0][c0 + 1] = -v[2]; ad[r0 + 0][c0 + 2] = v[1];
    ad[r0 + 1][c0 + 0] = v[2];  ad[r0 + 1][c0 + 2] = -v[0];
    ad[r0 + 2][c0 + 0] = -v[1]; ad[r0 + 2][c0 + 1] = v[0];
  }
  // J^T g = sum_n (ad^T)^n g / (n+1)!
  double term[6], acc[6];
  for (int k = 0; k < 6; ++k) { term[k] = g[k]; acc[k] = g[k]; }
  for (int n = 1; n < 30; ++n) {
    double nt[6];
    double mag = 0.0;
    for (int c = 0; c < 6; ++c) {
      double s = 0.0;
      for (int r = 0; r < 6; ++r) s += ad[r][c] * term[r];
      nt[c] = s / (double)(n + 1);
      mag += fabs(nt[c]);
    }
    for (int k = 0; k < 6; ++k) { term[k] = nt[k]; acc[k] += nt[k]; }
    if (mag == 0.0) break;
  }
  for (int k = 0; k < 6; ++k) out[k] = acc[k];
}

__global__ __launch_bounds__(64) void k_pose_finalize(const float* __restrict__ partials, int nblocks,
                                                      const float* __restrict__ view, const float* __restrict__ vel,
                                                      double dt, const float* __restrict__ lin,
                                                      const float* __restrict__ ang, float* __restrict__ out,
                                                      int accumulate) {
  const int lane = threadIdx.x;
  double s[18];
  for (int k = 0; k < 18; ++k) s[k] = 0.0;
  for (int b = lane; b < nblocks; b += WAVE)
    for (int k = 0; k < 18; ++k) s[k] += (double)partials[(size_t)b * 18 + k];
  for (int k = 0; k < 18; ++k)
    for (int o = 32; o >= 1; o >>= 1) s[k] += __shfl_xor(s[k], o, WAVE);
  if (lane != 0) return;
  double Wm[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) Wm[r][c] = (double)view[4 * c + r];
  double g[6];
  for (int r = 0; r < 3; ++r) g[r] = s[r] - (Wm[r][0] * s[15] + Wm[r][1] * s[16] + Wm[r][2] * s[17]);
  const double* GW = s + 6;
  double f0 = 0, f1 = 0, f2 = 0;
  for (int j = 0; j < 3; ++j) {
    f0 += -GW[3 + j] * Wm[2][j] + GW[6 + j] * Wm[1][j];
    f1 += GW[j] * Wm[2][j] - GW[6 + j] * Wm[0][j];
    f2 += -GW[j] * Wm[1][j] + GW[3 + j] * Wm[0][j];
  }
  g[3] = s[3] + f0; g[4] = s[4] + f1; g[5] = s[5] + f2;
  double RE[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}, tE[3] = {0, 0, 0};
  if (vel) {
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c) RE[r][c] = (double)vel[4 * c + r];
      tE[r] = (double)vel[12 + r];
    }
  }
  const double tg[3] = {tE[1] * g[2] - tE[2] * g[1], tE[2] * g[0] - tE[0] * g[2], tE[0] * g[1] - tE[1] * g[0]};
  double res[12];
  for (int c = 0; c < 3; ++c) {
    double rho = 0, th = 0;
    for (int r = 0; r < 3; ++r) {
      rho += RE[r][c] * g[r];
      th += RE[r][c] * (g[3 + r] - tg[r]);
    }
    res[3 + c] = rho;
    res[c] = th;
  }
  double xi[6] = {0, 0, 0, 0, 0, 0};
  if (lin && ang)
    for (int k = 0; k < 3; ++k) { xi[k] = dt * (double)lin[k]; xi[3 + k] = dt * (double)ang[k]; }
  double jg[6];
  se3_left_jacobian_T_times(xi, g, jg);
  for (int k = 0; k < 3; ++k) { res[9 + k] = dt * jg[k]; res[6 + k] = dt * jg[3 + k]; }
  for (int k = 0; k < 12; ++k) out[k] = (accumulate ? out[k] : 0.0f) + (float)res[k];
}

CamArgs make_cam_args(const float* view, const float* proj, const float* projraw, const float* campos, int W,
                      int H, float tfx, float tfy, int tw, int th) {
  CamArgs c;
  c.view = view; c.proj = proj; c.projraw = projraw; c.campos = campos;
  c.tanfovx = tfx; c.tanfovy = tfy; c.W = W; c.H = H; c.tile_w = tw; c.tile_h = th;
  c.grid_x = (W + tw - 1) / tw; c.grid_y = (H + th - 1) / th;
  return c;
}

MapArgs make_map_args(int P, int D, int M, const float* means, const float* scales, float mod, const float* rots,
                      const float* opac, const float* shs, const float* colors_pre, const float* cov_pre) {
  MapArgs m;
  m.P = P; m.D = D; m.M = M; m.means = means; m.scales = scales; m.mod = mod; m.rots = rots; m.opac = opac;
  m.shs = shs; m.colors_pre = colors_pre; m.cov_pre = cov_pre;
  return m;
}

int check_map(int P, int D, int M, const float* means, const float* scales, const float* rots, const float* opac,
              const float* shs, const float* colors_pre, const float* cov_pre) {
  GSEVT_REQUIRE(P >= 0, "P must be >= 0 (got %d)", P);
  GSEVT_REQUIRE(D >= 0 && D <= 3, "sh_degree must be in [0, 3] (got %d)", D);
  GSEVT_REQUIRE(P == 0 || means, "means3D is null");
  GSEVT_REQUIRE(P == 0 || opac, "opacities is null");
  GSEVT_REQUIRE((shs != nullptr) != (colors_pre != nullptr),
                "Please provide exactly one of either SHs or precomputed colors!");
  GSEVT_REQUIRE((cov_pre != nullptr) != (scales != nullptr && rots != nullptr),
                "Please provide exactly one of either scale/rotation pair or precomputed 3D covariance!");
  GSEVT_REQUIRE(!shs || M >= (D + 1) * (D + 1), "shs has %d coefficients, degree %d needs %d", M, D,
                (D + 1) * (D + 1));
  return 0;
}

int check_tile(int W, int H, int tw, int th) {
  GSEVT_REQUIRE(W > 0 && H > 0, "image size must be positive (%dx%d)", W, H);
  GSEVT_REQUIRE(tw > 0 && th > 0 && tw * th == 256 && tw % 8 == 0 && th % 8 == 0,
                "tile must hold 256 pixels in 8x8 blocks: 16x16, 8x32 or 32x8 (got %dx%d)", tw, th);
  return 0;
}

}  // namespace

extern "C" {

int gsevt_preprocess(int P, int D, int M, const float* means3D, const float* scales, float scale_modifier,
                     const float* rotations, const float* opacities, const float* shs, const float* colors_precomp,
                     const float* cov3D_precomp, const float* viewmatrix, const float* projmatrix,
                     const float* campos, int W, int H, float tanfovx, float tanfovy, int tile_w, int tile_h,
                     int channels, const float* intensity_weights, int* radii, uint32_t* tiles_touched,
                     float* splats, void* stream) {
  if (check_map(P, D, M, means3D, scales, rotations, opacities, shs, colors_precomp, cov3D_precomp)) return -1;
  if (check_tile(W, H, tile_w, tile_h)) return -1;
  GSEVT_REQUIRE(channels == 1 || channels == 3, "channels must be 1 or 3 (got %d)", channels);
  GSEVT_REQUIRE(viewmatrix && projmatrix && campos, "camera matrices are null");
  GSEVT_REQUIRE(radii && tiles_touched && splats, "output buffers are null");
  if (P == 0) return 0;
  float iw[3] = {0.2989f, 0.5870f, 0.1140f};
  if (intensity_weights) { iw[0] = intensity_weights[0]; iw[1] = intensity_weights[1]; iw[2] = intensity_weights[2]; }
  MapArgs m = make_map_args(P, D, M, means3D, scales, scale_modifier, rotations, opacities, shs, colors_precomp,
                            cov3D_precomp);
  CamArgs c = make_cam_args(viewmatrix, projmatrix, nullptr, campos, W, H, tanfovx, tanfovy, tile_w, tile_h);
  hipStream_t s = (hipStream_t)stream;
  if (channels == 3)
    hipLaunchKernelGGL(k_preprocess<3>, dim3(div_up(P, 256)), dim3(256), 0, s, m, c, iw[0], iw[1], iw[2], radii,
                       tiles_touched, (float4*)splats);
  else
    hipLaunchKernelGGL(k_preprocess<1>, dim3(div_up(P, 256)), dim3(256), 0, s, m, c, iw[0], iw[1], iw[2], radii,
                       tiles_touched, (float4*)splats);
  return check_launch("preprocess");
}

int gsevt_duplicate_with_keys(int P, const uint32_t* offsets_inclusive, const int* radii, const float* splats,
                              int channels, int W, int H, int tile_w, int tile_h, uint64_t* keys, uint32_t* values,
                              long long capacity, void* stream) {
  if (check_tile(W, H, tile_w, tile_h)) return -1;
  GSEVT_REQUIRE(channels == 1 || channels == 3, "channels must be 1 or 3");
  if (P == 0) return 0;
  CamArgs c = make_cam_args(nullptr, nullptr, nullptr, nullptr, W, H, 0.f, 0.f, tile_w, tile_h);
  hipStream_t s = (hipStream_t)stream;
  if (channels == 3)
    hipLaunchKernelGGL(k_duplicate<3>, dim3(div_up(P, 256)), dim3(256), 0, s, P, offsets_inclusive, radii,
                       (const float4*)splats, c, keys, values, capacity);
  else
    hipLaunchKernelGGL(k_duplicate<2>, dim3(div_up(P, 256)), dim3(256), 0, s, P, offsets_inclusive, radii,
                       (const float4*)splats, c, keys, values, capacity);
  return check_launch("duplicate_with_keys");
}

int gsevt_tile_ranges(const uint32_t* d_num_rendered, long long capacity, const uint64_t* sorted_keys, int n_tiles,
                      uint32_t* ranges, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_fill_u32, dim3(fill_blocks(2LL * n_tiles)), dim3(256), 0, s, ranges, 2LL * n_tiles, 0u);
  if (check_launch("tile_ranges/fill")) return -1;
  if (capacity <= 0) return 0;
  hipLaunchKernelGGL(k_tile_ranges, dim3(div_up(capacity, 256)), dim3(256), 0, s, d_num_rendered, capacity,
                     sorted_keys, (uint2*)ranges);
  return check_launch("tile_ranges");
}

int gsevt_blend_forward(int W, int H, int tile_w, int tile_h, const uint32_t* ranges, const uint32_t* values,
                        const float* splats, int channels, const float* bg, float* out_color, float* out_depth,
                        float* out_opacity, float* final_T, int* n_contrib, int* n_touched, void* stream) {
  if (check_tile(W, H, tile_w, tile_h)) return -1;
  GSEVT_REQUIRE(channels == 1 || channels == 3, "channels must be 1 or 3");
  BlendArgs a = {};
  a.W = W; a.H = H; a.tile_w = tile_w; a.tile_h = tile_h; a.grid_x = (W + tile_w - 1) / tile_w;
  a.ranges = (const uint2*)ranges; a.vals = values; a.gid_mask = 0xFFFFFFFFu; a.splats = (const float4*)splats;
  a.bgp = bg;
  a.out_color = out_color; a.out_depth = out_depth; a.out_opac = out_opacity; a.final_T = final_T;
  a.n_contrib = n_contrib; a.n_touched = n_touched;
  const int ntiles = a.grid_x * ((H + tile_h - 1) / tile_h);
  hipStream_t s = (hipStream_t)stream;
  if (channels == 3)
    hipLaunchKernelGGL(k_blend_fwd<3>, dim3(ntiles), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(k_blend_fwd<1>, dim3(ntiles), dim3(256), 0, s, a);
  return check_launch("blend_forward");
}

int gsevt_blend_backward(int W, int H, int tile_w, int tile_h, const uint32_t* ranges, const uint32_t* values,
                         const float* splats, int channels, const float* bg, const float* final_T,
                         const int* n_contrib, const float* dL_dcolor, const float* dL_ddepth,
                         const float* dL_dopacity, float* grad10, void* stream) {
  if (check_tile(W, H, tile_w, tile_h)) return -1;
  GSEVT_REQUIRE(channels == 1 || channels == 3, "channels must be 1 or 3");
  GSEVT_REQUIRE(dL_dcolor && grad10, "dL_dcolor / grad buffer is null");
  BlendArgs a = {};
  a.W = W; a.H = H; a.tile_w = tile_w; a.tile_h = tile_h; a.grid_x = (W + tile_w - 1) / tile_w;
  a.ranges = (const uint2*)ranges; a.vals = values; a.gid_mask = 0xFFFFFFFFu; a.splats = (const float4*)splats;
  a.bgp = bg;
  a.final_T = (float*)final_T; a.n_contrib = (int*)n_contrib;
  a.g_color = dL_dcolor; a.g_depth = dL_ddepth; a.g_opac = dL_dopacity; a.grad10 = grad10;
  const int ntiles = a.grid_x * ((H + tile_h - 1) / tile_h);
  hipStream_t s = (hipStream_t)stream;
  if (channels == 3)
    hipLaunchKernelGGL(k_blend_bwd<3>, dim3(ntiles), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(k_blend_bwd<1>, dim3(ntiles), dim3(256), 0, s, a);
  return check_launch("blend_backward");
}

int gsevt_pose_partials_count(int P) { return P <= 0 ? 1 : (int)div_up(P, 256); }

int gsevt_preprocess_backward(int P, int D, int M, const float* means3D, const float* scales, float scale_modifier,
                              const float* rotations, const float* shs, const float* colors_precomp,
                              const float* cov3D_precomp, const float* viewmatrix, const float* projmatrix,
                              const float* projmatrix_raw, const float* campos, int W, int H, float tanfovx,
                              float tanfovy, int tile_w, int tile_h, const int* radii, const float* grad10,
                              int channels, const float* intensity_weights, float* pose_partials,
                              float* dL_dmeans3D, float* dL_dscales, float* dL_drotations, float* dL_dshs,
                              float* dL_dopacity, float* dL_dmeans2D, float* dL_dcolors, float* dL_dcov3D,
                              void* stream) {
  const float one = 1.0f;
  if (check_map(P, D, M, means3D, scales, rotations, &one, shs, colors_precomp, cov3D_precomp)) return -1;
  if (check_tile(W, H, tile_w, tile_h)) return -1;
  GSEVT_REQUIRE(viewmatrix && projmatrix && projmatrix_raw && campos, "camera matrices are null");
  GSEVT_REQUIRE(radii && grad10 && pose_partials, "buffers are null");
  hipStream_t s = (hipStream_t)stream;
  if (P == 0) {
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, s, (uint32_t*)pose_partials, 18LL, 0u);
    return check_launch("preprocess_backward/empty");
  }
  MapArgs m = make_map_args(P, D, M, means3D, scales, scale_modifier, rotations, nullptr, shs, colors_precomp,
                            cov3D_precomp);
  CamArgs c = make_cam_args(viewmatrix, projmatrix, projmatrix_raw, campos, W, H, tanfovx, tanfovy, tile_w, tile_h);
  BwdArgs b = {};
  b.radii = radii; b.grad10 = grad10; b.C = channels;
  b.iw[0] = 0.2989f; b.iw[1] = 0.5870f; b.iw[2] = 0.1140f;
  if (intensity_weights) for (int k = 0; k < 3; ++k) b.iw[k] = intensity_weights[k];
  b.partials = pose_partials;
  b.g_means = dL_dmeans3D; b.g_scales = dL_dscales; b.g_rots = dL_drotations; b.g_shs = dL_dshs;
  b.g_opac = dL_dopacity; b.g_means2d = dL_dmeans2D; b.g_colors = dL_dcolors; b.g_cov = dL_dcov3D;
  hipLaunchKernelGGL(k_preprocess_bwd, dim3(div_up(P, 256)), dim3(256), 0, s, m, c, b);
  return check_launch("preprocess_backward");
}

int gsevt_pose_finalize(const float* pose_partials, int n_partials, const float* viewmatrix,
                        const float* vel_transform, double delta_time, const float* linear_vel,
                        const float* angular_vel, float* out12, int accumulate, void* stream) {
  GSEVT_REQUIRE(pose_partials && viewmatrix && out12, "buffers are null");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_pose_finalize, dim3(1), dim3(64), 0, s, pose_partials, n_partials, viewmatrix,
                     vel_transform, delta_time, linear_vel, angular_vel, out12, accumulate);
  return check_launch("pose_finalize");
}

int gsevt_mark_visible(int P, const float* means3D, const float* viewmatrix, const float* projmatrix,
                       uint8_t* present, void* stream) {
  GSEVT_REQUIRE(P >= 0, "P must be non-negative (got %d)", P);
  if (P == 0) return 0;
  GSEVT_REQUIRE(means3D && viewmatrix && projmatrix && present, "buffers are null");
  hipLaunchKernelGGL(k_mark_visible, dim3(div_up(P, 256)), dim3(256), 0, (hipStream_t)stream, P, means3D,
                     viewmatrix, present);
  return check_launch("mark_visible");
}

int gsevt_fill_u32(uint32_t* p, long long n, uint32_t value, void* stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_fill_u32, dim3(fill_blocks(n)), dim3(256), 0, (hipStream_t)stream, p, n, value);
  return check_launch("fill_u32");
}

}  // extern "C"
