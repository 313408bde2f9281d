// Tile-level blend bodies shared by the single-render and batched launches.
//
// One 256-thread workgroup per tile; each wave owns 64 pixels of it.  A wave
// loads the tile's fragments in 64-entry batches, each lane testing whether
// its fragment's alpha >= 1/255 ellipse can reach the wave's pixels; a ballot
// of those tests is the only set of fragments the wave then visits, and the
// forward leaves as soon as a ballot shows every lane saturated.  Skipped
// fragments contribute exactly nothing, so results are bit-identical to
// visiting every fragment.
#pragma once

#include "gsevt_device.h"

namespace gsevt {

struct BlendArgs {
  int W, H, tile_w, tile_h, grid_x;
  const uint2* ranges;
  const uint32_t* vals;
  uint32_t gid_mask;    // values may carry a tile id above the Gaussian index bits
  const float4* splats;
  const float* bgp;    // device background (channels floats) or null
  float bgv[3];        // background by value when bgp is null
  float* out_color;    // [C][H][W]
  float* out_depth;    // [H][W]
  float* out_opac;     // [H][W]
  float* final_T;      // [H][W]
  int* n_contrib;      // [H][W]
  int* n_touched;      // [P] or null
  // backward
  const float* g_color;
  const float* g_depth;
  const float* g_opac;
  float* grad10;       // [P][10]; [P][6] for the lean one-channel backward
};

// Opaque copy: later code that uses the value cannot be moved above this
// point (keeps the cull of a prefetched batch, and the wait for its loads,
// after the visits of the current one).
template <int S>
__device__ __forceinline__ void pin_after(float4 (&r)[S]) {
#pragma unroll
  for (int s = 0; s < S; ++s) asm volatile("" : "+v"(r[s].x), "+v"(r[s].y), "+v"(r[s].z), "+v"(r[s].w));
}

// Stage one batch of a wave: each valid lane culls its fragment against the
// wave's pixels (splat_may_touch) and writes its record to the wave's LDS
// slice; returns the ballot of fragments to visit.  NL_IN_REC (blends that
// read no depth): -L' replaces the third float of the record's second float4
// in LDS, so one 16-byte read per visit brings conic c, opacity, -L' and the
// colour; otherwise -L' goes to negl.
//
// LDS layout: one plane of 64 float4 per record slot (rec[s][lane]), so the
// 64 lanes' 16-byte stores of a slot are contiguous and conflict-free (with
// the records lane-major, a 32- or 48-byte lane stride made lanes 0 and 4 of
// each group of 8 share banks); the visits read one slot of one fragment,
// the same address for every lane (a broadcast).
template <int S, bool NL_IN_REC>
__device__ __forceinline__ unsigned long long blend_stage(bool valid, int g, const float4 (&r)[S], const Strip& sb,
                                                          float4 (*rec)[WAVE], int* gid, float* negl, int lane) {
  bool keep = false;
  if (valid) {
    float nl;
    keep = splat_may_touch(r[0], r[1], sb, nl);
    gid[lane] = g;
    if (!NL_IN_REC) negl[lane] = nl;
    rec[0][lane] = r[0];
    rec[1][lane] = NL_IN_REC ? make_float4(r[1].x, r[1].y, nl, r[1].w) : r[1];
#pragma unroll
    for (int s = 2; s < S; ++s) rec[s][lane] = r[s];
  }
  return __ballot(keep);
}

// Forward blend of one tile (one 256-thread workgroup).  AUX: also write depth,
// opacity and n_touched.
template <int C, bool AUX>
__device__ __forceinline__ void blend_fwd_tile(const BlendArgs& a, int tile) {
  constexpr int S = (C == 1) ? 2 : 3;
  constexpr bool NLR = !AUX;   // no depth output: -L' travels in the record
  __shared__ float4 s_rec[4][S][WAVE];
  __shared__ int s_gid[4][WAVE];
  __shared__ float s_negl[4][WAVE];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  int px, py;
  tile_pixel(tile, tid, a.grid_x, a.tile_w, a.tile_h, px, py);
  const bool inside = px < a.W && py < a.H;
  const uint2 rg = a.ranges[tile];
  const float pxf = (float)px, pyf = (float)py;
  const Strip sb = wave_strip(tile, wave, a.grid_x, a.tile_w, a.tile_h);
  // 1 while the pixel may still take fragments (an int, not a bool: a
  // loop-carried bool becomes a lane mask whose merges cost scalar
  // instructions on every visit)
  int alive = inside ? 1 : 0;
  float T = 1.0f, D = 0.0f, A = 0.0f;
  float col[C];
#pragma unroll
  for (int c = 0; c < C; ++c) col[c] = 0.0f;
  int last = 0;
  const bool count_touch = AUX && a.n_touched != nullptr;
  // Batch pipeline: batch i is staged in LDS (culled) by the end of iteration
  // i-1; during its visits the records of batch i+1 and the Gaussian ids of
  // batch i+2 are in flight.  The visits read LDS only, so no wait on those
  // loads falls inside them; loads are unconditional (clamped index) and ids
  // are masked only when consumed.
  const uint32_t rend = rg.y;
  unsigned long long mask = 0ull;
  bool v1 = false;
  uint32_t g1raw = 0u;
  if (rg.x < rend) {
    const bool v0 = rg.x + lane < rend;
    const int g0 = v0 ? (int)(a.vals[rg.x + lane] & a.gid_mask) : 0;
    float4 rc[S];
#pragma unroll
    for (int s = 0; s < S; ++s) rc[s] = a.splats[(size_t)g0 * S + s];
    mask = blend_stage<S, NLR>(v0, g0, rc, sb, s_rec[wave], s_gid[wave], s_negl[wave], lane);
    const uint32_t i1 = rg.x + WAVE + lane;
    v1 = i1 < rend;
    g1raw = a.vals[v1 ? i1 : rend - 1];
  }
  for (uint32_t base = rg.x; base < rend; base += WAVE) {
    if (__ballot(alive != 0) == 0) break;
    const int g1 = (int)(g1raw & a.gid_mask);
    float4 rn[S];
#pragma unroll
    for (int s = 0; s < S; ++s) rn[s] = a.splats[(size_t)g1 * S + s];
    const uint32_t i2 = base + 2 * WAVE + lane;
    const bool v2 = i2 < rend;
    const uint32_t g2raw = a.vals[v2 ? i2 : rend - 1];
    wave_sync();
    while (mask != 0ull) {
      const int j = __builtin_ctzll(mask);
      mask &= mask - 1ull;
      const float4 r0 = s_rec[wave][0][j];
      const float4 r1 = s_rec[wave][1][j];
      const float dx = r0.x - pxf, dy = r0.y - pyf;
      const float power = -0.5f * ((r0.z * dx) * dx + (r1.x * dy) * dy) - (r0.w * dx) * dy;
      const float nl = NLR ? r1.z : s_negl[wave][j];
      // non-short-circuit tests
      const bool inr = (alive != 0) & !(power > 0.0f) & !(power < nl);
      // The wave skips the fragment when no lane is in range (a uniform
      // branch); otherwise every lane evaluates it and lanes out of range, or
      // failing the alpha or saturation test, keep their state through the
      // selects (no divergent branch: its exec-mask bookkeeping cost more
      // scalar instructions than the lanes it idled saved).
      if (__ballot(inr) == 0ull) continue;
      const float G = det_exp_unclamped(power);
      const float oG = r1.y * G;
      const float alpha = __builtin_fminf(oG, 0.99f);   // = (0.99 < oG ? 0.99 : oG) for every non-NaN oG
      const bool ok = inr & !(alpha < 1.0f / 255.0f);
      const float testT = T * (1.0f - alpha);
      const bool fin = ok & (testT < 0.0001f);
      const bool contrib = ok & !fin;
      alive = fin ? 0 : alive;
      const float w = alpha * T;
      col[0] = contrib ? col[0] + r1.w * w : col[0];
      if (C > 1) {
        const float4 r2 = s_rec[wave][S - 1][j];
        col[C > 1 ? 1 : 0] = contrib ? col[C > 1 ? 1 : 0] + r2.x * w : col[C > 1 ? 1 : 0];
        col[C > 2 ? 2 : 0] = contrib ? col[C > 2 ? 2 : 0] + r2.y * w : col[C > 2 ? 2 : 0];
      }
      if (AUX) {
        D = contrib ? D + r1.z * w : D;
        A = contrib ? A + w : A;
      }
      T = contrib ? testT : T;
      last = contrib ? (int)(base - rg.x) + j + 1 : last;
      if (count_touch) {   // lane 0 adds the wave's count
        const unsigned long long b = __ballot(contrib);
        if (b != 0ull && lane == 0) atomicAdd(&a.n_touched[s_gid[wave][j]], (int)__popcll(b));
      }
      if (__ballot(alive != 0) == 0ull) break;
    }
    wave_sync();
    pin_after<S>(rn);
    mask = blend_stage<S, NLR>(v1, g1, rn, sb, s_rec[wave], s_gid[wave], s_negl[wave], lane);
    v1 = v2;
    g1raw = g2raw;
  }
  if (inside) {
    const int pix = py * a.W + px;
    const size_t HW = (size_t)a.W * a.H;
#pragma unroll
    for (int c = 0; c < C; ++c) a.out_color[c * HW + pix] = col[c] + T * (a.bgp ? a.bgp[c] : a.bgv[c]);
    if (AUX) {
      a.out_depth[pix] = D;
      a.out_opac[pix] = A;
    }
    a.final_T[pix] = T;
    a.n_contrib[pix] = last;
  }
}

// ---------------------------------------------------------------------------
// blend backward
// ---------------------------------------------------------------------------
// Backward blend of one tile; gC/gD/gA are the lane's pixel gradients.
// LEAN (one channel, no depth/opacity loss): only the mean, conic and colour
// gradients are formed and reduced (6-slot butterfly); opacity and depth
// slots stay zero.
template <int C, bool LEAN = false>
__device__ __forceinline__ void blend_bwd_tile(const BlendArgs& a, int tile, const float* gC, float gD, float gA) {
  static_assert(!LEAN || C == 1, "lean backward is one-channel");
  constexpr int S = (C == 1) ? 2 : 3;
  constexpr bool NLR = LEAN;   // no depth gradient: -L' travels in the record
  __shared__ float4 s_rec[4][S][WAVE];
  __shared__ int s_gid[4][WAVE];
  __shared__ float s_negl[4][WAVE];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  int px, py;
  tile_pixel(tile, tid, a.grid_x, a.tile_w, a.tile_h, px, py);
  const bool inside = px < a.W && py < a.H;
  const uint2 rg = a.ranges[tile];
  const float pxf = (float)px, pyf = (float)py;
  const Strip sb = wave_strip(tile, wave, a.grid_x, a.tile_w, a.tile_h);
  const int pix = py * a.W + px;
  const int last = inside ? a.n_contrib[pix] : 0;
  const float TF = inside ? a.final_T[pix] : 1.0f;
  float bgdot = 0.0f;
#pragma unroll
  for (int c = 0; c < C; ++c) bgdot = bgdot + (a.bgp ? a.bgp[c] : a.bgv[c]) * gC[c];
  const float tf_bg = TF * bgdot;   // background term of dL/dalpha, before the 1 / (1 - alpha) scale
  float Tr = TF;
  float accC[C];
#pragma unroll
  for (int c = 0; c < C; ++c) accC[c] = 0.0f;
  float accD = 0.0f, accA = 0.0f;
  const int wave_last = wave_max_i(last);
  // Batches run from the wave's last contributor down, pipelined as in the
  // forward (staged one batch ahead, records one and ids two batches ahead).
  int e0 = wave_last;
  int s0 = e0 > WAVE ? e0 - WAVE : 0;
  int e1 = s0, s1 = e1 > WAVE ? e1 - WAVE : 0;
  unsigned long long mask = 0ull;
  bool v1 = false;
  uint32_t g1raw = 0u;
  if (e0 > 0) {
    const bool v0 = lane < e0 - s0;
    const int g0 = v0 ? (int)(a.vals[rg.x + s0 + lane] & a.gid_mask) : 0;
    float4 rc[S];
#pragma unroll
    for (int s = 0; s < S; ++s) rc[s] = a.splats[(size_t)g0 * S + s];
    mask = blend_stage<S, NLR>(v0, g0, rc, sb, s_rec[wave], s_gid[wave], s_negl[wave], lane);
    v1 = lane < e1 - s1;
    g1raw = a.vals[rg.x + (v1 ? s1 + lane : 0)];
  }
  while (e0 > 0) {
    const int start = s0;
    const int last_rel = last - start;   // fragment j of this batch is before the lane's last iff j < last_rel
    const int g1 = (int)(g1raw & a.gid_mask);
    float4 rn[S];
#pragma unroll
    for (int s = 0; s < S; ++s) rn[s] = a.splats[(size_t)g1 * S + s];
    const int e2 = s1, s2 = e2 > WAVE ? e2 - WAVE : 0;
    const bool v2 = lane < e2 - s2;
    const uint32_t g2raw = a.vals[rg.x + (v2 ? s2 + lane : 0)];
    wave_sync();
    while (mask != 0ull) {
      const int j = 63 - __builtin_clzll(mask);
      mask &= ~(1ull << j);
      // Non-contributing lanes keep their state and add zeros (selects).
      const float4 r0 = s_rec[wave][0][j];
      const float4 r1 = s_rec[wave][1][j];
      const float dx = r0.x - pxf, dy = r0.y - pyf;
      const float ca = r0.z, cb = r0.w, cc = r1.x, o = r1.y, z = r1.z;
      // the forward's products, kept for the mean gradients below
      const float cadx = ca * dx, ccdy = cc * dy, cbdx = cb * dx;
      const float power = -0.5f * (cadx * dx + ccdy * dy) - cbdx * dy;
      const float nl = NLR ? r1.z : s_negl[wave][j];
      const bool inr = (j < last_rel) & !(power > 0.0f) & !(power < nl);
      if (__ballot(inr) == 0ull) continue;
      const float G = det_exp_unclamped(power);
      const float oG = o * G;
      const float alpha = __builtin_fminf(oG, 0.99f);   // as in the forward
      // lanes in range all but always pass the 1/255 test as well (the range
      // is the same test with a 1e-4 margin), so no second early exit
      const bool contrib = inr & !(alpha < 1.0f / 255.0f);
      float colj[C];
      if (C == 1) {
        colj[0] = r1.w;
      } else {
        const float4 r2 = s_rec[wave][S - 1][j];
        colj[0] = r1.w;
        colj[C > 1 ? 1 : 0] = r2.x;
        colj[C > 2 ? 2 : 0] = r2.y;
      }
      const float om = 1.0f - alpha;
      // 1 / (1 - alpha) once (alpha <= 0.99): T before this fragment and the
      // background term both scale by it
      const float rom = __builtin_amdgcn_rcpf(om);
      const float Trn = Tr * rom;
      const float w = alpha * Trn;
      float v[16];
      // gradient-only arithmetic below may fuse multiply-adds (the decisions
      // and the transmittance above keep the forward's exact operations)
      float dLda = (colj[0] - accC[0]) * gC[0];
      // LEAN: one select of the weight instead of one per product (its pixel
      // gradient is finite on every lane); otherwise a pixel gradient that is
      // not finite reaches only the fragments that contribute to its pixel
      const float wc = contrib ? w : 0.0f;
#pragma unroll
      for (int c = 0; c < C; ++c) {
        v[6 + c] = LEAN ? wc * gC[c] : (contrib ? w * gC[c] : 0.0f);
        if (c > 0) dLda = __builtin_fmaf(colj[c] - accC[c], gC[c], dLda);
      }
      if (!LEAN) {
        v[9] = contrib ? w * gD : 0.0f;
        dLda = dLda + (z - accD) * gD;
        dLda = dLda + (1.0f - accA) * gA;
      }
      dLda = dLda * Trn;
      dLda = __builtin_fmaf(-rom, tf_bg, dLda);
#pragma unroll
      for (int c = 0; c < C; ++c) accC[c] = contrib ? __builtin_fmaf(alpha, colj[c], om * accC[c]) : accC[c];
      if (!LEAN) {
        accD = contrib ? alpha * z + om * accD : accD;
        accA = contrib ? alpha + om * accA : accA;
      }
      Tr = contrib ? Trn : Tr;
      // the 0.99 clamp passes no gradient to G
      const bool nc = contrib & !(oG > 0.99f);
      // one select of the power gradient instead of one per product (the
      // products of a zero with the finite offsets are zeros)
      const float gpow = nc ? (o * dLda) * G : 0.0f;
      const float hg = -0.5f * gpow;
      v[0] = gpow * -__builtin_fmaf(cb, dy, cadx);
      v[1] = gpow * -(cbdx + ccdy);
      v[2] = hg * (dx * dx);
      v[3] = gpow * (-dx * dy);
      v[4] = hg * (dy * dy);
      v[5] = (!LEAN && nc) ? G * dLda : 0.0f;
#pragma unroll
      for (int q = 6 + C; q < 16; ++q)
        if (q != 9 || LEAN) v[q] = 0.0f;
      {
        if (LEAN) {
          // six gradients per Gaussian, stride 6: mean 2, conic 3, colour
          const float v6[6] = {v[0], v[1], v[2], v[3], v[4], v[6]};
          float sred = wave_reduce6(v6, lane);
          // complete the reduction on every lane here (its last step is a DPP
          // add) rather than inside the writers' branch
          asm volatile("" : "+v"(sred));
          if (reduce6_writer(lane)) atomicAdd(&a.grad10[(size_t)s_gid[wave][j] * 6 + reduce6_slot(lane)], sred);
        } else {
          const float sred = wave_reduce16(v, lane);
          const int slot = reduce16_slot(lane);
          if ((lane & 3) == 0 && slot < 10) atomicAdd(&a.grad10[(size_t)s_gid[wave][j] * 10 + slot], sred);
        }
      }
    }
    wave_sync();
    pin_after<S>(rn);
    mask = blend_stage<S, NLR>(v1, g1, rn, sb, s_rec[wave], s_gid[wave], s_negl[wave], lane);
    e0 = e1; s0 = s1;
    e1 = e2; s1 = s2;
    v1 = v2;
    g1raw = g2raw;
  }
}


}  // namespace gsevt
