// Batched, graph-capturable tracking iteration for gfx950.
//
// An item is one (pose hypothesis, event window) pair; it owns two renders
// (last / next camera of the window).  For B items, R = 2B renders go through
// every launch together, indexed by blockIdx.y, so each launch fills the GPU
// and no per-render host work or host synchronisation remains:
//
//   tr_cams          render poses Exp(+-tau/2 [v; w]) T_c, camera records
//   tr_preprocess    splats of every (Gaussian, render): grey intensity
//   depth sort       stable 4-pass radix sort of each render's Gaussians by
//                    depth bits (invisible ones keyed 0xFFFFFFFF sort last)
//   tr_emit          fragments in depth order (scan of tiles_touched in that
//                    order gives the offsets)
//   tile sort        stable 2-pass radix sort of each render's fragments by
//                    tile id, then tile ranges
//   blend forward    shared tile body (blend_impl.h), one channel
//   tr_loss          per-item sums of dI^2, dI e, e^2 in FP64 -> loss,
//                    norm and the scalars of the loss gradient
//   blend backward   each pixel forms dL/dI from I_last, I_next, the event
//                    frame and the item scalars, then replays in reverse
//   tr_preprocess_bwd per (Gaussian, render) 18 pose partials per workgroup
//   tr_update        per item: FP64 reduction, adjoint of the velocity
//                    transform, Adam (torch defaults), pose/velocity update
//
// Depth-then-tile binning yields exactly the order of a stable sort on the
// 64-bit key tile << 32 | depth bits: within a tile, fragments follow the
// depth order of their Gaussians, ties in Gaussian index order.
#include "gsevt_api.h"
#include "gsevt_device.h"
#include "blend_impl.h"

using namespace gsevt;

namespace {

// ---------------------------------------------------------------------------
// small FP64 Lie helpers (device)
// ---------------------------------------------------------------------------
__device__ void se3_exp_d(const double* xi, double R[3][3], double t[3]) {
  // xi = [rho; theta]; Rodrigues with the |theta| < 1e-5 series branch
  const double* rho = xi;
  const double* th = xi + 3;
  const double W[3][3] = {{0, -th[2], th[1]}, {th[2], 0, -th[0]}, {-th[1], th[0], 0}};
  double W2[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) W2[r][c] = W[r][0] * W[0][c] + W[r][1] * W[1][c] + W[r][2] * W[2][c];
  const double a = sqrt(th[0] * th[0] + th[1] * th[1] + th[2] * th[2]);
  double A, Bc, Vb, Vc;
  if (a < 1e-5) {
    A = 1.0; Bc = 0.5; Vb = 0.5; Vc = 1.0 / 6.0;
  } else {
    A = sin(a) / a; Bc = (1 - cos(a)) / (a * a); Vb = (1 - cos(a)) / (a * a); Vc = (a - sin(a)) / (a * a * a);
  }
  double V[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      const double I = (r == c) ? 1.0 : 0.0;
      R[r][c] = I + A * W[r][c] + Bc * W2[r][c];
      V[r][c] = I + Vb * W[r][c] + Vc * W2[r][c];
    }
  for (int r = 0; r < 3; ++r) t[r] = V[r][0] * rho[0] + V[r][1] * rho[1] + V[r][2] * rho[2];
}

struct Intr {
  float projraw[16];   // transposed storage
  float tanfovx, tanfovy;
  int W, H, tile_w, tile_h, grid_x, grid_y;
};

struct ItemState {
  double* pose;        // [B][12]: R row-major (9), t (3), world -> camera
  float* lin;          // [B][3]
  float* ang;          // [B][3]
  float* adam_m;       // [B][12]  (theta, rho, w, v)
  float* adam_v;       // [B][12]
  int* adam_t;         // [B][4]
  int* iter;           // [B]
  float* dtau;         // [B]
  const int* window;   // [B] event frame of the item
  float* loss;         // [B]
  float* loss_hist;    // [B][hist_len]
  int hist_len;
  float* grad12;       // [B][12] last gradient (diagnostics / tests)
  int* stage;          // [B] or null: 0 coarse, 1 fine, 2 done
  int* start_vel;      // [B] or null: iteration at which the fine stage started
};

// ---------------------------------------------------------------------------
// cameras
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_tr_cams(int B, ItemState st, Intr in, CamDev* __restrict__ cams,
                                                float* __restrict__ vel, float* __restrict__ dts) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= 2 * B) return;
  const int b = r >> 1, side = r & 1;
  const double s = (side ? 0.5 : -0.5) * (double)st.dtau[b];
  double xi[6];
  for (int k = 0; k < 3; ++k) { xi[k] = s * (double)st.lin[3 * b + k]; xi[3 + k] = s * (double)st.ang[3 * b + k]; }
  double ER[3][3], Et[3];
  se3_exp_d(xi, ER, Et);
  const double* pc = st.pose + 12 * b;
  double Rr[3][3], tr[3];
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) Rr[i][j] = ER[i][0] * pc[j] + ER[i][1] * pc[3 + j] + ER[i][2] * pc[6 + j];
    tr[i] = ER[i][0] * pc[9] + ER[i][1] * pc[10] + ER[i][2] * pc[11] + Et[i];
  }
  CamDev c;
  // view: M[r][c] = m[4c + r] with M = [R t; 0 1]
  double Tm[4][4];
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) Tm[i][j] = Rr[i][j];
    Tm[i][3] = tr[i];
  }
  Tm[3][0] = Tm[3][1] = Tm[3][2] = 0.0; Tm[3][3] = 1.0;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) c.view[4 * j + i] = (float)Tm[i][j];
  // full projection P * T (P from the transposed-storage projraw)
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      double acc = 0.0;
      for (int k = 0; k < 4; ++k) acc += (double)in.projraw[4 * k + i] * Tm[k][j];
      c.proj[4 * j + i] = (float)acc;
    }
  for (int k = 0; k < 16; ++k) c.projraw[k] = in.projraw[k];
  for (int i = 0; i < 3; ++i) c.campos[i] = (float)(-(Rr[0][i] * tr[0] + Rr[1][i] * tr[1] + Rr[2][i] * tr[2]));
  c.tanfovx = in.tanfovx; c.tanfovy = in.tanfovy;
  c.W = in.W; c.H = in.H; c.tile_w = in.tile_w; c.tile_h = in.tile_h; c.grid_x = in.grid_x; c.grid_y = in.grid_y;
  cams[r] = c;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      const double e = (i < 3 && j < 3) ? ER[i][j] : (i < 3 ? Et[i] : (j == 3 ? 1.0 : 0.0));
      vel[16 * r + 4 * j + i] = (float)e;
    }
  dts[r] = (float)s;
}

// ---------------------------------------------------------------------------
// preprocess (grey intensity, frozen map with precomputed 3D covariance)
// ---------------------------------------------------------------------------
struct TrMap {
  int P, D, M;
  const float* means;   // [P][3]
  const float* cov;     // [P][6]
  const float* opac;    // [P]
  const float* shs;     // [3K][P] coefficient-major (row 3k + c)
  float iw[3];
};

struct TrBufs {
  float4* splats;       // [R][P][2]
  uint2* rects;         // [R][P] packed x0|y0<<16, x1|y1<<16
  uint32_t* tiles;      // [R][P]
  uint32_t* dkey;       // [R][P]
  uint32_t* dval;       // [R][P] gaussian | tile count << count_shift (count_shift 0: gaussian only)
  int count_shift;
};

__global__ __launch_bounds__(256) void k_tr_map_cov(int P, const float* __restrict__ scales, float mod,
                                                    const float* __restrict__ rots, float* __restrict__ cov) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  float R[3][3], sig[6];
  cov3d(scales + 3 * i, mod, rots + 4 * i, sig, R);
#pragma unroll
  for (int k = 0; k < 6; ++k) cov[6 * i + k] = sig[k];
}

// One thread per Gaussian, RPT renders per thread: the map entry (mean,
// covariance, opacity, SH) is loaded once into registers.  SH is read from the
// coefficient-major copy shT[3K][P], so every load is lane-contiguous.
constexpr int RPT = 32;

template <int D>
__global__ __launch_bounds__(256) void k_tr_preprocess(TrMap m, const CamDev* __restrict__ cams, int R, TrBufs o) {
  constexpr int K = (D + 1) * (D + 1);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.P) return;
  const int r0 = blockIdx.y * RPT;
  const int r1 = (r0 + RPT < R) ? r0 + RPT : R;
  const float mu[3] = {m.means[3 * i], m.means[3 * i + 1], m.means[3 * i + 2]};
  float sig[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) sig[k] = m.cov[6 * i + k];
  const float op = m.opac[i];
  float sh[3 * K];
#pragma unroll
  for (int k = 0; k < 3 * K; ++k) sh[k] = m.shs[(size_t)k * m.P + i];
  for (int r = r0; r < r1; ++r) {
    const CamDev& c = cams[r];
    const size_t ri = (size_t)r * m.P + i;
    Fwd f;
    if (!project_one(mu, sig, c, f)) {
      if (o.tiles) o.tiles[ri] = 0;
      o.dkey[ri] = 0xFFFFFFFFu;
      o.dval[ri] = (uint32_t)i;
      continue;
    }
    float dir[3], dlen, rgb[3];
    bool cl[3];
    sh_color_reg<D>(mu, sh, c, dir, dlen, rgb, cl);
    const float inten = m.iw[0] * rgb[0] + m.iw[1] * rgb[1] + m.iw[2] * rgb[2];
    o.splats[2 * ri] = make_float4(f.px, f.py, f.conic[0], f.conic[1]);
    // third slot: the tile rect, one byte per bound (x0, y0, x1, y1); the
    // one-channel blends never read a depth
    const uint32_t rect8 = (uint32_t)f.rect[0] | ((uint32_t)f.rect[1] << 8) | ((uint32_t)f.rect[2] << 16) |
                           ((uint32_t)f.rect[3] << 24);
    o.splats[2 * ri + 1] = make_float4(f.conic[2], op, __uint_as_float(rect8), inten);
    if (o.rects)
      o.rects[ri] = make_uint2((uint32_t)f.rect[0] | ((uint32_t)f.rect[1] << 16),
                               (uint32_t)f.rect[2] | ((uint32_t)f.rect[3] << 16));
    const uint32_t count = (uint32_t)((f.rect[2] - f.rect[0]) * (f.rect[3] - f.rect[1]));
    if (o.tiles) o.tiles[ri] = count;
    o.dkey[ri] = __float_as_uint(f.pv[2]);
    o.dval[ri] = o.count_shift ? ((uint32_t)i | (count << o.count_shift)) : (uint32_t)i;
  }
}

// ---------------------------------------------------------------------------
// segmented stable radix sort (u32 keys, u32 values), one segment per render
// ---------------------------------------------------------------------------
// A pass over digit bits [shift, shift + nbits) is a histogram kernel, a scan
// of the digit table and a scatter kernel.  The table holds 2^nbits rows per
// segment, row d listing the count of digit d in each 4096-element block, so
// one inclusive scan over all segments' tables gives every block's output
// offset per digit.  Each thread keeps 16 keys (element q * 256 + tid of its
// block, q = 0..15) in registers; ranks come from per-wave ballots, so the
// order is stable and independent of scheduling.
constexpr int SEG_THREADS = 256;
constexpr int SEG_ROUNDS = 16;
constexpr int SEG_TILE = SEG_THREADS * SEG_ROUNDS;
// Key that a dropping sort pass removes (never a packed fragment: the host
// requires n_tiles << gid_bits < 2^32).
constexpr uint32_t DROPPED_FRAGMENT = 0xFFFFFFFFu;

__device__ __forceinline__ long long seg_count(const uint32_t* d_counts, long long n_fixed, long long stride, int r) {
  long long n = d_counts ? (long long)d_counts[r] : n_fixed;
  return n < stride ? n : stride;
}

// One count per valid lane into the wave's sub-histogram.  The digits most
// lanes share go first: up to two rounds take the digit of the first lane not
// yet counted and add all lanes with that digit in one atomic (a pass whose
// digits take a few values, as the top byte of depth keys does, would
// otherwise put most of a wave's atomics on one LDS word); the remaining
// lanes add one each.
__device__ __forceinline__ void hist_count(uint32_t* s_h, bool valid, uint32_t d) {
  const int lane = threadIdx.x & 63;
  unsigned long long left = __ballot(valid);
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    if (left == 0ull) return;   // wave-uniform
    const int leader = __builtin_ctzll(left);
    const uint32_t dl = (uint32_t)__builtin_amdgcn_readlane((int)d, leader);
    const unsigned long long same = __ballot(d == dl) & left;
    if (lane == leader) atomicAdd(&s_h[dl], (uint32_t)__popcll(same));
    left &= ~same;
  }
  if ((left >> lane) & 1ull) atomicAdd(&s_h[d], 1u);
}

template <int DB, bool DROP>
__global__ __launch_bounds__(SEG_THREADS) void k_seg_hist2(const uint32_t* __restrict__ keys, long long stride,
                                                           const uint32_t* __restrict__ d_counts, long long n_fixed,
                                                           int shift, int nbits, uint32_t* __restrict__ hist,
                                                           int nb) {
  constexpr int NB = 1 << DB;
  __shared__ uint32_t s_h[4][NB];   // one sub-histogram per wave (less atomic contention)
  const int r = blockIdx.y;
  const int tid = threadIdx.x, wave = tid >> 6;
  const long long rem = seg_count(d_counts, n_fixed, stride, r) - (long long)blockIdx.x * SEG_TILE;
  const uint32_t mask = (1u << nbits) - 1u;
  for (int e = tid; e < 4 * NB; e += SEG_THREADS) (&s_h[0][0])[e] = 0;
  __syncthreads();
  const uint32_t* k = keys + (size_t)r * stride + (size_t)blockIdx.x * SEG_TILE;
  if (rem >= SEG_TILE) {
    // full block: issue all loads before the first atomic
    uint32_t v[SEG_ROUNDS];
#pragma unroll
    for (int q = 0; q < SEG_ROUNDS; ++q) v[q] = k[q * SEG_THREADS + tid];
#pragma unroll
    for (int q = 0; q < SEG_ROUNDS; ++q)
      hist_count(s_h[wave], !DROP || v[q] != DROPPED_FRAGMENT, (v[q] >> shift) & mask);
  } else if (rem > 0) {
    const int m = (int)rem;
#pragma unroll
    for (int q = 0; q < SEG_ROUNDS; ++q) {
      const int i = q * SEG_THREADS + tid;
      const uint32_t x = (i < m) ? k[i] : 0u;
      hist_count(s_h[wave], i < m && (!DROP || x != DROPPED_FRAGMENT), (x >> shift) & mask);
    }
  }
  __syncthreads();
  const int rows = 1 << nbits;
  for (int d = tid; d < rows; d += SEG_THREADS)
    hist[((size_t)r * rows + d) * nb + blockIdx.x] = s_h[0][d] + s_h[1][d] + s_h[2][d] + s_h[3][d];
}

// Ranks of one 4096-element block.  FULL: every element in range (no bounds
// tests); drop: DROPPED_FRAGMENT keys are not ranked or counted.
// s_cnt[(q * 4 + wave) * NB + d] receives the count of digit d among the 64
// elements of (round q, wave); rank[q] is the element's rank among its wave's
// elements of the same digit in round q.  The peers of a lane (lanes of its
// wave with the same digit) come either from ballots - the intersection, over
// the digit's nbits bits, of the lanes with that bit set or of its complement
// (BALLOT) - or from one 64-bit LDS word per (wave, digit): cleared, OR-ed
// with 1 << lane by every lane of that digit, read back.  OR commutes, so the
// masks do not depend on the order of the atomics; the three steps are ordered
// by the wave's in-order LDS queue.  Both are independent of scheduling.
// Ballots serve the 8-bit depth passes (the 256 masks per wave conflicted on
// 48% of LDS cycles and their 8 KB cut occupancy): 2.97 -> 2.75 ms per pass of
// 512 renders; the masks serve the 5- and 6-bit tile passes, where ballots
// were slower (9.5 -> 12.2 ms for the dropping pass).
template <int DB, bool FULL, bool DROP, bool BALLOT>
__device__ __forceinline__ void seg_rank_block(const uint32_t (&key)[SEG_ROUNDS], uint32_t (&rank)[SEG_ROUNDS],
                                               int m, int shift, int nbits, uint32_t mask, uint16_t* s_cnt,
                                               unsigned long long* s_match) {
  constexpr int NB = 1 << DB;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
  for (int q = 0; q < SEG_ROUNDS; ++q) {
    const bool valid = (FULL || q * SEG_THREADS + tid < m) && !(DROP && key[q] == DROPPED_FRAGMENT);
    const uint32_t d = (key[q] >> shift) & mask;
    unsigned long long peers;
    if (BALLOT) {
      peers = __ballot(valid);
      for (int b = 0; b < nbits; ++b) {
        const bool bit = (d >> b) & 1u;
        const unsigned long long bb = __ballot(bit);
        peers &= bit ? bb : ~bb;
      }
      peers = valid ? peers : 0ull;
    } else {
      unsigned long long* mw = s_match + wave * NB;
      if (valid) mw[d] = 0ull;
      wave_sync();
      if (valid) atomicOr(&mw[d], 1ull << lane);
      wave_sync();
      peers = valid ? mw[d] : 0ull;
      wave_sync();
    }
    rank[q] = (uint32_t)__popcll(peers & lt_mask);
    if (valid && rank[q] == 0) s_cnt[(q * 4 + wave) * NB + d] = (uint16_t)__popcll(peers);
  }
}

// Scatter of one block: rank (seg_rank_block), per-digit offsets within the
// block, then the block is laid out in digit order in LDS and written back
// run by run, so consecutive lanes store to consecutive addresses.
// VALS: values travel with the keys (vin / vout non-null); a keys-only sort
// stages half as much in LDS, so more workgroups fit on a CU.
template <int DB, bool DROP, bool VALS>
__global__ __launch_bounds__(SEG_THREADS) void k_seg_scatter2(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout,
    uint32_t* __restrict__ vout, long long stride, const uint32_t* __restrict__ d_counts, long long n_fixed,
    int shift, int nbits, const uint32_t* __restrict__ incl, const uint32_t* __restrict__ hist, int nb) {
  constexpr int NB = 1 << DB;
  constexpr int CNT_WORDS = SEG_ROUNDS * 4 * NB / 2;   // u16 table in u32 words
  constexpr int STAGE_WORDS = (VALS ? 2 : 1) * SEG_TILE;
  constexpr int BUF_WORDS = (CNT_WORDS > STAGE_WORDS) ? CNT_WORDS : STAGE_WORDS;
  // [round][wave][digit] u16 counts (block offsets < 2^16); later reused for
  // the block's keys (and values) in digit order
  __shared__ __align__(16) uint32_t s_buf[BUF_WORDS];
  __shared__ uint32_t s_base[NB];    // segment-relative output start of digit d for this block
  __shared__ uint32_t s_bstart[NB];  // start of digit d within the block
  __shared__ uint32_t s_part[SEG_THREADS / NB][NB];
  __shared__ uint32_t s_nkeep;                     // elements of the block that are written
  constexpr bool RANK_BALLOT = DB == 8;
  // [wave][digit] peer masks (one word when the ranks come from ballots)
  __shared__ unsigned long long s_match[RANK_BALLOT ? 1 : 4 * NB];
  uint16_t* s_cnt = reinterpret_cast<uint16_t*>(s_buf);
  const int r = blockIdx.y;
  const int tid = threadIdx.x, wave = tid >> 6;
  const long long rem = seg_count(d_counts, n_fixed, stride, r) - (long long)blockIdx.x * SEG_TILE;
  if (rem <= 0) return;   // uniform over the workgroup
  const bool full = rem >= SEG_TILE;
  const int m = full ? SEG_TILE : (int)rem;
  const int rows = 1 << nbits;
  const uint32_t mask = (uint32_t)rows - 1u;
  const size_t tbl = (size_t)r * rows * nb;
  // inclusive scan over all segments' tables: subtract the running total at
  // this segment's start (u32 arithmetic is exact modulo 2^32)
  const uint32_t seg0 = (r == 0) ? 0u : incl[tbl - 1];
  for (int d = tid; d < NB; d += SEG_THREADS) {
    const size_t h = tbl + (size_t)d * nb + blockIdx.x;
    s_base[d] = (d < rows) ? incl[h] - hist[h] - seg0 : 0u;
  }
  for (int e = tid; e < CNT_WORDS; e += SEG_THREADS) s_buf[e] = 0u;
  const size_t off = (size_t)r * stride + (size_t)blockIdx.x * SEG_TILE;
  const uint32_t* ki = kin + off;
  const uint32_t* vi = VALS ? vin + off : nullptr;
  uint32_t* ko = kout + (size_t)r * stride;
  uint32_t* vo = VALS ? vout + (size_t)r * stride : nullptr;
  uint32_t key[SEG_ROUNDS], val[SEG_ROUNDS], rank[SEG_ROUNDS];
  if (full) {
#pragma unroll
    for (int q = 0; q < SEG_ROUNDS; ++q) key[q] = ki[q * SEG_THREADS + tid];
    if (vi) {
#pragma unroll
      for (int q = 0; q < SEG_ROUNDS; ++q) val[q] = vi[q * SEG_THREADS + tid];
    }
  } else {
#pragma unroll
    for (int q = 0; q < SEG_ROUNDS; ++q) {
      const int i = q * SEG_THREADS + tid;
      key[q] = (i < m) ? ki[i] : 0u;
      val[q] = (vi && i < m) ? vi[i] : 0u;
    }
  }
  __syncthreads();
  if (full)
    seg_rank_block<DB, true, DROP, RANK_BALLOT>(key, rank, m, shift, nbits, mask, s_cnt, s_match);
  else
    seg_rank_block<DB, false, DROP, RANK_BALLOT>(key, rank, m, shift, nbits, mask, s_cnt, s_match);
  __syncthreads();
  // exclusive prefix of each digit over the 64 (round, wave) pairs in element
  // order: PARTS threads per digit, each scanning PER consecutive pairs; the
  // last part also yields the digit's block total
  {
    constexpr int PAIRS = SEG_ROUNDS * 4;
    constexpr int PARTS = SEG_THREADS / NB;
    constexpr int PER = PAIRS / PARTS;
    const int d = tid % NB, part = tid / NB;
    uint32_t c[PER];
#pragma unroll
    for (int k2 = 0; k2 < PER; ++k2) c[k2] = s_cnt[(part * PER + k2) * NB + d];
    uint32_t run = 0;
#pragma unroll
    for (int k2 = 0; k2 < PER; ++k2) {
      const uint32_t t = c[k2];
      c[k2] = run;
      run += t;
    }
    s_part[part][d] = run;
    __syncthreads();
    uint32_t o = 0;
    for (int p2 = 0; p2 < part; ++p2) o += s_part[p2][d];
#pragma unroll
    for (int k2 = 0; k2 < PER; ++k2) s_cnt[(part * PER + k2) * NB + d] = (uint16_t)(c[k2] + o);
    if (part == PARTS - 1) s_bstart[d] = o + run;   // digit total for now
  }
  __syncthreads();
  // exclusive scan of the digit totals (one wave, NB / 64 digits per lane)
  if (wave == 0) {
    constexpr int PL = (NB + 63) / 64;
    const int lane = tid;
    uint32_t t[PL], sum = 0;
#pragma unroll
    for (int k2 = 0; k2 < PL; ++k2) {
      const int d = lane * PL + k2;
      t[k2] = (d < NB) ? s_bstart[d] : 0u;
      sum += t[k2];
    }
    uint32_t incl_s = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(incl_s, o, WAVE);
      if (lane >= o) incl_s += y;
    }
    uint32_t run = incl_s - sum;
#pragma unroll
    for (int k2 = 0; k2 < PL; ++k2) {
      const int d = lane * PL + k2;
      if (d < NB) s_bstart[d] = run;
      run += t[k2];
    }
    if (lane == 63) s_nkeep = incl_s;
  }
  __syncthreads();
  uint32_t lp[SEG_ROUNDS];
#pragma unroll
  for (int q = 0; q < SEG_ROUNDS; ++q) {
    const uint32_t d = (key[q] >> shift) & mask;
    lp[q] = s_bstart[d] + s_cnt[(q * 4 + wave) * NB + d] + rank[q];
  }
  __syncthreads();   // the counts table is overwritten below
#pragma unroll
  for (int q = 0; q < SEG_ROUNDS; ++q) {
    if ((full || q * SEG_THREADS + tid < m) && !(DROP && key[q] == DROPPED_FRAGMENT)) {
      s_buf[lp[q]] = key[q];
      if (vo) s_buf[SEG_TILE + lp[q]] = val[q];
    }
  }
  __syncthreads();
  const int nkeep = (int)s_nkeep;
#pragma unroll 4
  for (int i = tid; i < nkeep; i += SEG_THREADS) {
    const uint32_t k = s_buf[i];
    const uint32_t d = (k >> shift) & mask;
    const uint32_t pos = s_base[d] + ((uint32_t)i - s_bstart[d]);
    ko[pos] = k;
    if (vo) vo[pos] = s_buf[SEG_TILE + i];
  }
}

// Elements each segment kept in a dropping pass: the sum of the segment's
// table rows, from the inclusive scan over all segments' tables.
__global__ __launch_bounds__(256) void k_seg_counts(int R, long long rows_nb, const uint32_t* __restrict__ incl,
                                                    uint32_t* __restrict__ counts) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  const uint32_t lo = (r == 0) ? 0u : incl[(size_t)r * rows_nb - 1];
  counts[r] = incl[(size_t)(r + 1) * rows_nb - 1] - lo;
}

// ---------------------------------------------------------------------------
// fragment emission in depth order
// ---------------------------------------------------------------------------
// Tile counts in depth order: read from the sorted values themselves when the
// preprocess packed them there (count_shift > 0), else gathered by index.
__global__ __launch_bounds__(256) void k_tr_gather_tiles(int R, int P, const uint32_t* __restrict__ dval,
                                                         const uint32_t* __restrict__ tiles,
                                                         uint32_t* __restrict__ out, int count_shift) {
  const int r = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= P) return;
  const size_t base = (size_t)r * P;
  const uint32_t v = dval[base + j];
  out[base + j] = count_shift ? (v >> count_shift) : tiles[base + v];
}

__global__ __launch_bounds__(256) void k_tr_nfrag(int R, int P, const uint32_t* __restrict__ incl,
                                                  long long cap, uint32_t* __restrict__ nfrag,
                                                  int* __restrict__ overflow) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  const uint32_t seg0 = (r == 0) ? 0u : incl[(size_t)r * P - 1];
  const uint32_t n = incl[(size_t)r * P + P - 1] - seg0;
  nfrag[r] = n;
  if ((long long)n > cap) atomicMax(overflow, 1);
}

__global__ __launch_bounds__(256) void k_tr_emit(int R, int P, const uint32_t* __restrict__ dval,
                                                 const uint32_t* __restrict__ incl, const uint2* __restrict__ rects,
                                                 const uint32_t* __restrict__ tiles, int grid_x, long long cap,
                                                 uint32_t* __restrict__ ftile, uint32_t* __restrict__ fgid) {
  const int r = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= P) return;
  const size_t base = (size_t)r * P;
  const uint32_t g = dval[base + j];
  if (tiles[base + g] == 0) return;
  const uint32_t seg0 = (r == 0) ? 0u : incl[base - 1];
  long long off = (j == 0) ? 0 : (long long)(uint32_t)(incl[base + j - 1] - seg0);
  const uint2 rc = rects[base + g];
  const int x0 = rc.x & 0xFFFF, y0 = rc.x >> 16, x1 = rc.y & 0xFFFF, y1 = rc.y >> 16;
  uint32_t* ft = ftile + (size_t)r * cap;
  uint32_t* fg = fgid + (size_t)r * cap;
  for (int y = y0; y < y1; ++y)
    for (int x = x0; x < x1; ++x) {
      if (off < cap) {
        ft[off] = (uint32_t)(y * grid_x + x);
        fg[off] = g;
      }
      ++off;
    }
}

__global__ __launch_bounds__(256) void k_tr_ranges(int R, long long cap, const uint32_t* __restrict__ ftile,
                                                   const uint32_t* __restrict__ nfrag, int n_tiles,
                                                   uint2* __restrict__ ranges) {
  const int r = blockIdx.y;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long n = (long long)nfrag[r];
  if (n > cap) n = cap;
  if (i >= n) return;
  const uint32_t* ft = ftile + (size_t)r * cap;
  uint2* rg = ranges + (size_t)r * n_tiles;
  const uint32_t t = ft[i];
  if (i == 0 || ft[i - 1] != t) rg[t].x = (uint32_t)i;
  if (i == n - 1 || ft[i + 1] != t) rg[t].y = (uint32_t)(i + 1);
}

// ---------------------------------------------------------------------------
// blend forward / backward over renders
// ---------------------------------------------------------------------------
struct TrImg {
  float* inten;          // [R][HW]
  float* final_T;        // [R][HW]
  int* n_contrib;        // [R][HW]
  const float* events;   // [n_windows][HW]
  float* grad6;          // [R][P][6]: mean 2, conic 3, intensity
};

__device__ __forceinline__ BlendArgs tr_blend_args(const Intr& in, int r, int P, long long cap,
                                                   const uint2* ranges, const uint32_t* fgid,
                                                   const float4* splats, const TrImg& im, float bg,
                                                   uint32_t gid_mask) {
  BlendArgs a = {};
  const size_t HW = (size_t)in.W * in.H;
  a.W = in.W; a.H = in.H; a.tile_w = in.tile_w; a.tile_h = in.tile_h; a.grid_x = in.grid_x;
  a.ranges = ranges + (size_t)r * in.grid_x * in.grid_y;
  a.vals = fgid + (size_t)r * cap;
  a.gid_mask = gid_mask;
  a.splats = splats + (size_t)r * P * 2;
  a.bgp = nullptr;
  a.bgv[0] = bg; a.bgv[1] = 0.0f; a.bgv[2] = 0.0f;
  a.out_color = im.inten + (size_t)r * HW;
  a.final_T = im.final_T + (size_t)r * HW;
  a.n_contrib = im.n_contrib + (size_t)r * HW;
  a.grad10 = nullptr;
  return a;
}

__global__ __launch_bounds__(256) void k_tr_blend_fwd(Intr in, int P, long long cap, const uint2* ranges,
                                                      const uint32_t* fgid, const float4* splats, TrImg im,
                                                      float bg, uint32_t gid_mask) {
  const BlendArgs a = tr_blend_args(in, blockIdx.y, P, cap, ranges, fgid, splats, im, bg, gid_mask);
  blend_fwd_tile<1, false>(a, blockIdx.x);
}

// per-item loss sums: S[b] = {sum dI^2, sum u e, sum e^2} with u = dI (signed)
// or |dI| (unsigned), partials per workgroup in FP64
__global__ __launch_bounds__(256) void k_tr_loss_partial(int HW, const float* __restrict__ inten,
                                                         const float* __restrict__ events, const int* __restrict__ window,
                                                         int unsigned_mode, const int* __restrict__ stage,
                                                         double* __restrict__ part) {
  __shared__ double s_w[4][3];
  const int b = blockIdx.y;
  const float* Il = inten + (size_t)(2 * b) * HW;
  const float* In = inten + (size_t)(2 * b + 1) * HW;
  const float* e = events + (size_t)window[b] * HW;
  // unsigned: |dI| against |e| (the |.| of the signed event frame is the
  // unsigned one, tracker.py:78-91 and event.py:126-127)
  const bool uns = unsigned_mode < 0 ? stage[b] == 0 : unsigned_mode != 0;
  double a0 = 0, a1 = 0, a2 = 0;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < HW; p += gridDim.x * blockDim.x) {
    const double d = (double)In[p] - (double)Il[p];
    const double u = uns ? fabs(d) : d;
    const double ev = uns ? fabs((double)e[p]) : (double)e[p];
    a0 += d * d;
    a1 += u * ev;
    a2 += ev * ev;
  }
  for (int o = 32; o >= 1; o >>= 1) {
    a0 += __shfl_xor(a0, o, WAVE);
    a1 += __shfl_xor(a1, o, WAVE);
    a2 += __shfl_xor(a2, o, WAVE);
  }
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { s_w[wave][0] = a0; s_w[wave][1] = a1; s_w[wave][2] = a2; }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int k = threadIdx.x;
    part[((size_t)b * gridDim.x + blockIdx.x) * 3 + k] = s_w[0][k] + s_w[1][k] + s_w[2][k] + s_w[3][k];
  }
}

// scal[b] = {1/n, L, (1 - S2/n)/L, 1/L}: n = ||dI||, L = || u/n - e ||
__global__ __launch_bounds__(64) void k_tr_loss_final(int B, int nparts, const double* __restrict__ part,
                                                      float* __restrict__ scal, ItemState st) {
  const int b = blockIdx.x;
  const int lane = threadIdx.x;
  double s0 = 0, s1 = 0, s2 = 0;
  for (int k = lane; k < nparts; k += WAVE) {
    s0 += part[((size_t)b * nparts + k) * 3];
    s1 += part[((size_t)b * nparts + k) * 3 + 1];
    s2 += part[((size_t)b * nparts + k) * 3 + 2];
  }
  for (int o = 32; o >= 1; o >>= 1) {
    s0 += __shfl_xor(s0, o, WAVE);
    s1 += __shfl_xor(s1, o, WAVE);
    s2 += __shfl_xor(s2, o, WAVE);
  }
  if (lane != 0) return;
  const double n = sqrt(s0);
  const double re = n > 0 ? s1 / n : 0.0;
  // ||dI / n - e||^2 with ||dI / n|| = 1; a render that does not change at
  // all (n == 0) contributes nothing, leaving ||e|| (the reference's 0 / 0
  // would give NaN) and a zero gradient (invn = 0 below)
  double L2 = n > 0 ? 1.0 - 2.0 * re + s2 : s2;
  if (L2 < 0) L2 = 0;
  const double L = sqrt(L2);
  const double invn = n > 0 ? 1.0 / n : 0.0;
  const double invL = L > 0 ? 1.0 / L : 0.0;
  scal[4 * b + 0] = (float)invn;
  scal[4 * b + 1] = (float)L;
  scal[4 * b + 2] = (float)((1.0 - re) * invL);
  scal[4 * b + 3] = (float)invL;
  st.loss[b] = (float)L;
  const int it = st.iter[b];
  if (it >= 0 && it < st.hist_len) st.loss_hist[(size_t)b * st.hist_len + it] = (float)L;
}

__device__ __forceinline__ void tr_blend_bwd_item(const Intr& in, int r, int tile, int P, long long cap,
                                                  const uint2* ranges, const uint32_t* fgid, const float4* splats,
                                                  const TrImg& im, const int* __restrict__ window,
                                                  const float* __restrict__ scal, int unsigned_mode,
                                                  const int* __restrict__ stage, float bg, uint32_t gid_mask) {
  const int b = r >> 1;
  const float sign = (r & 1) ? 1.0f : -1.0f;
  BlendArgs a = tr_blend_args(in, r, P, cap, ranges, fgid, splats, im, bg, gid_mask);
  a.grad10 = nullptr;
  const int tid = threadIdx.x;
  int px, py;
  tile_pixel(tile, tid, a.grid_x, a.tile_w, a.tile_h, px, py);
  const bool inside = px < a.W && py < a.H;
  const size_t HW = (size_t)a.W * a.H;
  float g = 0.0f;
  if (inside) {
    const int pix = py * a.W + px;
    const float invn = scal[4 * b], rg = scal[4 * b + 2], invL = scal[4 * b + 3];
    const float d = im.inten[(size_t)(2 * b + 1) * HW + pix] - im.inten[(size_t)(2 * b) * HW + pix];
    const float rr = d * invn;
    const float e = im.events[(size_t)window[b] * HW + pix];
    float G;
    if (unsigned_mode < 0 ? stage[b] == 0 : unsigned_mode != 0) {
      const float u = fabsf(rr);
      G = (rr > 0.0f ? 1.0f : (rr < 0.0f ? -1.0f : 0.0f)) * (u - fabsf(e)) * invL;
    } else {
      G = (rr - e) * invL;
    }
    g = sign * (G - rr * rg) * invn;
  }
  // per-Gaussian gradients of this render land in grad6[r][P][6] (the lean
  // one-channel body writes stride 6)
  a.grad10 = im.grad6 + (size_t)r * P * 6;
  float gC[1] = {g};
  blend_bwd_tile<1, true>(a, tile, gC, 0.0f, 0.0f);
}

__global__ __launch_bounds__(256) void k_tr_blend_bwd(Intr in, int P, long long cap, const uint2* ranges,
                                                      const uint32_t* fgid, const float4* splats, TrImg im,
                                                      const int* __restrict__ window, const float* __restrict__ scal,
                                                      int unsigned_mode, const int* __restrict__ stage, float bg,
                                                      uint32_t gid_mask) {
  tr_blend_bwd_item(in, blockIdx.y, blockIdx.x, P, cap, ranges, fgid, splats, im, window, scal, unsigned_mode, stage,
                    bg, gid_mask);
}

// Persistent form of both blends: a fixed grid of workgroups pulls (render,
// tile) items from device counters until every item is taken.  With nq = 8
// queues, queue q holds the renders r with r % 8 == q and workgroup w starts on
// queue w % 8 (the XCD the hardware places it on, as workgroups are dealt to
// the 8 XCDs round-robin), so a render's records and fragment lists are read
// through one XCD's L2; a workgroup whose queue is empty takes items from the
// next queue.  Queue q hands out k = 0, 1, ... as render q + nq (k / n_tiles),
// tile k % n_tiles; a k past the last render means the queue is empty.
// Every workgroup leaves once it has found all nq queues empty (each probe
// moves a counter forward, so no wait is involved).
__device__ __forceinline__ int tr_queue_next(uint32_t* queue, int nq, int R, int n_tiles, int& q, int& exhausted) {
  while (exhausted < nq) {
    const uint32_t k = atomicAdd(&queue[q], 1u);
    const uint32_t rr = (uint32_t)q + (uint32_t)nq * (k / (uint32_t)n_tiles);
    if (rr < (uint32_t)R) return (int)(rr * (uint32_t)n_tiles + k % (uint32_t)n_tiles);
    q = (q + 1) % nq;
    ++exhausted;
  }
  return -1;
}

constexpr int TR_QUEUES = 8;

// Eight waves per SIMD: the persistent backward then fits 64 VGPRs (71
// otherwise, seven waves) without scratch; measured -0.9% on the backward and
// -1.3% on the forward (the grid kernels fit eight waves as they are).
template <bool BWD>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_tr_blend_queue(Intr in, int R, int P, long long cap, const uint2* ranges,
                                                        const uint32_t* fgid, const float4* splats, TrImg im,
                                                        const int* __restrict__ window,
                                                        const float* __restrict__ scal, int unsigned_mode,
                                                        const int* __restrict__ stage, float bg, uint32_t gid_mask,
                                                        uint32_t* queue, int nq) {
  __shared__ int s_item;
  const int n_tiles = in.grid_x * in.grid_y;
  int q = (int)(blockIdx.x % (unsigned)nq), exhausted = 0;   // used by thread 0 only
  for (;;) {
    if (threadIdx.x == 0) s_item = tr_queue_next(queue, nq, R, n_tiles, q, exhausted);
    __syncthreads();
    const int item = s_item;
    __syncthreads();   // every wave has finished the previous item and read this one
    if (item < 0) break;
    const int r = item / n_tiles, tile = item - r * n_tiles;
    if (BWD) {
      tr_blend_bwd_item(in, r, tile, P, cap, ranges, fgid, splats, im, window, scal, unsigned_mode, stage, bg,
                        gid_mask);
    } else {
      const BlendArgs a = tr_blend_args(in, r, P, cap, ranges, fgid, splats, im, bg, gid_mask);
      blend_fwd_tile<1, false>(a, tile);
    }
  }
}

// ---------------------------------------------------------------------------
// preprocess backward -> 18 pose partials per (workgroup, render)
// ---------------------------------------------------------------------------
template <int D>
__device__ __forceinline__ void tr_pose_partials(const TrMap& m, const CamDev& c, const float* mu, const float* sig,
                                                 const float* sh, const float* g, float* part) {
  constexpr int K = (D + 1) * (D + 1);
  const float gm0 = g[0], gm1 = g[1], ga = g[2], gb = g[3], gc = g[4], gI = g[5];
  Fwd f;
  project_one(mu, sig, c, f);
  const float A = f.cov[0], B = f.cov[1], Cc = f.cov[2];
  const float d2 = f.det * f.det;
  const float gA = (ga * (-Cc * Cc) + gb * (B * Cc) + gc * (-B * B)) / d2;
  const float gB = (ga * (2.0f * B * Cc) + gb * (-(A * Cc + B * B)) + gc * (2.0f * A * B)) / d2;
  const float gCc = (ga * (-B * B) + gb * (A * B) + gc * (-A * A)) / d2;
  const float G2[2][2] = {{gA, 0.5f * gB}, {0.5f * gB, gCc}};
  float gM2[2][3];
#pragma unroll
  for (int q = 0; q < 2; ++q)
#pragma unroll
    for (int k = 0; k < 3; ++k) gM2[q][k] = 2.0f * (G2[q][0] * f.B[0][k] + G2[q][1] * f.B[1][k]);
  float Wm[3][3];
#pragma unroll
  for (int q = 0; q < 3; ++q)
#pragma unroll
    for (int k = 0; k < 3; ++k) Wm[q][k] = mat(c.view, q, k);
  const float gJ00 = gM2[0][0] * Wm[0][0] + gM2[0][1] * Wm[0][1] + gM2[0][2] * Wm[0][2];
  const float gJ02 = gM2[0][0] * Wm[2][0] + gM2[0][1] * Wm[2][1] + gM2[0][2] * Wm[2][2];
  const float gJ11 = gM2[1][0] * Wm[1][0] + gM2[1][1] * Wm[1][1] + gM2[1][2] * Wm[1][2];
  const float gJ12 = gM2[1][0] * Wm[2][0] + gM2[1][1] * Wm[2][1] + gM2[1][2] * Wm[2][2];
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
  const float ghx = gm0 * (0.5f * (float)c.W) * f.pw;
  const float ghy = gm1 * (0.5f * (float)c.H) * f.pw;
  const float ghw = -(gm0 * (0.5f * (float)c.W) * f.hx + gm1 * (0.5f * (float)c.H) * f.hy) * f.pw * f.pw;
#pragma unroll
  for (int k = 0; k < 3; ++k)
    gpv[k] += ghx * mat(c.projraw, 0, k) + ghy * mat(c.projraw, 1, k) + ghw * mat(c.projraw, 3, k);
  float gcol[3] = {m.iw[0] * gI, m.iw[1] * gI, m.iw[2] * gI};
  float dir[3], dlen, rgb[3];
  bool cl[3];
  sh_color_reg<D>(mu, sh, c, dir, dlen, rgb, cl);
#pragma unroll
  for (int ch = 0; ch < 3; ++ch)
    if (cl[ch]) gcol[ch] = 0.0f;
  float gbasis[16];
#pragma unroll
  for (int k = 0; k < K; ++k) gbasis[k] = gcol[0] * sh[3 * k] + gcol[1] * sh[3 * k + 1] + gcol[2] * sh[3 * k + 2];
  float gdir[3] = {0.0f, 0.0f, 0.0f};
  sh_basis_grad(D, dir[0], dir[1], dir[2], gbasis, gdir);
  const float dd = dir[0] * gdir[0] + dir[1] * gdir[1] + dir[2] * gdir[2];
  float gd[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) gd[k] = (gdir[k] - dir[k] * dd) / dlen;
  part[0] += gpv[0]; part[1] += gpv[1]; part[2] += gpv[2];
  part[3] += f.pv[1] * gpv[2] - f.pv[2] * gpv[1];
  part[4] += f.pv[2] * gpv[0] - f.pv[0] * gpv[2];
  part[5] += f.pv[0] * gpv[1] - f.pv[1] * gpv[0];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    part[6 + k] += f.J00 * gM2[0][k];
    part[9 + k] += f.J11 * gM2[1][k];
    part[12 + k] += f.J02 * gM2[0][k] + f.J12 * gM2[1][k];
  }
  part[15] -= gd[0]; part[16] -= gd[1]; part[17] -= gd[2];
}

// One thread per Gaussian, RPT renders per thread; per render a workgroup
// reduction of the 18 pose partials.
template <int D>
__global__ __launch_bounds__(256) void k_tr_preprocess_bwd(TrMap m, const CamDev* __restrict__ cams, int R,
                                                           float* __restrict__ grad6, float* __restrict__ partials,
                                                           int clear) {
  constexpr int K = (D + 1) * (D + 1);
  __shared__ float s_part[RPT][4][18];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const bool valid = i < m.P;
  const int r0 = blockIdx.y * RPT;
  const int r1 = (r0 + RPT < R) ? r0 + RPT : R;
  float mu[3] = {0, 0, 0}, sig[6] = {0, 0, 0, 0, 0, 0};
  float sh[3 * K];
  if (valid) {
    mu[0] = m.means[3 * i]; mu[1] = m.means[3 * i + 1]; mu[2] = m.means[3 * i + 2];
#pragma unroll
    for (int k = 0; k < 6; ++k) sig[k] = m.cov[6 * i + k];
#pragma unroll
    for (int k = 0; k < 3 * K; ++k) sh[k] = m.shs[(size_t)k * m.P + i];
  }
  // gradients of the next render are loaded while this one is processed; a
  // Gaussian with all-zero gradients (culled, or reaching no pixel) adds
  // nothing.  Each render's wave totals go to LDS; the workgroup sums them
  // once, after the last render.
  float gn[6] = {0, 0, 0, 0, 0, 0};
  if (valid) {
#pragma unroll
    for (int k = 0; k < 6; ++k) gn[k] = grad6[((size_t)r0 * m.P + i) * 6 + k];
  }
  for (int r = r0; r < r1; ++r) {
    float gq[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) gq[k] = gn[k];
    if (valid && r + 1 < r1) {
#pragma unroll
      for (int k = 0; k < 6; ++k) gn[k] = grad6[((size_t)(r + 1) * m.P + i) * 6 + k];
    }
    float part[18];
#pragma unroll
    for (int k = 0; k < 18; ++k) part[k] = 0.0f;
    if (gq[0] != 0.0f || gq[1] != 0.0f || gq[2] != 0.0f || gq[3] != 0.0f || gq[4] != 0.0f || gq[5] != 0.0f) {
      tr_pose_partials<D>(m, cams[r], mu, sig, sh, gq, part);
      if (clear) {   // leave the buffer zero for the next backward blend
        float2* g2 = reinterpret_cast<float2*>(grad6 + ((size_t)r * m.P + i) * 6);
        g2[0] = make_float2(0.0f, 0.0f);
        g2[1] = make_float2(0.0f, 0.0f);
        g2[2] = make_float2(0.0f, 0.0f);
      }
    }
    float v16[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) v16[k] = part[k];
    const float s16 = wave_reduce16(v16, lane);
    const float s16b = wave_sum(part[16]);
    const float s17 = wave_sum(part[17]);
    if ((lane & 3) == 0) s_part[r - r0][wave][reduce16_slot(lane)] = s16;
    if (lane == 0) { s_part[r - r0][wave][16] = s16b; s_part[r - r0][wave][17] = s17; }
  }
  __syncthreads();
  for (int e = tid; e < (r1 - r0) * 18; e += blockDim.x) {
    const int rr = e / 18, k = e % 18;
    partials[((size_t)(r0 + rr) * gridDim.x + blockIdx.x) * 18 + k] =
        s_part[rr][0][k] + s_part[rr][1][k] + s_part[rr][2][k] + s_part[rr][3][k];
  }
}

// ---------------------------------------------------------------------------
// per-item adjoint + Adam + pose/velocity update (one wave per item)
// ---------------------------------------------------------------------------
struct OptCfg {
  float lr[4];          // base LR: theta (rot), rho (trans), w, v
  int max_optim_iter;   // LR ramp length = max_optim_iter / 2
  int opt_vel;          // 1: fine stage (all four groups, LR ramp), 0: coarse (theta, rho only)
  int start_vel_iter;   // start of the fine stage inside the level
  int staged;           // 1: opt_vel / start_vel_iter per item from st.stage / st.start_vel
  float conv_thr;       // staged: convergence threshold (tracker.py:65-76)
};

__device__ void se3_jl_T_times(const double* xi, const double* g, double* out) {
  double ad[6][6];
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 6; ++c) ad[r][c] = 0.0;
  for (int blk = 0; blk < 3; ++blk) {
    const int r0 = (blk == 2) ? 3 : 0, c0 = (blk == 0) ? 0 : 3;
    const double* v = (blk == 1) ? xi : xi + 3;
    ad[r0 + 0][c0 + 1] = -v[2]; ad[r0 + 0][c0 + 2] = v[1];
    ad[r0 + 1][c0 + 0] = v[2];  ad[r0 + 1][c0 + 2] = -v[0];
    ad[r0 + 2][c0 + 0] = -v[1]; ad[r0 + 2][c0 + 1] = v[0];
  }
  double term[6], acc[6];
  for (int k = 0; k < 6; ++k) { term[k] = g[k]; acc[k] = g[k]; }
  for (int n = 1; n < 30; ++n) {
    double nt[6], mag = 0.0;
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

__global__ __launch_bounds__(64) void k_tr_update(int B, int nblk, const float* __restrict__ partials,
                                                  const CamDev* __restrict__ cams, const float* __restrict__ vel,
                                                  const float* __restrict__ dts, ItemState st, OptCfg cfg,
                                                  int do_step) {
  const int b = blockIdx.x;
  const int lane = threadIdx.x;
  double grad[12];
  for (int k = 0; k < 12; ++k) grad[k] = 0.0;
  for (int side = 0; side < 2; ++side) {
    const int r = 2 * b + side;
    double s[18];
    for (int k = 0; k < 18; ++k) s[k] = 0.0;
    for (int q = lane; q < nblk; q += WAVE)
      for (int k = 0; k < 18; ++k) s[k] += (double)partials[((size_t)r * nblk + q) * 18 + k];
    for (int k = 0; k < 18; ++k)
      for (int o = 32; o >= 1; o >>= 1) s[k] += __shfl_xor(s[k], o, WAVE);
    const float* view = cams[r].view;
    double Wm[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) Wm[i][j] = (double)view[4 * j + i];
    double g[6];
    for (int i = 0; i < 3; ++i) g[i] = s[i] - (Wm[i][0] * s[15] + Wm[i][1] * s[16] + Wm[i][2] * s[17]);
    const double* GW = s + 6;
    double f0 = 0, f1 = 0, f2 = 0;
    for (int j = 0; j < 3; ++j) {
      f0 += -GW[3 + j] * Wm[2][j] + GW[6 + j] * Wm[1][j];
      f1 += GW[j] * Wm[2][j] - GW[6 + j] * Wm[0][j];
      f2 += -GW[j] * Wm[1][j] + GW[3 + j] * Wm[0][j];
    }
    g[3] = s[3] + f0; g[4] = s[4] + f1; g[5] = s[5] + f2;
    const float* E = vel + 16 * r;
    double RE[3][3], tE[3];
    for (int i = 0; i < 3; ++i) {
      for (int j = 0; j < 3; ++j) RE[i][j] = (double)E[4 * j + i];
      tE[i] = (double)E[12 + i];
    }
    const double tg[3] = {tE[1] * g[2] - tE[2] * g[1], tE[2] * g[0] - tE[0] * g[2], tE[0] * g[1] - tE[1] * g[0]};
    for (int c = 0; c < 3; ++c) {
      double rho = 0, th = 0;
      for (int i = 0; i < 3; ++i) {
        rho += RE[i][c] * g[i];
        th += RE[i][c] * (g[3 + i] - tg[i]);
      }
      grad[3 + c] += rho;
      grad[c] += th;
    }
    const double dt = (double)dts[r];
    double xi[6];
    for (int k = 0; k < 3; ++k) { xi[k] = dt * (double)st.lin[3 * b + k]; xi[3 + k] = dt * (double)st.ang[3 * b + k]; }
    double jg[6];
    se3_jl_T_times(xi, g, jg);
    for (int k = 0; k < 3; ++k) { grad[9 + k] += dt * jg[k]; grad[6 + k] += dt * jg[3 + k]; }
  }
  if (lane != 0) return;
  for (int k = 0; k < 12; ++k) st.grad12[12 * b + k] = (float)grad[k];
  if (!do_step) return;
  const int stage = cfg.staged ? st.stage[b] : (cfg.opt_vel ? 1 : 0);
  if (stage == 2) return;   // done: no further updates
  const int opt_vel = stage == 1;
  const int start_vel = cfg.staged ? st.start_vel[b] : cfg.start_vel_iter;
  // learning-rate schedule of the fine stage
  const int it = st.iter[b];
  float lr_scale[4] = {1.0f, 1.0f, 1.0f, 1.0f};
  bool active[4] = {true, true, opt_vel != 0, opt_vel != 0};
  if (opt_vel) {
    const float fn = (float)cfg.max_optim_iter / 2.0f;
    const int k = it - start_vel;
    const float frac = (1 <= k && (float)k <= fn) ? (float)k / fn : 1.0f;
    lr_scale[0] = frac; lr_scale[1] = frac; lr_scale[2] = 1.0f - frac; lr_scale[3] = 1.0f - frac;
  }
  // Adam (beta 0.9 / 0.999, eps 1e-8) on parameters that are reset to 0 after
  // every step, as utils/render_camera/camera.py:146-155 does
  float delta[12];
  for (int grp = 0; grp < 4; ++grp) {
    if (!active[grp]) {
      for (int k = 0; k < 3; ++k) delta[3 * grp + k] = 0.0f;
      continue;
    }
    const int t = ++st.adam_t[4 * b + grp];
    const float bc1 = 1.0f - powf(0.9f, (float)t);
    const float bc2 = 1.0f - powf(0.999f, (float)t);
    const float step = (cfg.lr[grp] * lr_scale[grp]) / bc1;
    const float bc2s = sqrtf(bc2);
    for (int k = 0; k < 3; ++k) {
      const int q = 12 * b + 3 * grp + k;
      const float gk = (float)grad[3 * grp + k];
      const float m = 0.9f * st.adam_m[q] + (1.0f - 0.9f) * gk;
      const float v = 0.999f * st.adam_v[q] + (1.0f - 0.999f) * gk * gk;
      st.adam_m[q] = m;
      st.adam_v[q] = v;
      const float denom = sqrtf(v) / bc2s + 1e-8f;
      delta[3 * grp + k] = -step * (m / denom);
    }
  }
  // velocity first, then pose (update_vwRT)
  for (int k = 0; k < 3; ++k) {
    st.ang[3 * b + k] += delta[6 + k];
    st.lin[3 * b + k] += delta[9 + k];
  }
  double xi[6] = {delta[3], delta[4], delta[5], delta[0], delta[1], delta[2]};
  double DR[3][3], Dt[3];
  se3_exp_d(xi, DR, Dt);
  double* pc = st.pose + 12 * b;
  double nR[9], nt[3];
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) nR[3 * i + j] = DR[i][0] * pc[j] + DR[i][1] * pc[3 + j] + DR[i][2] * pc[6 + j];
    nt[i] = DR[i][0] * pc[9] + DR[i][1] * pc[10] + DR[i][2] * pc[11] + Dt[i];
  }
  for (int k = 0; k < 9; ++k) pc[k] = nR[k];
  for (int k = 0; k < 3; ++k) pc[9 + k] = nt[k];
  if (cfg.staged) {
    // tracker.py:213-236: the convergence test on this level's losses (the
    // loss of iteration it is loss_hist[it]), then the stage switch or stop,
    // then the iteration caps
    bool conv = false;
    if (it >= 10 && it < st.hist_len) {
      const float* h = st.loss_hist + (size_t)b * st.hist_len;
      double sum = 0.0;
      for (int k = it - 9; k <= it; ++k) sum += fabs((double)h[k] - (double)h[k - 1]);
      conv = sum / 10.0 < (double)cfg.conv_thr;
    }
    int next = stage;
    if (conv) {
      if (stage == 0) { next = 1; st.start_vel[b] = it; }
      else next = 2;
    }
    if (next == 0 && it >= cfg.max_optim_iter) next = 2;
    if (next == 1 && conv == false && it >= st.start_vel[b] + cfg.max_optim_iter) next = 2;
    st.stage[b] = next;
  }
  st.iter[b] = it + 1;
}

constexpr int LOSS_BLOCKS = 64;

// Grid-stride fills (any n; the grid is capped so the launch stays within
// the total-thread limit).  k_fill_f32 stores float4 where p is 16 B aligned.
__global__ __launch_bounds__(256) void k_fill_f32(float* p, long long n, float v) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  const long long t0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if ((reinterpret_cast<uintptr_t>(p) & 15) == 0) {
    float4* q = reinterpret_cast<float4*>(p);
    const long long n4 = n >> 2;
    for (long long i = t0; i < n4; i += stride) q[i] = make_float4(v, v, v, v);
    for (long long i = (n4 << 2) + t0; i < n; i += stride) p[i] = v;
  } else {
    for (long long i = t0; i < n; i += stride) p[i] = v;
  }
}

__global__ __launch_bounds__(256) void k_fill_u32b(uint32_t* p, long long n, uint32_t v) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

}  // namespace

// Packed fragments (tile << gid_bits | gid, when both fit in 32 bits) of 256
// depth-consecutive Gaussians per workgroup.  Every tile of a Gaussian's rect
// has a slot; the block's slots [lo, hi) are produced in chunks of EMIT_CHUNK:
// each Gaussian marks the first slot of its segment inside the chunk, an
// inclusive prefix maximum over the chunk gives every slot its Gaussian, and
// consecutive threads then write consecutive slots.  With splats given, a slot
// whose tile the splat cannot reach (alpha < 1/255 at every pixel, the blend's
// own conservative cull on the whole tile) receives DROPPED_FRAGMENT, which
// the first pass of the tile sort removes.
constexpr int EMIT_CHUNK = 2048;

// Workgroups are dealt to the 8 XCDs round-robin (linear id w runs on XCD
// w % 8).  This maps w to a logical workgroup so that each XCD runs one
// contiguous range of logical workgroups, in order: neighbouring workgroups,
// which gather neighbouring splat records, then share one XCD's L2 instead of
// fetching the same 128-byte lines into eight.  A bijection on [0, total).
__device__ __forceinline__ uint32_t xcd_contiguous(uint32_t w, uint32_t total) {
  constexpr uint32_t NX = 8;
  const uint32_t q = total / NX, rem = total % NX, x = w % NX;
  return x * q + (x < rem ? x : rem) + w / NX;
}

__global__ __launch_bounds__(256) void k_tr_emit_packed(int R, int P, const uint32_t* __restrict__ dval,
                                                        uint32_t dval_mask, const uint32_t* __restrict__ incl,
                                                        const float4* __restrict__ splats, int cull, int grid_x,
                                                        int tile_w, int tile_h, int gid_bits, long long cap,
                                                        uint32_t* __restrict__ frag, uint32_t* __restrict__ sort_hist,
                                                        int hist_bits, int hist_nb) {
  constexpr int PER = EMIT_CHUNK / 256;
  static_assert(EMIT_CHUNK <= SEG_TILE, "a chunk must meet at most two sort blocks");
  __shared__ uint32_t s_off[257];
  __shared__ uint32_t s_g[256];
  // per Gaussian, one 16-byte read per slot: first slot (block-relative),
  // rect x0 | y0 << 16, rect x1 | cull kind << 16 (0: reaches no pixel, 1:
  // test, 2: open conic), 1 / rect width
  __shared__ uint4 s_info[256];
  // owner of each slot of the chunk as its Gaussian's block index + 1 (0:
  // none yet); 16 bits keep the workgroup at 18 KB of LDS, eight per CU
  __shared__ uint16_t s_own[EMIT_CHUNK];
  __shared__ int s_wmax[4];
  __shared__ float4 s_r0[256];   // mean x, y; the band test's ey, mg
  __shared__ float4 s_cc[256];   // the band test's p, u, w, kr
  // sort_hist: the kept fragments by the tile sort's first digit, for a
  // window of HIST_WIN consecutive sort blocks (a chunk meets at most two)
  constexpr int HIST_WIN = 4;
  __shared__ uint32_t s_hc[HIST_WIN * 64];
  // logical workgroup: render r, depth-ordered Gaussians [j0, j0 + 256)
  const uint32_t lb = xcd_contiguous(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
  const int r = (int)(lb / gridDim.x);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int j0 = (int)(lb % gridDim.x) * 256;
  const int j = j0 + tid;
  const size_t base = (size_t)r * P;
  const uint32_t seg0 = (r == 0) ? 0u : incl[base - 1];
  if (j < P) {
    const uint32_t g = dval[base + j] & dval_mask;
    s_g[tid] = g;
    const uint32_t off = (j == 0) ? 0u : (uint32_t)(incl[base + j - 1] - seg0);
    s_off[tid] = off;
    // records of Gaussians with no fragments may be stale; nothing below reads
    // them (such a Gaussian owns no slot)
    const float4 r0 = splats[2 * (base + g)], r1 = splats[2 * (base + g) + 1];
    const uint32_t rect8 = __float_as_uint(r1.z);
    const uint32_t x0 = rect8 & 0xFFu, y0 = (rect8 >> 8) & 0xFFu, x1 = (rect8 >> 16) & 0xFFu;
    uint32_t kind = 1u;
    if (cull) {
      const BandRec br = band_prep(r0, r1);
      kind = (uint32_t)br.q3.z;
      s_r0[tid] = make_float4(r0.x, r0.y, br.q3.x, br.q3.y);
      s_cc[tid] = br.q2;
    }
    // 1 / rect width (approximate reciprocal; the row division below stays exact)
    s_info[tid] = make_uint4(off, x0 | (y0 << 16), x1 | (kind << 16),
                             __float_as_uint(__builtin_amdgcn_rcpf((float)(x1 - x0))));
  } else {
    s_off[tid] = (uint32_t)(incl[base + P - 1] - seg0);
  }
  if (tid == 255) {
    const int jl = (j0 + 255 < P) ? j0 + 255 : P - 1;
    s_off[256] = (uint32_t)(incl[base + jl] - seg0);
  }
  __syncthreads();
  const uint32_t lo = s_off[0];
  const uint32_t hi = ((long long)s_off[256] < cap) ? s_off[256] : (uint32_t)cap;
  uint32_t* out = frag + (size_t)r * cap;
  // this workgroup's share of the window's rows (a block that straddles
  // workgroups collects several shares); every count lies below cap
  uint32_t wb = lo / SEG_TILE;   // first sort block of the window
  const auto flush_hist = [&]() {
    for (int e = tid; e < HIST_WIN * 64; e += 256) {
      const uint32_t n = s_hc[e];
      if (n) atomicAdd(&sort_hist[((size_t)r * (1u << hist_bits) + (e & 63)) * hist_nb + wb + (e >> 6)], n);
      s_hc[e] = 0u;
    }
  };
  for (int e = tid; e < HIST_WIN * 64; e += 256) s_hc[e] = 0u;
  for (uint32_t c = lo; c < hi; c += EMIT_CHUNK) {
    const uint32_t ce = (hi - c > (uint32_t)EMIT_CHUNK) ? c + EMIT_CHUNK : hi;
    if (sort_hist && (ce - 1) / SEG_TILE - wb >= (uint32_t)HIST_WIN) {   // uniform
      __syncthreads();
      flush_hist();
      wb = c / SEG_TILE;
    }
#pragma unroll
    for (int i = 0; i < PER; ++i) s_own[i * 256 + tid] = 0;
    __syncthreads();
    {
      const uint32_t b0 = s_off[tid], b1 = s_off[tid + 1];
      if (b0 < b1 && b1 > c && b0 < ce) s_own[(b0 > c ? b0 : c) - c] = (uint16_t)(tid + 1);
    }
    __syncthreads();
    // inclusive prefix maximum: PER consecutive entries per thread, then the
    // maxima of the earlier threads (wave scan, then earlier waves)
    int v[PER];
    int m = 0;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int x = s_own[tid * PER + i];
      m = x > m ? x : m;
      v[i] = m;
    }
    int run = m;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(run, o, WAVE);
      if (lane >= o) run = y > run ? y : run;
    }
    if (lane == 63) s_wmax[wave] = run;
    int prev = __shfl_up(run, 1, WAVE);
    if (lane == 0) prev = 0;
    __syncthreads();
    for (int w2 = 0; w2 < wave; ++w2) prev = s_wmax[w2] > prev ? s_wmax[w2] : prev;
#pragma unroll
    for (int i = 0; i < PER; ++i) s_own[tid * PER + i] = (uint16_t)(v[i] > prev ? v[i] : prev);
    __syncthreads();
    for (uint32_t k = c + tid; k < ce; k += 256) {
      // every slot of the chunk has its owner marked at or before it (an
      // owner whose slots began in an earlier chunk at the chunk's first slot)
      const int a = (int)s_own[k - c] - 1;
      const uint4 inf = s_info[a];
      const uint32_t q = k - inf.x;
      const uint32_t x0 = inf.y & 0xFFFF, y0 = inf.y >> 16, x1 = inf.z & 0xFFFF, kind = inf.z >> 16;
      // q / w: (q + 0.5) / w lies at least 0.5 / w from an integer and the float
      // product errs by < 2^-22 (q + 0.5) / w, so the truncation is exact for q < 2^21
      uint32_t qy = (uint32_t)(((float)q + 0.5f) * __uint_as_float(inf.w));
      if (q >= (1u << 21)) qy = q / (x1 - x0);
      const uint32_t tx = x0 + (q - qy * (x1 - x0)), ty = y0 + qy;
      const uint32_t tile = ty * (uint32_t)grid_x + tx;
      bool keep = true;
      if (cull) {
        // the blend's cull, in its band form, on the tile's rectangle of pixel
        // centres: a dropped fragment has alpha < 1/255 at every pixel of its tile
        keep = kind != 0u;
        if (kind == 1u) {
          Strip sb;
          sb.x0 = (float)(tx * (uint32_t)tile_w);
          sb.x1 = sb.x0 + (float)(tile_w - 1);
          sb.y0 = (float)(ty * (uint32_t)tile_h);
          sb.y1 = sb.y0 + (float)(tile_h - 1);
          const float4 m = s_r0[a];
          keep = band_reach(m.x, m.y, s_cc[a], m.z, m.w, sb);
        }
      }
      out[k] = keep ? ((tile << gid_bits) | s_g[a]) : DROPPED_FRAGMENT;
      if (sort_hist && keep)
        atomicAdd(&s_hc[((k / SEG_TILE - wb) << 6) | (tile & ((1u << hist_bits) - 1u))], 1u);
    }
    __syncthreads();
  }
  if (sort_hist) flush_hist();
}

// Tile ranges of tile-sorted packed fragments by binary search, one thread
// per (render, tile): [x, y) runs from the first fragment whose tile is >= t to
// the first whose tile is > t; an empty tile gets (0, 0).
__global__ __launch_bounds__(256) void k_tr_ranges_packed(int R, long long cap, const uint32_t* __restrict__ frag,
                                                          const uint32_t* __restrict__ nfrag, int gid_bits,
                                                          int n_tiles, uint2* __restrict__ ranges) {
  const int r = blockIdx.y;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_tiles) return;
  long long n = (long long)nfrag[r];
  if (n > cap) n = cap;
  const uint32_t* f = frag + (size_t)r * cap;
  uint32_t bound[2];
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const uint32_t tt = (uint32_t)t + (uint32_t)e;
    uint32_t lo = 0, hi = (uint32_t)n;
    while (lo < hi) {
      const uint32_t mid = lo + ((hi - lo) >> 1);
      if ((f[mid] >> gid_bits) < tt) lo = mid + 1; else hi = mid;
    }
    bound[e] = lo;
  }
  ranges[(size_t)r * n_tiles + t] = (bound[0] < bound[1]) ? make_uint2(bound[0], bound[1]) : make_uint2(0u, 0u);
}

// ===========================================================================
// C ABI (one entry point per stage; the host sequences them, the sequence is
// free of host synchronisation and can be captured in a graph)
// ===========================================================================
// gsevt_item_state (per-item state arrays) is declared in gsevt.h.

namespace {

ItemState to_state(const gsevt_item_state* s) {
  ItemState st;
  st.pose = s->pose; st.lin = s->lin; st.ang = s->ang; st.adam_m = s->adam_m; st.adam_v = s->adam_v;
  st.adam_t = s->adam_t; st.iter = s->iter; st.dtau = s->dtau; st.window = s->window; st.loss = s->loss;
  st.loss_hist = s->loss_hist; st.hist_len = s->hist_len; st.grad12 = s->grad12;
  st.stage = s->stage; st.start_vel = s->start_vel;
  return st;
}

Intr make_intr(const float* projraw16, float tfx, float tfy, int W, int H, int tw, int th) {
  Intr in;
  for (int k = 0; k < 16; ++k) in.projraw[k] = projraw16[k];
  in.tanfovx = tfx; in.tanfovy = tfy; in.W = W; in.H = H; in.tile_w = tw; in.tile_h = th;
  in.grid_x = (W + tw - 1) / tw; in.grid_y = (H + th - 1) / th;
  return in;
}

int check_intr(int W, int H, int tw, int th) {
  GSEVT_REQUIRE(W > 0 && H > 0, "image size must be positive");
  GSEVT_REQUIRE(tw > 0 && th > 0 && tw * th == 256 && tw % 8 == 0 && th % 8 == 0,
                "tile must hold 256 pixels in 8x8 blocks (16x16, 8x32 or 32x8)");
  // the one-channel splat record holds the tile rect in one byte per bound
  GSEVT_REQUIRE(((W + tw - 1) / tw) < 256 && ((H + th - 1) / th) < 256,
                "tile grid %dx%d too large for the fused path (at most 255 tiles per row and column)",
                (W + tw - 1) / tw, (H + th - 1) / th);
  return 0;
}

}  // namespace

extern "C" {

int gsevt_tr_camera_record_bytes(void) { return (int)sizeof(CamDev); }

int gsevt_tr_map_cov(int P, const float* scales, float scale_modifier, const float* rotations, float* cov6,
                     void* stream) {
  GSEVT_REQUIRE(P >= 0 && (P == 0 || (scales && rotations && cov6)), "bad arguments");
  if (P == 0) return 0;
  hipLaunchKernelGGL(k_tr_map_cov, dim3(div_up(P, 256)), dim3(256), 0, (hipStream_t)stream, P, scales,
                     scale_modifier, rotations, cov6);
  return check_launch("tr_map_cov");
}

int gsevt_tr_cams(int B, const gsevt_item_state* state, const float* projraw16, float tanfovx, float tanfovy,
                  int W, int H, int tile_w, int tile_h, void* cams, float* vel, float* dts, void* stream) {
  if (check_intr(W, H, tile_w, tile_h)) return -1;
  GSEVT_REQUIRE(B > 0 && state && projraw16 && cams && vel && dts, "bad arguments");
  const Intr in = make_intr(projraw16, tanfovx, tanfovy, W, H, tile_w, tile_h);
  hipLaunchKernelGGL(k_tr_cams, dim3(div_up(2 * B, 64)), dim3(64), 0, (hipStream_t)stream, B, to_state(state), in,
                     (CamDev*)cams, vel, dts);
  return check_launch("tr_cams");
}

int gsevt_tr_preprocess(int R, int P, int D, int M, const float* means, const float* cov6, const float* opac,
                        const float* shs, const float* intensity_weights3, const void* cams, float* splats,
                        uint32_t* rects, uint32_t* tiles, uint32_t* dkey, uint32_t* dval, int count_shift,
                        void* stream) {
  GSEVT_REQUIRE(R > 0 && P > 0 && D >= 0 && D <= 3, "bad sizes (R=%d P=%d D=%d)", R, P, D);
  GSEVT_REQUIRE(R < 65536, "too many renders in one launch");
  GSEVT_REQUIRE(count_shift == 0 || (count_shift >= 1 && count_shift <= 24 && (1LL << count_shift) >= P),
                "count_shift %d cannot hold %d gaussian indices", count_shift, P);
  GSEVT_REQUIRE(tiles || count_shift, "tiles may be null only when the counts travel in dval (count_shift > 0)");
  TrMap m;
  m.P = P; m.D = D; m.M = M; m.means = means; m.cov = cov6; m.opac = opac; m.shs = shs;
  m.iw[0] = intensity_weights3 ? intensity_weights3[0] : 0.2989f;
  m.iw[1] = intensity_weights3 ? intensity_weights3[1] : 0.5870f;
  m.iw[2] = intensity_weights3 ? intensity_weights3[2] : 0.1140f;
  TrBufs o;
  o.splats = (float4*)splats; o.rects = (uint2*)rects; o.tiles = tiles; o.dkey = dkey; o.dval = dval;
  o.count_shift = count_shift;
  const dim3 grid(div_up(P, 256), div_up(R, RPT));
  hipStream_t s = (hipStream_t)stream;
  switch (D) {
    case 0: hipLaunchKernelGGL(k_tr_preprocess<0>, grid, dim3(256), 0, s, m, (const CamDev*)cams, R, o); break;
    case 1: hipLaunchKernelGGL(k_tr_preprocess<1>, grid, dim3(256), 0, s, m, (const CamDev*)cams, R, o); break;
    case 2: hipLaunchKernelGGL(k_tr_preprocess<2>, grid, dim3(256), 0, s, m, (const CamDev*)cams, R, o); break;
    default: hipLaunchKernelGGL(k_tr_preprocess<3>, grid, dim3(256), 0, s, m, (const CamDev*)cams, R, o); break;
  }
  return check_launch("tr_preprocess");
}

// Scratch (uint32 elements) for gsevt_seg_sort_u32.
long long gsevt_seg_sort_scratch_elems(int R, long long stride) {
  const long long nb = stride <= 0 ? 1 : (long long)div_up(stride, SEG_TILE);
  const long long table = (long long)R * 256 * nb;
  return 2 * table + gsevt::scan_scratch_elems(table);
}

// Stable LSD sort of R segments (stride elements apart) of u32 keys/values on
// key bits [0, key_bits) in an even number of passes (result in keys/vals).
// counts: per-segment element counts on the device (null: every segment holds
// n_fixed elements).
int gsevt_seg_sort_u32_bits(int R, long long stride, const uint32_t* d_counts, long long n_fixed, uint32_t* keys,
                            uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, int begin_bit, int end_bit,
                            uint32_t* scratch, uint32_t* kept_counts, int first_hist_ready, void* stream);

int gsevt_seg_sort_u32(int R, long long stride, const uint32_t* d_counts, long long n_fixed, uint32_t* keys,
                       uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, int key_bits, uint32_t* scratch,
                       void* stream) {
  return gsevt_seg_sort_u32_bits(R, stride, d_counts, n_fixed, keys, vals, keys_alt, vals_alt, 0, key_bits, scratch,
                                 nullptr, 0, stream);
}

// As gsevt_seg_sort_u32 on key bits [begin_bit, end_bit); vals may be null
// (keys-only sort).  kept_counts non-null: the first pass removes every key
// equal to 0xFFFFFFFF (with its value), kept_counts[r] receives the number of
// keys segment r keeps, and the sorted keys of segment r are its first
// kept_counts[r] slots.  first_hist_ready: the first pass's histogram table
// is already at the start of scratch (written by the packed emission).
int gsevt_seg_sort_u32_bits(int R, long long stride, const uint32_t* d_counts, long long n_fixed, uint32_t* keys,
                            uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, int begin_bit, int end_bit,
                            uint32_t* scratch, uint32_t* kept_counts, int first_hist_ready, void* stream) {
  const int key_bits = end_bit - begin_bit;
  GSEVT_REQUIRE(R > 0 && R < 65536 && stride > 0 && begin_bit >= 0 && key_bits > 0 && end_bit <= 32,
                "bad arguments");
  GSEVT_REQUIRE((long long)R * stride < (1LL << 40), "too large");
  GSEVT_REQUIRE((vals == nullptr) == (vals_alt == nullptr), "vals and vals_alt must both be given or both be null");
  hipStream_t s = (hipStream_t)stream;
  const int nb = (int)div_up(stride, SEG_TILE);
  const long long table = (long long)R * 256 * nb;
  uint32_t* hist = scratch;
  uint32_t* incl = scratch + table;
  uint32_t* sscr = scratch + 2 * table;
  int passes = (key_bits + 7) / 8;
  if (passes & 1) ++passes;
  const int dbits = (key_bits + passes - 1) / passes;
  // the first pass takes the remainder (key_bits - (passes - 1) * dbits bits)
  // so that the last pass, whose digit has the fewest distinct values in
  // practice (high tile bits), is the wide one
  const int first = key_bits - (passes - 1) * dbits;
  uint32_t *ki = keys, *vi = vals, *ko = keys_alt, *vo = vals_alt;
  for (int p = 0; p < passes; ++p) {
    int nbt = (p == 0) ? first : dbits;
    if (nbt < 1) nbt = 1;
    const int shift = begin_bit + (p == 0 ? 0 : first + (p - 1) * dbits);
    // 2^nbt table rows per segment
    const long long ptable = (long long)R * (1LL << nbt) * nb;
    const bool drop = p == 0 && kept_counts;
    const uint32_t* cnt = (p > 0 && kept_counts) ? kept_counts : d_counts;
    const dim3 grid(nb, R);
    if (p == 0 && first_hist_ready) {
      GSEVT_REQUIRE(nbt <= 6, "a first pass with a given histogram takes at most 6 bits (got %d)", nbt);
    } else if (drop) {
      GSEVT_REQUIRE(nbt <= 6, "a dropping pass takes at most 6 bits (got %d)", nbt);
      hipLaunchKernelGGL((k_seg_hist2<6, true>), grid, dim3(SEG_THREADS), 0, s, ki, stride, cnt, n_fixed, shift, nbt,
                         hist, nb);
    } else if (nbt <= 6) {
      hipLaunchKernelGGL((k_seg_hist2<6, false>), grid, dim3(SEG_THREADS), 0, s, ki, stride, cnt, n_fixed, shift, nbt,
                         hist, nb);
    } else {
      hipLaunchKernelGGL((k_seg_hist2<8, false>), grid, dim3(SEG_THREADS), 0, s, ki, stride, cnt, n_fixed, shift, nbt,
                         hist, nb);
    }
    if (gsevt::scan_inclusive_u32(hist, incl, ptable, sscr, nullptr, s)) return -1;
#define GSEVT_SCATTER(DB_, DROP_, VALS_)                                                                       \
  hipLaunchKernelGGL((k_seg_scatter2<DB_, DROP_, VALS_>), grid, dim3(SEG_THREADS), 0, s, ki, vi, ko, vo, stride, \
                     cnt, n_fixed, shift, nbt, incl, hist, nb)
    const bool vals_on = vi != nullptr;
    if (drop) {
      if (vals_on) GSEVT_SCATTER(6, true, true); else GSEVT_SCATTER(6, true, false);
    } else if (nbt <= 6) {
      if (vals_on) GSEVT_SCATTER(6, false, true); else GSEVT_SCATTER(6, false, false);
    } else {
      if (vals_on) GSEVT_SCATTER(8, false, true); else GSEVT_SCATTER(8, false, false);
    }
#undef GSEVT_SCATTER
    if (drop)
      hipLaunchKernelGGL(k_seg_counts, dim3(div_up(R, 256)), dim3(256), 0, s, R, (long long)(1LL << nbt) * nb, incl,
                         kept_counts);
    if (check_launch("seg_sort_u32")) return -1;
    uint32_t* t = ki; ki = ko; ko = t;
    t = vi; vi = vo; vo = t;
  }
  return 0;
}

// Packed fragments in depth order: as gsevt_tr_emit, output tile << gid_bits |
// gid in one array (requires gid_bits + tile bits <= 32).
int gsevt_tr_emit_packed(int R, int P, const uint32_t* dval_sorted, const uint32_t* tiles, const float* splats,
                         int cull, int W, int H, int tile_w, int tile_h, int gid_bits, int counts_packed,
                         long long cap,
                         uint32_t* tiles_sorted, uint32_t* incl, uint32_t* scan_scratch, uint32_t* nfrag,
                         int* overflow, uint32_t* frag, uint32_t* sort_hist, int hist_bits, void* stream) {
  GSEVT_REQUIRE(R > 0 && P > 0 && cap > 0 && gid_bits > 0 && gid_bits < 32 && (1LL << gid_bits) >= P,
                "bad sizes");
  GSEVT_REQUIRE(!sort_hist || (hist_bits >= 1 && hist_bits <= 6 && cap < (1LL << 32)),
                "the first-pass histogram takes 1 to 6 digit bits (got %d)", hist_bits);
  GSEVT_REQUIRE(W > 0 && H > 0 && tile_w > 0 && tile_h > 0, "bad image or tile size");
  GSEVT_REQUIRE(splats, "splats is null (the records hold the tile rects)");
  hipStream_t s = (hipStream_t)stream;
  const int grid_x = (W + tile_w - 1) / tile_w;
  const long long n_tiles = (long long)grid_x * ((H + tile_h - 1) / tile_h);
  GSEVT_REQUIRE((n_tiles << gid_bits) < (1LL << 32), "%lld tiles and %d gaussian bits do not fit a 32-bit key "
                "below the dropped-fragment key", n_tiles, gid_bits);
  GSEVT_REQUIRE(!counts_packed || (n_tiles >> (32 - gid_bits)) == 0,
                "%lld tiles do not fit above %d gaussian bits of the sorted values", n_tiles, gid_bits);
  const int count_shift = counts_packed ? gid_bits : 0;
  const uint32_t dval_mask = counts_packed ? (uint32_t)((1ull << gid_bits) - 1) : 0xFFFFFFFFu;
  GSEVT_REQUIRE(counts_packed || (tiles && tiles_sorted), "tiles and tiles_sorted are needed without counts_packed");
  if (count_shift) {
    // the counts sit above the index bits of the sorted values: scan them there
    if (gsevt::scan_inclusive_u32(dval_sorted, incl, (long long)R * P, scan_scratch, nullptr, s, count_shift))
      return -1;
  } else {
    hipLaunchKernelGGL(k_tr_gather_tiles, dim3(div_up(P, 256), R), dim3(256), 0, s, R, P, dval_sorted, tiles,
                       tiles_sorted, 0);
    if (gsevt::scan_inclusive_u32(tiles_sorted, incl, (long long)R * P, scan_scratch, nullptr, s)) return -1;
  }
  hipLaunchKernelGGL(k_tr_nfrag, dim3(div_up(R, 256)), dim3(256), 0, s, R, P, incl, cap, nfrag, overflow);
  const int hist_nb = (int)div_up(cap, SEG_TILE);
  if (sort_hist) {
    const long long n = (long long)R * (1LL << hist_bits) * hist_nb;
    hipLaunchKernelGGL(k_fill_u32b, dim3(fill_blocks(n)), dim3(256), 0, s, sort_hist, n, 0u);
  }
  hipLaunchKernelGGL(k_tr_emit_packed, dim3(div_up(P, 256), R), dim3(256), 0, s, R, P, dval_sorted, dval_mask, incl,
                     (const float4*)splats, cull, grid_x, tile_w, tile_h, gid_bits, cap, frag, sort_hist, hist_bits,
                     hist_nb);
  return check_launch("tr_emit_packed");
}

int gsevt_tr_ranges_packed(int R, long long cap, const uint32_t* frag, const uint32_t* nfrag, int gid_bits,
                           int n_tiles, uint32_t* ranges, void* stream) {
  GSEVT_REQUIRE(R > 0 && cap > 0 && cap < (1LL << 32) && n_tiles > 0 && gid_bits > 0 && gid_bits < 32,
                "bad sizes");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_tr_ranges_packed, dim3(div_up(n_tiles, 256), R), dim3(256), 0, s, R, cap, frag, nfrag,
                     gid_bits, n_tiles, (uint2*)ranges);
  return check_launch("tr_ranges_packed");
}

long long gsevt_tr_bin_scratch_elems(int R, int P) { return gsevt::scan_scratch_elems((long long)R * P); }

// Fragments of every render in depth order: gathers tiles_touched in sorted
// order, scans it, records per-render counts (overflow flag when a count
// exceeds cap) and writes (tile, gaussian) pairs.
int gsevt_tr_emit(int R, int P, const uint32_t* dval_sorted, const uint32_t* tiles, const uint32_t* rects, int W,
                  int tile_w, long long cap, uint32_t* tiles_sorted, uint32_t* incl, uint32_t* scan_scratch,
                  uint32_t* nfrag, int* overflow, uint32_t* ftile, uint32_t* fgid, void* stream) {
  GSEVT_REQUIRE(R > 0 && P > 0 && cap > 0, "bad sizes");
  hipStream_t s = (hipStream_t)stream;
  const int grid_x = (W + tile_w - 1) / tile_w;
  hipLaunchKernelGGL(k_tr_gather_tiles, dim3(div_up(P, 256), R), dim3(256), 0, s, R, P, dval_sorted, tiles,
                     tiles_sorted, 0);
  if (gsevt::scan_inclusive_u32(tiles_sorted, incl, (long long)R * P, scan_scratch, nullptr, s)) return -1;
  hipLaunchKernelGGL(k_tr_nfrag, dim3(div_up(R, 256)), dim3(256), 0, s, R, P, incl, cap, nfrag, overflow);
  hipLaunchKernelGGL(k_tr_emit, dim3(div_up(P, 256), R), dim3(256), 0, s, R, P, dval_sorted, incl,
                     (const uint2*)rects, tiles, grid_x, cap, ftile, fgid);
  return check_launch("tr_emit");
}

int gsevt_tr_ranges(int R, long long cap, const uint32_t* ftile_sorted, const uint32_t* nfrag, int n_tiles,
                    uint32_t* ranges, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const long long n = 2LL * R * n_tiles;
  hipLaunchKernelGGL(k_fill_u32b, dim3(fill_blocks(n)), dim3(256), 0, s, ranges, n, 0u);
  hipLaunchKernelGGL(k_tr_ranges, dim3(div_up(cap, 256), R), dim3(256), 0, s, R, cap, ftile_sorted, nfrag, n_tiles,
                     (uint2*)ranges);
  return check_launch("tr_ranges");
}

int gsevt_tr_blend_forward(int R, const float* projraw16, float tanfovx, float tanfovy, int W, int H, int tile_w,
                           int tile_h, int P, long long cap, const uint32_t* ranges, const uint32_t* fgid,
                           uint32_t gid_mask, const float* splats, float bg_intensity, float* inten, float* final_T,
                           int* n_contrib, void* stream) {
  if (check_intr(W, H, tile_w, tile_h)) return -1;
  const Intr in = make_intr(projraw16, tanfovx, tanfovy, W, H, tile_w, tile_h);
  TrImg im = {};
  im.inten = inten; im.final_T = final_T; im.n_contrib = n_contrib;
  hipLaunchKernelGGL(k_tr_blend_fwd, dim3(in.grid_x * in.grid_y, R), dim3(256), 0, (hipStream_t)stream, in, P, cap,
                     (const uint2*)ranges, fgid, (const float4*)splats, im, bg_intensity, gid_mask);
  return check_launch("tr_blend_forward");
}

int gsevt_tr_blend_queue_count(void) { return TR_QUEUES; }

// Persistent launches of the two blends (k_tr_blend_queue): `workgroups`
// workgroups, xcd_queues != 0 for the 8 per-XCD queues (else one queue);
// queue holds gsevt_tr_blend_queue_count() uint32 counters, cleared here.
int gsevt_tr_blend_forward_queue(int R, const float* projraw16, float tanfovx, float tanfovy, int W, int H,
                                 int tile_w, int tile_h, int P, long long cap, const uint32_t* ranges,
                                 const uint32_t* fgid, uint32_t gid_mask, const float* splats, float bg_intensity,
                                 float* inten, float* final_T, int* n_contrib, uint32_t* queue, int workgroups,
                                 int xcd_queues, void* stream) {
  if (check_intr(W, H, tile_w, tile_h)) return -1;
  GSEVT_REQUIRE(R > 0 && queue && workgroups > 0 && workgroups <= 65536, "bad queue arguments");
  hipStream_t s = (hipStream_t)stream;
  const Intr in = make_intr(projraw16, tanfovx, tanfovy, W, H, tile_w, tile_h);
  TrImg im = {};
  im.inten = inten; im.final_T = final_T; im.n_contrib = n_contrib;
  hipLaunchKernelGGL(k_fill_u32b, dim3(1), dim3(256), 0, s, queue, (long long)TR_QUEUES, 0u);
  hipLaunchKernelGGL(k_tr_blend_queue<false>, dim3(workgroups), dim3(256), 0, s, in, R, P, cap,
                     (const uint2*)ranges, fgid, (const float4*)splats, im, nullptr, nullptr, 0, nullptr,
                     bg_intensity, gid_mask, queue, xcd_queues ? TR_QUEUES : 1);
  return check_launch("tr_blend_forward_queue");
}

int gsevt_tr_blend_backward_queue(int R, const float* projraw16, float tanfovx, float tanfovy, int W, int H,
                                  int tile_w, int tile_h, int P, long long cap, const uint32_t* ranges,
                                  const uint32_t* fgid, uint32_t gid_mask, const float* splats, float bg_intensity,
                                  const float* inten, const float* final_T, const int* n_contrib,
                                  const float* events, const gsevt_item_state* state, const float* scal,
                                  int unsigned_mode, float* grad6, int clear_grad6, uint32_t* queue, int workgroups,
                                  int xcd_queues, void* stream) {
  if (check_intr(W, H, tile_w, tile_h)) return -1;
  GSEVT_REQUIRE(R > 0 && queue && workgroups > 0 && workgroups <= 65536, "bad queue arguments");
  GSEVT_REQUIRE(state && (unsigned_mode >= 0 || state->stage), "unsigned_mode -1 needs state->stage");
  hipStream_t s = (hipStream_t)stream;
  const Intr in = make_intr(projraw16, tanfovx, tanfovy, W, H, tile_w, tile_h);
  const long long ng = (long long)R * P * 6;
  if (clear_grad6) hipLaunchKernelGGL(k_fill_f32, dim3(fill_blocks(ng)), dim3(256), 0, s, grad6, ng, 0.0f);
  hipLaunchKernelGGL(k_fill_u32b, dim3(1), dim3(256), 0, s, queue, (long long)TR_QUEUES, 0u);
  TrImg im = {};
  im.inten = (float*)inten; im.final_T = (float*)final_T; im.n_contrib = (int*)n_contrib; im.events = events;
  im.grad6 = grad6;
  hipLaunchKernelGGL(k_tr_blend_queue<true>, dim3(workgroups), dim3(256), 0, s, in, R, P, cap,
                     (const uint2*)ranges, fgid, (const float4*)splats, im, state->window, scal, unsigned_mode,
                     state->stage, bg_intensity, gid_mask, queue, xcd_queues ? TR_QUEUES : 1);
  return check_launch("tr_blend_backward_queue");
}

int gsevt_tr_loss_partials_count(void) { return LOSS_BLOCKS; }

int gsevt_tr_loss(int B, int W, int H, const float* inten, const float* events, const gsevt_item_state* state,
                  int unsigned_mode, double* partials, float* scal, void* stream) {
  GSEVT_REQUIRE(B > 0 && inten && events && state && partials && scal, "bad arguments");
  GSEVT_REQUIRE(unsigned_mode >= 0 || state->stage, "unsigned_mode -1 needs state->stage");
  hipStream_t s = (hipStream_t)stream;
  const ItemState st = to_state(state);
  hipLaunchKernelGGL(k_tr_loss_partial, dim3(LOSS_BLOCKS, B), dim3(256), 0, s, W * H, inten, events, st.window,
                     unsigned_mode, st.stage, partials);
  hipLaunchKernelGGL(k_tr_loss_final, dim3(B), dim3(64), 0, s, B, LOSS_BLOCKS, partials, scal, st);
  return check_launch("tr_loss");
}

int gsevt_tr_blend_backward(int R, const float* projraw16, float tanfovx, float tanfovy, int W, int H, int tile_w,
                            int tile_h, int P, long long cap, const uint32_t* ranges, const uint32_t* fgid,
                            uint32_t gid_mask, const float* splats, float bg_intensity, const float* inten,
                            const float* final_T,
                            const int* n_contrib, const float* events, const gsevt_item_state* state,
                            const float* scal, int unsigned_mode, float* grad6, int clear_grad6, void* stream) {
  if (check_intr(W, H, tile_w, tile_h)) return -1;
  GSEVT_REQUIRE(state && (unsigned_mode >= 0 || state->stage), "unsigned_mode -1 needs state->stage");
  hipStream_t s = (hipStream_t)stream;
  const Intr in = make_intr(projraw16, tanfovx, tanfovy, W, H, tile_w, tile_h);
  const long long ng = (long long)R * P * 6;
  if (clear_grad6) hipLaunchKernelGGL(k_fill_f32, dim3(fill_blocks(ng)), dim3(256), 0, s, grad6, ng, 0.0f);
  TrImg im = {};
  im.inten = (float*)inten; im.final_T = (float*)final_T; im.n_contrib = (int*)n_contrib; im.events = events;
  im.grad6 = grad6;
  hipLaunchKernelGGL(k_tr_blend_bwd, dim3(in.grid_x * in.grid_y, R), dim3(256), 0, s, in, P, cap,
                     (const uint2*)ranges, fgid, (const float4*)splats, im, state->window, scal, unsigned_mode,
                     state->stage, bg_intensity, gid_mask);
  return check_launch("tr_blend_backward");
}

int gsevt_tr_pose_partials_count(int P) { return (int)div_up(P, 256); }

int gsevt_tr_preprocess_backward(int R, int P, int D, int M, const float* means, const float* cov6,
                                 const float* opac, const float* shs, const float* intensity_weights3,
                                 const void* cams, float* grad6, float* partials, int clear_grad6,
                                 void* stream) {
  GSEVT_REQUIRE(R > 0 && P > 0 && D >= 0 && D <= 3, "bad sizes");
  TrMap m;
  m.P = P; m.D = D; m.M = M; m.means = means; m.cov = cov6; m.opac = opac; m.shs = shs;
  m.iw[0] = intensity_weights3 ? intensity_weights3[0] : 0.2989f;
  m.iw[1] = intensity_weights3 ? intensity_weights3[1] : 0.5870f;
  m.iw[2] = intensity_weights3 ? intensity_weights3[2] : 0.1140f;
  const dim3 grid(div_up(P, 256), div_up(R, RPT));
  hipStream_t s = (hipStream_t)stream;
  const CamDev* cm = (const CamDev*)cams;
  switch (D) {
    case 0: hipLaunchKernelGGL(k_tr_preprocess_bwd<0>, grid, dim3(256), 0, s, m, cm, R, grad6, partials,
                                 clear_grad6); break;
    case 1: hipLaunchKernelGGL(k_tr_preprocess_bwd<1>, grid, dim3(256), 0, s, m, cm, R, grad6, partials,
                                 clear_grad6); break;
    case 2: hipLaunchKernelGGL(k_tr_preprocess_bwd<2>, grid, dim3(256), 0, s, m, cm, R, grad6, partials,
                                 clear_grad6); break;
    default: hipLaunchKernelGGL(k_tr_preprocess_bwd<3>, grid, dim3(256), 0, s, m, cm, R, grad6, partials,
                                 clear_grad6); break;
  }
  return check_launch("tr_preprocess_backward");
}

int gsevt_tr_update(int B, int n_partials, const float* partials, const void* cams, const float* vel,
                    const float* dts, const gsevt_item_state* state, const float* lr4, int max_optim_iter,
                    int opt_vel, int start_vel_iter, int do_step, void* stream) {
  GSEVT_REQUIRE(B > 0 && partials && cams && vel && dts && state && lr4, "bad arguments");
  OptCfg cfg;
  for (int k = 0; k < 4; ++k) cfg.lr[k] = lr4[k];
  cfg.max_optim_iter = max_optim_iter; cfg.opt_vel = opt_vel; cfg.start_vel_iter = start_vel_iter;
  cfg.staged = 0; cfg.conv_thr = 0.0f;
  hipLaunchKernelGGL(k_tr_update, dim3(B), dim3(64), 0, (hipStream_t)stream, B, n_partials, partials,
                     (const CamDev*)cams, vel, dts, to_state(state), cfg, do_step);
  return check_launch("tr_update");
}

int gsevt_tr_update_staged(int B, int n_partials, const float* partials, const void* cams, const float* vel,
                           const float* dts, const gsevt_item_state* state, const float* lr4, int max_optim_iter,
                           float converged_threshold, void* stream) {
  GSEVT_REQUIRE(B > 0 && partials && cams && vel && dts && state && lr4, "bad arguments");
  GSEVT_REQUIRE(state->stage && state->start_vel, "staged update needs state->stage and state->start_vel");
  GSEVT_REQUIRE(max_optim_iter >= 0, "bad max_optim_iter");
  OptCfg cfg;
  for (int k = 0; k < 4; ++k) cfg.lr[k] = lr4[k];
  cfg.max_optim_iter = max_optim_iter; cfg.opt_vel = 0; cfg.start_vel_iter = 0;
  cfg.staged = 1; cfg.conv_thr = converged_threshold;
  hipLaunchKernelGGL(k_tr_update, dim3(B), dim3(64), 0, (hipStream_t)stream, B, n_partials, partials,
                     (const CamDev*)cams, vel, dts, to_state(state), cfg, 1);
  return check_launch("tr_update_staged");
}

}  // extern "C"
