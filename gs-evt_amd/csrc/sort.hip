// Device-wide inclusive scan and stable LSD radix sort of (u64 key, u32 value)
// pairs for gfx950.
//
// Scan: reduce-then-scan in three launches (tile scan with per-tile totals,
// one-workgroup scan of the totals, carry add).  2048 elements per tile.
//
// Radix sort: up to 8 bits per pass, 4096 elements per tile, three launches
// per pass (tile histograms, scan of the digit-major histogram table, stable
// scatter).  The scatter ranks elements of equal digit inside a wave with
// ballots (one per digit bit: the peers of a lane are the lanes whose digit
// bits all agree), then across the four waves and across the 16 rounds of a
// tile through LDS counters, so equal keys keep their input order.  The
// element count is read from device memory and the grid is sized for the
// capacity, so the whole sort can be captured in a graph without a host
// round trip.
#include "gsevt_api.h"
#include "gsevt_device.h"

using namespace gsevt;

namespace {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

// Inclusive scan of one value per thread across the workgroup; returns the
// workgroup total through *total.
__device__ __forceinline__ uint32_t block_inclusive_scan(uint32_t v, uint32_t* s_wave, uint32_t* total) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nw = blockDim.x >> 6;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    const uint32_t t = __shfl_up(v, o, WAVE);
    if (lane >= o) v += t;
  }
  if (lane == WAVE - 1) s_wave[wave] = v;
  __syncthreads();
  uint32_t carry = 0, tot = 0;
  for (int w = 0; w < nw; ++w) {
    if (w < wave) carry += s_wave[w];
    tot += s_wave[w];
  }
  __syncthreads();
  *total = tot;
  return v + carry;
}

// Loads the SCAN_ITEMS consecutive elements of thread t of a tile (as two
// 16-byte loads in a full tile with 16-byte aligned input), shifted right by
// in_shift; elements past n read as 0.
__device__ __forceinline__ void scan_load(const uint32_t* __restrict__ in, long long n, long long t0, int in_shift,
                                          uint32_t (&v)[SCAN_ITEMS]) {
  static_assert(SCAN_ITEMS == 8, "the vector path moves 8 elements as two uint4");
  const long long base = t0 + (long long)threadIdx.x * SCAN_ITEMS;
  if (t0 + SCAN_TILE <= n && (reinterpret_cast<uintptr_t>(in) & 15) == 0) {   // uniform over the workgroup
    const uint4 a = *reinterpret_cast<const uint4*>(in + base);
    const uint4 b = *reinterpret_cast<const uint4*>(in + base + 4);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
  } else {
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) v[k] = (base + k < n) ? in[base + k] : 0u;
  }
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) v[k] >>= in_shift;
}

// Reduce then scan: tile_sums[b] = the sum of tile b's elements, then (after
// their exclusive scan) each tile is read again and written once, scanned
// and offset; the output is written once instead of written, read and
// written again.
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(const uint32_t* __restrict__ in, long long n,
                                                              uint32_t* __restrict__ tile_sums, int in_shift) {
  __shared__ uint32_t s_wave[SCAN_THREADS / WAVE];
  uint32_t v[SCAN_ITEMS];
  scan_load(in, n, (long long)blockIdx.x * SCAN_TILE, in_shift, v);
  uint32_t t = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) t += v[k];
#pragma unroll
  for (int o = WAVE / 2; o >= 1; o >>= 1) t += __shfl_xor(t, o, WAVE);
  if ((threadIdx.x & (WAVE - 1)) == 0) s_wave[threadIdx.x / WAVE] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
    for (int w = 0; w < SCAN_THREADS / WAVE; ++w) tot += s_wave[w];
    tile_sums[blockIdx.x] = tot;
  }
}

// Inclusive scan of tile b's elements plus offsets[b] (the exclusive scan of
// the tile sums); total_out (optional) receives the last element.
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_tiles(const uint32_t* __restrict__ in,
                                                             uint32_t* __restrict__ out, long long n,
                                                             const uint32_t* __restrict__ offsets, int in_shift,
                                                             uint32_t* __restrict__ total_out) {
  __shared__ uint32_t s_wave[SCAN_THREADS / WAVE];
  const long long t0 = (long long)blockIdx.x * SCAN_TILE;
  const long long base = t0 + (long long)threadIdx.x * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS];
  scan_load(in, n, t0, in_shift, v);
  uint32_t run = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    run += v[k];
    v[k] = run;
  }
  uint32_t total;
  const uint32_t incl = block_inclusive_scan(run, s_wave, &total);
  const uint32_t excl = incl - run + offsets[blockIdx.x];
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) v[k] += excl;
  if (t0 + SCAN_TILE <= n && (reinterpret_cast<uintptr_t>(out) & 15) == 0) {   // uniform over the workgroup
    *reinterpret_cast<uint4*>(out + base) = make_uint4(v[0], v[1], v[2], v[3]);
    *reinterpret_cast<uint4*>(out + base + 4) = make_uint4(v[4], v[5], v[6], v[7]);
  } else {
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k)
      if (base + k < n) out[base + k] = v[k];
  }
  if (total_out) {
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k)
      if (base + k == n - 1) *total_out = v[k];
  }
}

// Exclusive scan of the tile sums in one workgroup (any count), 8
// consecutive sums per thread and step.
__global__ __launch_bounds__(1024) void k_scan_sums(uint32_t* __restrict__ sums, int ntiles) {
  constexpr int K = 8;
  __shared__ uint32_t s_wave[1024 / WAVE];
  uint32_t carry = 0;
  for (int b = 0; b < ntiles; b += (int)blockDim.x * K) {
    const int i0 = b + (int)threadIdx.x * K;
    uint32_t v[K], run = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      v[k] = (i0 + k < ntiles) ? sums[i0 + k] : 0u;
      run += v[k];
    }
    uint32_t total;
    uint32_t excl = block_inclusive_scan(run, s_wave, &total) - run + carry;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (i0 + k < ntiles) sums[i0 + k] = excl;
      excl += v[k];
    }
    carry += total;
  }
}

// ---------------------------------------------------------------------------
// radix sort
// ---------------------------------------------------------------------------
constexpr int SORT_THREADS = 256;
constexpr int SORT_ROUNDS = 16;
constexpr int SORT_TILE = SORT_THREADS * SORT_ROUNDS;
constexpr int RADIX = 256;

__global__ __launch_bounds__(SORT_THREADS) void k_radix_hist(const uint64_t* __restrict__ keys,
                                                             const uint32_t* __restrict__ d_n, long long cap,
                                                             int shift, uint32_t mask,
                                                             uint32_t* __restrict__ hist, int ntiles) {
  __shared__ uint32_t s_h[RADIX];
  long long n = (long long)*d_n;
  if (n > cap) n = cap;
  s_h[threadIdx.x] = 0;
  __syncthreads();
  const long long base = (long long)blockIdx.x * SORT_TILE;
  for (int r = 0; r < SORT_ROUNDS; ++r) {
    const long long i = base + (long long)r * SORT_THREADS + threadIdx.x;
    if (i < n) {
      const uint32_t d = (uint32_t)(keys[i] >> shift) & mask;
      atomicAdd(&s_h[d], 1u);
    }
  }
  __syncthreads();
  hist[(size_t)threadIdx.x * ntiles + blockIdx.x] = s_h[threadIdx.x];
}

__global__ __launch_bounds__(SORT_THREADS) void k_radix_scatter(
    const uint64_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in, uint64_t* __restrict__ keys_out,
    uint32_t* __restrict__ vals_out, const uint32_t* __restrict__ d_n, long long cap, int shift, int nbits,
    const uint32_t* __restrict__ offsets_incl, const uint32_t* __restrict__ hist, int ntiles) {
  __shared__ uint32_t s_wcount[4][RADIX];
  __shared__ uint32_t s_run[RADIX];
  __shared__ uint32_t s_base[RADIX];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  long long n = (long long)*d_n;
  if (n > cap) n = cap;
  const uint32_t mask = (1u << nbits) - 1u;
  s_run[tid] = 0;
  {
    const size_t h = (size_t)tid * ntiles + blockIdx.x;
    s_base[tid] = offsets_incl[h] - hist[h];  // exclusive start of this tile's digit run
  }
  const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const long long base = (long long)blockIdx.x * SORT_TILE;
  for (int r = 0; r < SORT_ROUNDS; ++r) {
    const long long i = base + (long long)r * SORT_THREADS + tid;
    const bool valid = i < n;
    if (base + (long long)r * SORT_THREADS >= n) break;  // uniform across the workgroup
    uint64_t key = 0;
    uint32_t val = 0, d = 0;
    if (valid) {
      key = keys_in[i];
      val = vals_in[i];
      d = (uint32_t)(key >> shift) & mask;
    }
#pragma unroll
    for (int w = 0; w < 4; ++w) s_wcount[w][tid] = 0;
    __syncthreads();
    unsigned long long peers = __ballot(valid);
    for (int b = 0; b < nbits; ++b) {
      const unsigned long long bb = __ballot(valid && ((d >> b) & 1u));
      peers &= ((d >> b) & 1u) ? bb : ~bb;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & lt_mask);
    if (valid && rank == 0) s_wcount[wave][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    {
      // digit `tid`: exclusive prefix across waves, then advance the running count
      const uint32_t c0 = s_wcount[0][tid], c1 = s_wcount[1][tid], c2 = s_wcount[2][tid], c3 = s_wcount[3][tid];
      const uint32_t run = s_run[tid];
      s_wcount[0][tid] = run;
      s_wcount[1][tid] = run + c0;
      s_wcount[2][tid] = run + c0 + c1;
      s_wcount[3][tid] = run + c0 + c1 + c2;
      s_run[tid] = run + c0 + c1 + c2 + c3;
    }
    __syncthreads();
    if (valid) {
      const uint32_t pos = s_base[d] + s_wcount[wave][d] + rank;
      keys_out[pos] = key;
      vals_out[pos] = val;
    }
    __syncthreads();
  }
}

}  // namespace

namespace gsevt {

__global__ void k_zero1(uint32_t* p) { *p = 0u; }

int scan_inclusive_u32(const uint32_t* in, uint32_t* out, long long n, uint32_t* scratch, uint32_t* total,
                       hipStream_t s, int in_shift) {
  if (n <= 0) {
    if (total) hipLaunchKernelGGL(k_zero1, dim3(1), dim3(1), 0, s, total);
    return check_launch("scan/empty");
  }
  const int ntiles = (int)div_up(n, SCAN_TILE);
  hipLaunchKernelGGL(k_scan_reduce, dim3(ntiles), dim3(SCAN_THREADS), 0, s, in, n, scratch, in_shift);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, s, scratch, ntiles);
  hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles), dim3(SCAN_THREADS), 0, s, in, out, n, scratch, in_shift, total);
  return check_launch("scan");
}

long long scan_scratch_elems(long long n) { return n <= 0 ? 1 : (long long)div_up(n, SCAN_TILE); }

}  // namespace gsevt

extern "C" {

// Scratch (uint32 elements) needed by gsevt_scan_inclusive_u32 for n inputs.
long long gsevt_scan_scratch_elems(long long n) { return gsevt::scan_scratch_elems(n); }

int gsevt_scan_inclusive_u32(const uint32_t* in, uint32_t* out, long long n, uint32_t* scratch,
                             uint32_t* total_out, void* stream) {
  GSEVT_REQUIRE(n >= 0, "negative length");
  GSEVT_REQUIRE(n == 0 || (in && out && scratch), "null buffer");
  GSEVT_REQUIRE(n < (1LL << 32), "scan length exceeds 2^32");
  return gsevt::scan_inclusive_u32(in, out, n, scratch, total_out, (hipStream_t)stream);
}

// Scratch (uint32 elements) needed by gsevt_sort_pairs for a capacity.
long long gsevt_sort_scratch_elems(long long capacity) {
  const long long nt = capacity <= 0 ? 1 : (long long)div_up(capacity, SORT_TILE);
  const long long table = nt * RADIX;
  return 2 * table + gsevt_scan_scratch_elems(table) + 1;
}

// Number of LSD passes for a key width (always even, so the result lands
// back in the first buffer pair).
int gsevt_sort_passes(int key_bits) {
  int p = (key_bits + 7) / 8;
  if (p & 1) ++p;
  return p;
}

// Stable sort of (keys, values) on bits [0, key_bits).  keys/values hold the
// input and receive the output; keys_alt/values_alt are ping-pong buffers of
// the same capacity.  The element count is read from d_count (device).
int gsevt_sort_pairs(uint64_t* keys, uint32_t* values, uint64_t* keys_alt, uint32_t* values_alt,
                     const uint32_t* d_count, long long capacity, int key_bits, uint32_t* scratch, void* stream) {
  GSEVT_REQUIRE(key_bits > 0 && key_bits <= 64, "key_bits must be in [1, 64] (got %d)", key_bits);
  GSEVT_REQUIRE(capacity >= 0 && capacity < (1LL << 32), "capacity out of range");
  if (capacity == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  const int ntiles = (int)div_up(capacity, SORT_TILE);
  const long long table = (long long)ntiles * RADIX;
  uint32_t* hist = scratch;
  uint32_t* offs = scratch + table;
  uint32_t* sscr = scratch + 2 * table;
  const int passes = gsevt_sort_passes(key_bits);
  const int dbits = (key_bits + passes - 1) / passes;
  uint64_t* kin = keys;
  uint32_t* vin = values;
  uint64_t* kout = keys_alt;
  uint32_t* vout = values_alt;
  for (int p = 0; p < passes; ++p) {
    const int shift = p * dbits;
    int nb = key_bits - shift;
    if (nb > dbits) nb = dbits;
    if (nb < 1) nb = 1;
    const uint32_t mask = (1u << nb) - 1u;
    hipLaunchKernelGGL(k_radix_hist, dim3(ntiles), dim3(SORT_THREADS), 0, s, kin, d_count, capacity, shift, mask,
                       hist, ntiles);
    if (gsevt::scan_inclusive_u32(hist, offs, table, sscr, nullptr, s)) return -1;
    hipLaunchKernelGGL(k_radix_scatter, dim3(ntiles), dim3(SORT_THREADS), 0, s, kin, vin, kout, vout, d_count,
                       capacity, shift, nb, offs, hist, ntiles);
    if (check_launch("sort_pairs")) return -1;
    uint64_t* tk = kin; kin = kout; kout = tk;
    uint32_t* tv = vin; vin = vout; vout = tv;
  }
  return 0;
}

}  // extern "C"
