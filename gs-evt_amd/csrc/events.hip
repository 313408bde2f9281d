// Event-frame construction on the GPU (utils/event_camera/event.py:116-128 and
// utils/tracker.py:78-91 of the reference):
//   scatter      count[y, x] += (p ? +1 : -1), int32 atomics (exact)
//   undistort    Brown-Conrady (k1, k2, p1, p2, k3) inverse map with the same
//                camera matrix, coordinates quantised to 1/32 pixel and
//                bilinear weights from the 32x32 weight table, zero border
//                (the cv2.undistort algorithm)
//   blur         separable 9-tap Gaussian, sigma = 0.3((k-1)/2 - 1) + 0.8,
//                replicate border (cv2.GaussianBlur with sigma 0)
//   normalise    unit L2 norm, sum of squares in FP64 (cv2.normalize)
//   pyramid      nearest downsample, src = floor(dst * src_w / dst_w)
#include <mutex>

#include "gsevt_api.h"
#include "gsevt_device.h"

using namespace gsevt;

// Number of sum-of-squares partials of the event-frame normalisation
// (NORM_BLOCKS below).
constexpr int NOALLOC_PARTIALS = 256;

// Static scratch of gsevt_event_frame_noalloc: one copy per GPU, held until
// the process ends.  Host code never uses this name as an address (on the
// host it only names the symbol); the device address of each GPU's copy
// comes from hipGetSymbolAddress on that device and is kept per device.
static __device__ __align__(256) unsigned char scratch[GSEVT_EVENT_FRAME_NOALLOC_MAX_PIXELS * 8LL + NOALLOC_PARTIALS * 8];
static_assert(sizeof(scratch) == GSEVT_EVENT_FRAME_NOALLOC_MAX_PIXELS * 8LL + NOALLOC_PARTIALS * 8, "scratch size");

namespace {

constexpr int INTER_BITS = 5;
constexpr int INTER_TAB = 1 << INTER_BITS;

__global__ __launch_bounds__(256) void k_scatter(int n, const int* __restrict__ x, const int* __restrict__ y,
                                                 const int* __restrict__ p, int W, int H, int* __restrict__ count,
                                                 int* __restrict__ rejected) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int xi = x[i], yi = y[i];
  if ((unsigned)xi >= (unsigned)W || (unsigned)yi >= (unsigned)H) {
    if (rejected) atomicAdd(rejected, 1);
    return;
  }
  atomicAdd(&count[yi * W + xi], p[i] ? 1 : -1);
}

// Undistortion map: for each destination pixel the source position in
// 1/32-pixel fixed point (x in the low 16 bits... stored as two int32).
__global__ __launch_bounds__(256) void k_undistort_map(int W, int H, double fx, double fy, double cx, double cy,
                                                       double k1, double k2, double p1, double p2, double k3,
                                                       int2* __restrict__ map) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= W * H) return;
  const int j = idx % W, i = idx / W;
  // inverse of the (new) camera matrix, which equals the camera matrix
  const double ifx = 1.0 / fx, ify = 1.0 / fy;
  const double x = (double)j * ifx - cx * ifx;
  const double y = (double)i * ify - cy * ify;
  const double x2 = x * x, y2 = y * y;
  const double r2 = x2 + y2, _2xy = 2 * x * y;
  const double kr = 1 + ((k3 * r2 + k2) * r2 + k1) * r2;
  const double xd = x * kr + p1 * _2xy + p2 * (r2 + 2 * x2);
  const double yd = y * kr + p1 * (r2 + 2 * y2) + p2 * _2xy;
  const double u = fx * xd + cx;
  const double v = fy * yd + cy;
  // saturate_cast<int>(u * 32): round half to even, then saturate
  const double uu = rint(u * INTER_TAB), vv = rint(v * INTER_TAB);
  const int iu = uu > 2147483647.0 ? 2147483647 : (uu < -2147483648.0 ? (-2147483647 - 1) : (int)uu);
  const int iv = vv > 2147483647.0 ? 2147483647 : (vv < -2147483648.0 ? (-2147483647 - 1) : (int)vv);
  map[idx] = make_int2(iu, iv);
}

__global__ __launch_bounds__(256) void k_remap(int W, int H, const int* __restrict__ src,
                                               const int2* __restrict__ map, float* __restrict__ dst) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= W * H) return;
  const int2 m = map[idx];
  const int sx = m.x >> INTER_BITS, sy = m.y >> INTER_BITS;
  const int fxq = m.x & (INTER_TAB - 1), fyq = m.y & (INTER_TAB - 1);
  const float ax = (float)fxq * (1.0f / INTER_TAB), ay = (float)fyq * (1.0f / INTER_TAB);
  const float wx0 = 1.0f - ax, wx1 = ax, wy0 = 1.0f - ay, wy1 = ay;
  const float w00 = wy0 * wx0, w01 = wy0 * wx1, w10 = wy1 * wx0, w11 = wy1 * wx1;
  auto at = [&](int yy, int xx) -> float {
    return ((unsigned)xx < (unsigned)W && (unsigned)yy < (unsigned)H) ? (float)src[yy * W + xx] : 0.0f;
  };
  dst[idx] = at(sy, sx) * w00 + at(sy, sx + 1) * w01 + at(sy + 1, sx) * w10 + at(sy + 1, sx + 1) * w11;
}

struct Kern9 { float k[9]; };

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__global__ __launch_bounds__(256) void k_blur_rows(int W, int H, const float* __restrict__ src, Kern9 kr,
                                                   float* __restrict__ dst) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= W * H) return;
  const int x = idx % W, y = idx / W;
  const float* row = src + (size_t)y * W;
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < 9; ++k) s = s + kr.k[k] * row[clampi(x + k - 4, 0, W - 1)];
  dst[idx] = s;
}

__global__ __launch_bounds__(256) void k_blur_cols(int W, int H, const float* __restrict__ src, Kern9 kr,
                                                   float* __restrict__ dst) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= W * H) return;
  const int x = idx % W, y = idx / W;
  float s = kr.k[4] * src[(size_t)y * W + x];
#pragma unroll
  for (int k = 1; k <= 4; ++k)
    s = s + kr.k[4 + k] * (src[(size_t)clampi(y + k, 0, H - 1) * W + x] + src[(size_t)clampi(y - k, 0, H - 1) * W + x]);
  dst[idx] = s;
}

// sum of squares in FP64, one partial per workgroup
__global__ __launch_bounds__(256) void k_sumsq(int n, const float* __restrict__ a, double* __restrict__ part) {
  __shared__ double s_w[4];
  double v = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double t = (double)a[i];
    v += t * t;
  }
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, WAVE);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

// out_signed = a * (1/norm), out_abs = |out_signed| (norm from the partials)
__global__ __launch_bounds__(256) void k_normalize(int n, const float* __restrict__ a, const double* __restrict__ part,
                                                   int nparts, float* __restrict__ out_signed,
                                                   float* __restrict__ out_abs) {
  __shared__ float s_scale;
  if (threadIdx.x == 0) {
    double ss = 0.0;
    for (int k = 0; k < nparts; ++k) ss += part[k];
    const double nrm = sqrt(ss);
    s_scale = nrm > 0.0 ? (float)(1.0 / nrm) : 0.0f;
  }
  __syncthreads();
  const float sc = s_scale;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float v = a[i] * sc;
    out_signed[i] = v;
    if (out_abs) out_abs[i] = fabsf(v);
  }
}

__global__ __launch_bounds__(256) void k_resize_nearest(int sw, int sh, const float* __restrict__ src, int dw, int dh,
                                                        double ifx, double ify, float* __restrict__ dst) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= dw * dh) return;
  const int x = idx % dw, y = idx / dw;
  int sx = (int)floor((double)x * ifx);
  int sy = (int)floor((double)y * ify);
  sx = sx < sw - 1 ? sx : sw - 1;
  sy = sy < sh - 1 ? sy : sh - 1;
  dst[idx] = src[(size_t)sy * sw + sx];
}

constexpr int NORM_BLOCKS = NOALLOC_PARTIALS;

// Scratch layout shared by both event-frame entry points: the remapped frame
// and the row-blurred frame (W * H floats each), then NORM_BLOCKS doubles of
// sum-of-squares partials.  Every byte is written before it is read, so no
// part of it has to start at zero.
long long event_frame_scratch_bytes(long long pixels) { return pixels * 4 * 2 + NORM_BLOCKS * 8; }

int event_frame_launch(int W, int H, const int* count, const int* map_xy, int ksize, double sigma, void* workspace,
                       float* out_signed, float* out_abs, hipStream_t s, const char* what) {
  GSEVT_REQUIRE(ksize == 9, "only the 9x9 kernel of the tracking configs is built (got %d)", ksize);
  float* und = (float*)workspace;
  float* tmp = und + (size_t)W * H;
  double* part = (double*)(tmp + (size_t)W * H);
  // kernel taps: computed in double, rounded to float, renormalised in double
  if (sigma <= 0) sigma = ((ksize - 1) * 0.5 - 1) * 0.3 + 0.8;
  Kern9 kr;
  {
    const double scale2 = -0.5 / (sigma * sigma);
    double sum = 0.0;
    float cf[9];
    for (int i = 0; i < 9; ++i) {
      const double xx = i - (ksize - 1) * 0.5;
      cf[i] = (float)exp(scale2 * xx * xx);
      sum += cf[i];
    }
    sum = 1.0 / sum;
    for (int i = 0; i < 9; ++i) kr.k[i] = (float)(cf[i] * sum);
  }
  const unsigned nb = div_up((long long)W * H, 256);
  hipLaunchKernelGGL(k_remap, dim3(nb), dim3(256), 0, s, W, H, count, (const int2*)map_xy, und);
  hipLaunchKernelGGL(k_blur_rows, dim3(nb), dim3(256), 0, s, W, H, und, kr, tmp);
  hipLaunchKernelGGL(k_blur_cols, dim3(nb), dim3(256), 0, s, W, H, tmp, kr, und);
  hipLaunchKernelGGL(k_sumsq, dim3(NORM_BLOCKS), dim3(256), 0, s, W * H, und, part);
  hipLaunchKernelGGL(k_normalize, dim3(NORM_BLOCKS), dim3(256), 0, s, W * H, und, part, NORM_BLOCKS, out_signed,
                     out_abs);
  return check_launch(what);
}

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[MAX_DEVICES] = {};

// Device address of the current device's scratch, looked up once per device
// (with that device current, as the caller's is).
int static_scratch_ptr(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  GSEVT_REQUIRE(e == hipSuccess, "event_frame_noalloc: hipGetDevice: %s", hipGetErrorString(e));
  GSEVT_REQUIRE(dev >= 0 && dev < MAX_DEVICES, "event_frame_noalloc: device %d beyond %d", dev, MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    GSEVT_REQUIRE(e == hipSuccess && p, "event_frame_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

}  // namespace

extern "C" {

int gsevt_event_scatter(int n, const int* x, const int* y, const int* polarity, int W, int H, int* count,
                        int* rejected, void* stream) {
  GSEVT_REQUIRE(n >= 0 && W > 0 && H > 0, "bad sizes");
  if (n == 0) return 0;
  GSEVT_REQUIRE(x && y && polarity && count, "null buffer");
  hipLaunchKernelGGL(k_scatter, dim3(div_up(n, 256)), dim3(256), 0, (hipStream_t)stream, n, x, y, polarity, W, H,
                     count, rejected);
  return check_launch("event_scatter");
}

int gsevt_undistort_map(int W, int H, const double* K, const double* dist5, int* map_xy, void* stream) {
  GSEVT_REQUIRE(W > 0 && H > 0 && K && dist5 && map_xy, "bad arguments");
  hipLaunchKernelGGL(k_undistort_map, dim3(div_up((long long)W * H, 256)), dim3(256), 0, (hipStream_t)stream, W, H,
                     K[0], K[4], K[2], K[5], dist5[0], dist5[1], dist5[2], dist5[3], dist5[4], (int2*)map_xy);
  return check_launch("undistort_map");
}

// Scratch: 2 * W * H floats + NORM_BLOCKS doubles.
long long gsevt_event_frame_scratch_bytes(int W, int H) { return event_frame_scratch_bytes((long long)W * H); }

// count (int32 [H,W]) -> signed [H,W] and |signed| [H,W] float frames.
int gsevt_event_frame(int W, int H, const int* count, const int* map_xy, int ksize, double sigma, void* workspace,
                      float* out_signed, float* out_abs, void* stream) {
  GSEVT_REQUIRE(W > 0 && H > 0 && count && map_xy && workspace && out_signed, "bad arguments");
  return event_frame_launch(W, H, count, map_xy, ksize, sigma, workspace, out_signed, out_abs, (hipStream_t)stream,
                            "event_frame");
}

// As gsevt_event_frame with the scratch in this library's per-device array
// (no allocation, no workspace argument).
int gsevt_event_frame_noalloc(int W, int H, const int* count, const int* map_xy, int ksize, double sigma,
                              float* out_signed, float* out_abs, void* stream) {
  GSEVT_REQUIRE(W > 0 && H > 0 && count && map_xy && out_signed, "bad arguments");
  GSEVT_REQUIRE((long long)W * H <= GSEVT_EVENT_FRAME_NOALLOC_MAX_PIXELS,
                "event_frame_noalloc: %d x %d = %lld pixels exceeds the static scratch limit of %d pixels",
                W, H, (long long)W * H, GSEVT_EVENT_FRAME_NOALLOC_MAX_PIXELS);
  void* ws = nullptr;
  if (static_scratch_ptr(&ws) != 0) return -1;
  return event_frame_launch(W, H, count, map_xy, ksize, sigma, ws, out_signed, out_abs, (hipStream_t)stream,
                            "event_frame_noalloc");
}

int gsevt_resize_nearest(int sw, int sh, const float* src, int dw, int dh, float* dst, void* stream) {
  GSEVT_REQUIRE(sw > 0 && sh > 0 && dw > 0 && dh > 0 && src && dst, "bad arguments");
  const double ifx = 1.0 / ((double)dw / sw), ify = 1.0 / ((double)dh / sh);
  hipLaunchKernelGGL(k_resize_nearest, dim3(div_up((long long)dw * dh, 256)), dim3(256), 0, (hipStream_t)stream, sw,
                     sh, src, dw, dh, ifx, ify, dst);
  return check_launch("resize_nearest");
}

}  // extern "C"
