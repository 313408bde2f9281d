// Host-side helpers shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

// The public C ABI: every extern "C" definition in this library must match it.
#include "../../include/gsevt.h"

namespace gsevt {

void set_error(const char* fmt, ...);

// Returns 0 after a successful launch, otherwise records the HIP error.
inline int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s: %s", what, hipGetErrorString(e));
    return (int)e;
  }
  return 0;
}

inline unsigned div_up(long long a, long long b) { return (unsigned)((a + b - 1) / b); }
// Blocks of 256 threads for a grid-stride loop over n elements (1..65536).
inline unsigned fill_blocks(long long n) {
  const long long b = (n + 255) / 256;
  return (unsigned)(b < 1 ? 1 : (b > 65536 ? 65536 : b));
}

// Device-wide inclusive scan of n uint32 (sort.hip), of in[i] >> in_shift;
// scratch holds scan_scratch_elems(n) uint32, total (optional) receives the
// last value.
int scan_inclusive_u32(const uint32_t* in, uint32_t* out, long long n, uint32_t* scratch, uint32_t* total,
                       hipStream_t s, int in_shift = 0);
long long scan_scratch_elems(long long n);

}  // namespace gsevt

#define GSEVT_REQUIRE(cond, ...)        \
  do {                                  \
    if (!(cond)) {                      \
      gsevt::set_error(__VA_ARGS__);    \
      return -1;                        \
    }                                   \
  } while (0)
