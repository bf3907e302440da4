// Host-side helpers of the C ABI: error reporting and argument checks.
#pragma once
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>

#include "sph_common.h"

namespace sph {

void set_error(const char* fmt, ...);

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline int32_t launch_status(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s: %s", what, hipGetErrorString(e));
    return (int32_t)e;
  }
  return SPH_OK;
}

inline int nblocks_for(int64_t n, int per_block = kBlock) {
  return (int)((n + per_block - 1) / per_block);
}

// Validate a view for a call that needs the listed fields.
int32_t check_view(const sph_view* a, const char* who, bool need_vel, bool need_state,
                   bool need_bins);
int32_t check_grid(const sph_grid* g, const char* who);
Grid make_grid(const sph_grid& g, double rmax);

}  // namespace sph

#define SPH_TRY(expr)                  \
  do {                                 \
    int32_t _rc = (expr);              \
    if (_rc != SPH_OK) return _rc;     \
  } while (0)
