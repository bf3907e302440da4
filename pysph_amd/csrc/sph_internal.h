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
// A periodic axis must be longer than twice the search radius.
int32_t check_period(const sph_grid* g, double rmax, const char* who);
Grid make_grid(const sph_grid& g, double rmax);

// Grid of one scene of a multi-scene binning.
struct BatchBinGrid {
  double cs;
  int lo[3];
  int dims[3];
  int cell_base;
};
int32_t fill_i32(hipStream_t st, int32_t* p, int64_t n, int32_t v);
int32_t batch_bin(hipStream_t st, int64_t n, int64_t ncells, const float* x, const float* y,
                  const float* z, const int* scene, const BatchBinGrid* grids, int32_t* perm,
                  int32_t* cell, int32_t* cell_start, void* ws);
int32_t permute_fields(hipStream_t st, int64_t n, const int32_t* perm, int32_t nfields,
                       const void* const* src, void* const* dst, const int32_t* esz);

}  // namespace sph

#define SPH_TRY(expr)                  \
  do {                                 \
    int32_t _rc = (expr);              \
    if (_rc != SPH_OK) return _rc;     \
  } while (0)
