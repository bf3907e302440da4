// ABI plumbing: version, thread-local error message, argument validation.
#include <math.h>
#include <string.h>

#include "sph_internal.h"

namespace sph {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int32_t check_view(const sph_view* a, const char* who, bool need_vel, bool need_state,
                   bool need_bins) {
  if (a == nullptr) {
    set_error("%s: NULL view", who);
    return SPH_EINVAL;
  }
  if (a->n < 0 || a->n > (int64_t)0x7fffffff) {
    set_error("%s: particle count %lld out of range", who, (long long)a->n);
    return SPH_EINVAL;
  }
  if (a->dtype != SPH_F32 && a->dtype != SPH_F64) {
    set_error("%s: dtype %d unsupported", who, a->dtype);
    return SPH_EUNSUPPORTED;
  }
  if (a->n == 0) return SPH_OK;
  if (!a->x || !a->y || !a->z || !a->h) {
    set_error("%s: x, y, z and h are required", who);
    return SPH_EINVAL;
  }
  if (need_vel && (!a->u || !a->v || !a->w)) {
    set_error("%s: u, v, w are required", who);
    return SPH_EINVAL;
  }
  if (need_state && (!a->m || !a->rho)) {
    set_error("%s: m and rho are required", who);
    return SPH_EINVAL;
  }
  if (need_bins && (!a->cell || !a->cell_start)) {
    set_error("%s: the array is not binned (cell / cell_start missing)", who);
    return SPH_EINVAL;
  }
  return SPH_OK;
}

int32_t check_period(const sph_grid* g, double rmax, const char* who) {
  for (int d = 0; d < 3; ++d) {
    if (g->period[d] > 0.0 && !(g->period[d] > 2.0 * rmax)) {
      set_error("%s: periodic axis %d: period %g must exceed twice the search radius %g", who,
                d, g->period[d], rmax);
      return SPH_EINVAL;
    }
  }
  return SPH_OK;
}

int32_t check_grid(const sph_grid* g, const char* who) {
  if (g == nullptr) {
    set_error("%s: NULL grid", who);
    return SPH_EINVAL;
  }
  if (!(g->cell_size > 0.0) || !isfinite(g->cell_size)) {
    set_error("%s: cell size %g must be positive", who, g->cell_size);
    return SPH_EINVAL;
  }
  for (int d = 0; d < 3; ++d) {
    if (g->dims[d] < 1) {
      set_error("%s: grid dims must be >= 1", who);
      return SPH_EINVAL;
    }
  }
  for (int d = 0; d < 3; ++d) {
    if (g->period[d] < 0.0 || !isfinite(g->period[d]) || !isfinite(g->pmin[d])) {
      set_error("%s: periodic axis %d: period must be finite and >= 0", who, d);
      return SPH_EINVAL;
    }
  }
  const double nc = (double)g->dims[0] * g->dims[1] * g->dims[2];
  if (nc > 2.0e9) {
    set_error("%s: %g cells exceed the int32 cell index", who, nc);
    return SPH_EINVAL;
  }
  return SPH_OK;
}

Grid make_grid(const sph_grid& g, double rmax) {
  Grid G;
  G.cs = g.cell_size;
  G.cs_f = (float)g.cell_size;
  G.inv_cs_f = (float)(1.0 / g.cell_size);
  for (int d = 0; d < 3; ++d) {
    G.lo[d] = g.lo[d];
    G.dims[d] = g.dims[d];
  }
  G.rmax = rmax;
  int r = (int)ceil(rmax / g.cell_size);
  G.ring0 = r < 1 ? 1 : r;
  G.cell_base = 0;
  G.periodic = 0;
  for (int d = 0; d < 3; ++d) {
    G.per[d] = g.period[d] > 0.0 ? g.period[d] : 0.0;
    G.pmin[d] = g.pmin[d];
    if (G.per[d] > 0.0) G.periodic = 1;
  }
  return G;
}

}  // namespace sph

extern "C" {

int32_t sph_abi_version(void) { return SPH_ABI_VERSION; }

const char* sph_last_error(void) { return sph::g_err; }

int32_t sph_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

}  // extern "C"
