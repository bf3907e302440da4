// Shared between the SoA neighbour passes (sph_passes.hip) and the packed
// FP32 fast path (sph_fast.hip): the pass arguments, the exact cut-off test
// and the Tait equation of state.
#pragma once
#include "sph_internal.h"

namespace sph {

// Exact cut-off test in storage precision T, returning the squared distance
// in CT for the pair arithmetic.
template <typename T, typename CT>
struct Cutoff;

template <typename CT>
struct Cutoff<float, CT> {
  __device__ __forceinline__ static bool test(float xs, float ys, float zs, float xd, float yd,
                                              float zd, float r2f, double r2d, CT& d2) {
    const float dx = xs - xd, dy = ys - yd, dz = zs - zd;
    const float d2f = dx * dx + dy * dy + dz * dz;
    if (!Within32::test(xs, ys, zs, xd, yd, zd, d2f, r2f, r2d)) return false;
    if (sizeof(CT) == 8) {
      const double ex = (double)xd - (double)xs, ey = (double)yd - (double)ys,
                   ez = (double)zd - (double)zs;
      d2 = (CT)(ex * ex + ey * ey + ez * ez);
    } else {
      d2 = (CT)d2f;
    }
    return true;
  }
};

template <typename CT>
struct Cutoff<double, CT> {
  __device__ __forceinline__ static bool test(double xs, double ys, double zs, double xd,
                                              double yd, double zd, float, double r2d,
                                              CT& d2) {
#pragma clang fp contract(off)
    const double dx = xs - xd, dy = ys - yd, dz = zs - zd;
    const double dd = (dx * dx + dy * dy) + dz * dz;
    d2 = (CT)dd;
    return dd < r2d;
  }
};

template <typename T>
struct PassArgs {
  View<T> dst;
  View<T> src[SPH_MAX_SRC];
  unsigned flags[SPH_MAX_SRC];
  int nsrc;
  int dim;
  double rs;   // radius_scale
  double fac1; // cubic spline fac(1)
  const float* disp;
  // density / Tait
  T* rho;
  T* p;
  T* cs;
  double co, ro, gamma, B;
  // momentum / XSPH
  double alpha, beta, eta, eps;
  T* acc[3];
  T* dv[3];
  int do_mom, do_xsph;  // which of the two pair sums the call computes
  // integrator stage
  double dt;
  double g[3];
  int has_gravity;
  const T* s0[6];  // x0,y0,z0,u0,v0,w0
  T* k1[6];        // k1 of momentum (3) and XSPH (3)
  T* out[6];       // new x,y,z,u,v,w
  float* disp_out;
  // neighbour list (superset within radius*(1+nl_skin)) built by a
  // density scan and read by the later passes of the same step
  int* nl_idx;
  int* nl_count;
  int nl_max;
  int nl_mode;  // 0 build, 1 use
  float nl_skin;
  int symmetric;  // variable-h symmetric pairs
  unsigned excl_self;  // density: sources whose self pair is dropped
  T* rate;        // SPHDensityRate output
  // density: stored partial sums of source cache_src (mode 1 fill, 2 use)
  double* cache_sums;
  int cache_src;
  int cache_mode;
  int lds_rows;  // density: stencil rows staged in LDS (sph_params.lds_rows)
};

// Scheme parameters of a pass from the ABI struct.
template <typename T>
inline void set_params(PassArgs<T>& a, const sph_params* prm) {
  a.dim = prm->dim;
  a.rs = prm->radius_scale;
  a.fac1 = fac_unit(prm->dim);
  a.co = prm->co;
  a.ro = prm->ro;
  a.gamma = prm->gamma;
  a.B = prm->co * prm->co * prm->ro / prm->gamma;
  a.alpha = prm->alpha;
  a.beta = prm->beta;
  a.eta = prm->eta;
  a.eps = prm->eps;
  for (int c = 0; c < 3; ++c) a.g[c] = prm->g[c];
  a.has_gravity = prm->has_gravity;
  a.symmetric = prm->symmetric != 0;
  a.excl_self = (unsigned)prm->exclude_self;
  a.lds_rows = prm->lds_rows != 0;
}

// Distance to a periodic image source at (sx,sy,sz), FP64 without
// contraction like the test on a ghost copy (point.pxd:35-37).
template <typename CT>
__device__ __forceinline__ bool image_geom(double sx, double sy, double sz, double xd, double yd,
                                           double zd, double r2d, CT& d2, CT& rx, CT& ry,
                                           CT& rz) {
#pragma clang fp contract(off)
  const double ex = sx - xd, ey = sy - yd, ez = sz - zd;
  const double dd = (ex * ex + ey * ey) + ez * ez;
  if (!(dd < r2d)) return false;
  d2 = (CT)dd;
  rx = (CT)(xd - sx);
  ry = (CT)(yd - sy);
  rz = (CT)(zd - sz);
  return true;
}

// The pair test and geometry of the SoA passes.  Cut-off: the
// destination's radius, or with symmetric pairs the larger of the two radii
// (d2 < r_d^2 || d2 < r_s^2 is d2 < max(r_d^2, r_s^2)); r2f / r2d: the
// destination's radius squared.  (ox, oy, oz) shifts the source to a
// periodic image: the shifted coordinate and the distance are then formed
// in FP64 like a ghost copy's ((x_j + L) - x_i), whatever the storage type.
// On acceptance d2 and r = x_d - x_s (in CT) are set.
template <typename T, typename CT>
__device__ __forceinline__ bool pair_geom(const PassArgs<T>& a, const View<T>& S, int j, T xd,
                                          T yd, T zd, float r2f, double r2d, bool image,
                                          double ox, double oy, double oz, CT& d2, CT& rx,
                                          CT& ry, CT& rz) {
  if (a.symmetric) {
    const double rs = (double)S.h[j] * a.rs;
    const double r2s = rs * rs;
    if (r2s > r2d) {
      r2d = r2s;
      r2f = (float)r2s;
    }
  }
  if (!image) {
    if (!Cutoff<T, CT>::test(S.x[j], S.y[j], S.z[j], xd, yd, zd, r2f, r2d, d2)) return false;
    rx = (CT)xd - (CT)S.x[j];
    ry = (CT)yd - (CT)S.y[j];
    rz = (CT)zd - (CT)S.z[j];
    return true;
  }
  return image_geom<CT>((double)S.x[j] + ox, (double)S.y[j] + oy, (double)S.z[j] + oz,
                        (double)xd, (double)yd, (double)zd, r2d, d2, rx, ry, rz);
}

// Tait equation of state in FP64 (eos_funcs.pyx:105-116).
__device__ __forceinline__ void tait(double rho, double co, double ro, double gamma, double B,
                                     double& p, double& cs) {
  const double ratio = rho / ro;
  const double gamma2 = 0.5 * (gamma - 1.0);
  p = (pow(ratio, gamma) - 1.0) * B;
  cs = pow(ratio, gamma2) * co;
}

// Largest array the packed fast path serves: its gathers address a float4
// array with 32-bit byte offsets.
constexpr int kPackedMaxN = 1 << 27;

// True when the destination and every source carry packed FP32 inputs (and
// are small enough for 32-bit byte offsets into them).
template <typename T>
static inline bool packed_inputs(const PassArgs<T>& a, bool need_aux) {
  if (sizeof(T) != 4) return false;
  if (!a.dst.pos4 || !a.dst.vel4 || (need_aux && !a.dst.aux4)) return false;
  if (a.dst.n > kPackedMaxN) return false;
  for (int s = 0; s < a.nsrc; ++s) {
    if (!a.src[s].pos4 || !a.src[s].vel4 || (need_aux && !a.src[s].aux4)) return false;
    if (a.src[s].n > kPackedMaxN) return false;
  }
  return true;
}

// Fast-path launchers (sph_fast.hip).  rmax is the largest search radius.
int32_t fast_density(hipStream_t st, const sph_grid& g, double rmax, PassArgs<float> a,
                     bool tait, int pair_prec);
// The FP32-pair density passes of two destinations over the same sources
// in one launch (no lists, no source cache).
int32_t fast_density_pair(hipStream_t st, const sph_grid& g, double rmax, PassArgs<float> a0,
                          PassArgs<float> a1, bool tait);
int32_t fast_momentum(hipStream_t st, const sph_grid& g, double rmax, PassArgs<float> a,
                      int mode, int pair_prec);

}  // namespace sph
