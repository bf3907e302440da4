// Neighbour passes of the WCSPH step on gfx950:
//   * summation density (SPHRho, density_funcs.pyx:38-84) with an optional
//     Tait equation-of-state epilogue (eos_funcs.pyx:105-116);
//   * momentum with Monaghan artificial viscosity (pressure_funcs.pyx:
//     154-245) fused with the XSPH correction (xsph_funcs.pyx:31-88);
//   * the same pass fused with the Euler / RK2 stage update of the particle
//     (integrator.py:472-514, :534-551, :597-619).
//
// One thread per destination particle.  Destinations are visited in the
// cell-sorted layout, so the 64 lanes of a wave share the same few stencil
// rows and their source reads hit in L1/L2.  Each stencil (dy,dz) row is a
// single contiguous source range.  Candidates are screened in the storage
// precision and accepted by the reference's exact FP64 test.
#include <math.h>

#include "sph_passes.h"

namespace sph {

// Every accepted source j of destination (xd,yd,zd) in cell (cx,cy,cz):
// the stencil rows of the destination and of its periodic images (one ring
// wider, their cells come from FP64 image points), each pair through
// pair_geom; body(j, d2, rx, ry, rz).
template <typename T, typename CT, class F>
__device__ __forceinline__ void scan_pairs(const Grid& g, const PassArgs<T>& a, const View<T>& S,
                                           T xd, T yd, T zd, int cx, int cy, int cz, int ring,
                                           float r2f, double r2d, F&& body) {
  for_each_image(g, (double)xd, (double)yd, (double)zd, cx, cy, cz, ring * g.cs,
                 [&](double ox, double oy, double oz, int icx, int icy, int icz, bool image) {
                   for_each_row(g, S.cstart, icx, icy, icz, image ? ring + 1 : ring,
                                [&](int j0, int j1) {
                                  for (int j = j0; j < j1; ++j) {
                                    CT d2, rx, ry, rz;
                                    if (!pair_geom<T, CT>(a, S, j, xd, yd, zd, r2f, r2d, image,
                                                          ox, oy, oz, d2, rx, ry, rz))
                                      continue;
                                    body(j, d2, rx, ry, rz);
                                  }
                                });
                 });
}

template <typename T, typename CT, bool TAIT>
__global__ __launch_bounds__(kBlock) void density_kernel(Grid g, PassArgs<T> a) {
  const int bid = xcd_block(blockIdx.x, gridDim.x);
  const int i = bid * kBlock + threadIdx.x;
  if (i >= a.dst.n) return;
  if (a.dst.tag != nullptr && a.dst.tag[i] != 0) {
    a.rho[i] = T(0);
    if (TAIT) {
      a.p[i] = T(0);
      a.cs[i] = T(0);
    }
    return;
  }
  const T xd = a.dst.x[i], yd = a.dst.y[i], zd = a.dst.z[i];
  const T hd = a.dst.h[i];
  const double r = (double)hd * a.rs;
  const double r2d = r * r;
  const float r2f = (float)r2d;
  int cx, cy, cz;
  cell_xyz(a.dst.cell[i], g, cx, cy, cz);
  const int ring = ring_of(g, a.disp);
  const int dim = a.dim;
  CT rho = CT(0);
  for (int s = 0; s < a.nsrc; ++s) {
    const View<T> S = a.src[s];
    CT nr = CT(0);
    if (s == a.cache_src && a.cache_mode == 2) {
      nr = (CT)a.cache_sums[i];  // stored partial sum of a static source
    } else {
      scan_pairs<T, CT>(g, a, S, xd, yd, zd, cx, cy, cz, ring, r2f, r2d,
                        [&](int j, CT d2, CT, CT, CT) {
                          const CT hab = CT(0.5) * ((CT)hd + (CT)S.h[j]);
                          const CT inv = CT(1) / hab;
                          const CT q = sqrt(d2) * inv;
                          nr += (CT)S.m[j] * (w_shape<CT>(q) * h_dim_inv<CT>(inv, dim));
                        });
      if (s == a.cache_src && a.cache_mode == 1) a.cache_sums[i] = (double)nr;
    }
    if ((a.excl_self >> s) & 1u) {
      // the self pair (d2 = 0, h_ab = h_a) as the scan added it
      const CT hab = CT(0.5) * ((CT)hd + (CT)hd);
      const CT inv = CT(1) / hab;
      nr -= (CT)S.m[i] * (w_shape<CT>(sqrt(CT(0)) * inv) * h_dim_inv<CT>(inv, dim));
    }
    rho += nr;
  }
  rho *= (CT)a.fac1;
  a.rho[i] = (T)rho;
  if (TAIT) {
    double p, cs;
    tait((double)(T)rho, a.co, a.ro, a.gamma, a.B, p, cs);
    a.p[i] = (T)p;
    a.cs[i] = (T)cs;
  }
}

// SPHDensityRate (density_funcs.pyx:89-158): sum_b m_b (v_a - v_b) .
// grad_a W_ab with h_ab = (h_a + h_b)/2; non-LocalReal destinations get 0.
template <typename T, typename CT>
__global__ __launch_bounds__(kBlock) void density_rate_kernel(Grid g, PassArgs<T> a) {
  const int bid = xcd_block(blockIdx.x, gridDim.x);
  const int i = bid * kBlock + threadIdx.x;
  if (i >= a.dst.n) return;
  if (a.dst.tag != nullptr && a.dst.tag[i] != 0) {
    a.rate[i] = T(0);
    return;
  }
  const T xd = a.dst.x[i], yd = a.dst.y[i], zd = a.dst.z[i];
  const T ud = a.dst.u[i], vd = a.dst.v[i], wd = a.dst.w[i];
  const CT ha = (CT)a.dst.h[i];
  const double r = (double)a.dst.h[i] * a.rs;
  const double r2d = r * r;
  const float r2f = (float)r2d;
  int cx, cy, cz;
  cell_xyz(a.dst.cell[i], g, cx, cy, cz);
  const int ring = ring_of(g, a.disp);
  const int dim = a.dim;
  CT rate = CT(0);
  for (int s = 0; s < a.nsrc; ++s) {
    const View<T> S = a.src[s];
    CT nr = CT(0);
    scan_pairs<T, CT>(g, a, S, xd, yd, zd, cx, cy, cz, ring, r2f, r2d,
                      [&](int j, CT d2, CT rx, CT ry, CT rz) {
                        const CT hab = CT(0.5) * (ha + (CT)S.h[j]);
                        const CT inv = CT(1) / hab;
                        const CT rr = sqrt(d2);
                        const CT q = rr * inv;
                        const CT gf = grad_shape<CT>(q, rr, hab, inv) * h_dim_inv<CT>(inv, dim);
                        const CT vx = (CT)ud - (CT)S.u[j], vy = (CT)vd - (CT)S.v[j],
                                 vz = (CT)wd - (CT)S.w[j];
                        nr += (vx * rx + vy * ry + vz * rz) * gf * (CT)S.m[j];
                      });
    rate += nr;
  }
  a.rate[i] = (T)(rate * (CT)a.fac1);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void tait_kernel(int n, const T* __restrict__ rho,
                                                      const int64_t* __restrict__ tag,
                                                      double co, double ro, double gamma,
                                                      double B, T* __restrict__ p,
                                                      T* __restrict__ cs) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  if (tag != nullptr && tag[i] != 0) {
    p[i] = T(0);
    cs[i] = T(0);
    return;
  }
  double pp, cc;
  tait((double)rho[i], co, ro, gamma, B, pp, cc);
  p[i] = (T)pp;
  cs[i] = (T)cc;
}

// MODE 0: write acc / dv.  MODE 1: RK2 predictor.  MODE 2: RK2 final.
// MODE 3: Euler.
template <typename T, typename CT, int MODE>
__global__ __launch_bounds__(kBlock) void momentum_kernel(Grid g, PassArgs<T> a) {
  const int bid = xcd_block(blockIdx.x, gridDim.x);
  const int i = bid * kBlock + threadIdx.x;
  if (i >= a.dst.n) return;
  const int dim = a.dim;
  const bool local = a.dst.tag == nullptr || a.dst.tag[i] == 0;
  CT acc[3] = {CT(0), CT(0), CT(0)};
  CT dv[3] = {CT(0), CT(0), CT(0)};
  const T xd = a.dst.x[i], yd = a.dst.y[i], zd = a.dst.z[i];
  const T ud = a.dst.u[i], vd = a.dst.v[i], wd = a.dst.w[i];
  if (local) {
    const CT ha = (CT)a.dst.h[i];
    const CT rhoa = (CT)a.dst.rho[i];
    const CT Pa = (CT)a.dst.p[i];
    const CT ca = (CT)a.dst.cs[i];
    const CT Aa = Pa / (rhoa * rhoa);
    const double r = (double)a.dst.h[i] * a.rs;
    const double r2d = r * r;
    const float r2f = (float)r2d;
    int cx, cy, cz;
    cell_xyz(a.dst.cell[i], g, cx, cy, cz);
    const int ring = ring_of(g, a.disp);
    const CT alpha = (CT)a.alpha, beta = (CT)a.beta, eta2 = (CT)(a.eta * a.eta);
    for (int s = 0; s < a.nsrc; ++s) {
      const View<T> S = a.src[s];
      const bool do_m = (a.flags[s] & SPH_SRC_MOMENTUM) != 0 && a.do_mom;
      const bool do_x = (a.flags[s] & SPH_SRC_XSPH) != 0 && a.do_xsph;
      if (!do_m && !do_x) continue;
      CT am[3] = {CT(0), CT(0), CT(0)};
      CT ax[3] = {CT(0), CT(0), CT(0)};
      scan_pairs<T, CT>(g, a, S, xd, yd, zd, cx, cy, cz, ring, r2f, r2d,
                        [&](int j, CT d2, CT rx, CT ry, CT rz) {
          const CT hab = CT(0.5) * (ha + (CT)S.h[j]);
          const CT inv = CT(1) / hab;
          const CT rr = sqrt(d2);
          const CT q = rr * inv;
          const CT fac = h_dim_inv<CT>(inv, dim);
          const CT mb = (CT)S.m[j];
          const CT rhob = (CT)S.rho[j];
          if (do_m) {
            const CT vx = (CT)ud - (CT)S.u[j], vy = (CT)vd - (CT)S.v[j],
                     vz = (CT)wd - (CT)S.w[j];
            const CT dot = vx * rx + vy * ry + vz * rz;
            const CT Pb = (CT)S.p[j];
            CT tmp = Aa + Pb / (rhob * rhob);
            if (dot < CT(0)) {
              const CT cab = CT(0.5) * (ca + (CT)S.cs[j]);
              const CT rhoab = CT(0.5) * (rhoa + rhob);
              const CT mu = hab * dot / (d2 + eta2 * hab * hab);
              tmp += (-alpha * cab * mu + beta * mu * mu) / rhoab;
            }
            const CT gf = -mb * tmp * (grad_shape<CT>(q, rr, hab, inv) * fac);
            am[0] += gf * rx;
            am[1] += gf * ry;
            am[2] += gf * rz;
          }
          if (do_x) {
            const CT rhoab = CT(0.5) * (rhob + rhoa);
            const CT temp = mb * (w_shape<CT>(q) * fac) / rhoab;
            ax[0] += temp * ((CT)S.u[j] - (CT)ud);
            ax[1] += temp * ((CT)S.v[j] - (CT)vd);
            ax[2] += temp * ((CT)S.w[j] - (CT)wd);
          }
      });
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        acc[c] += am[c];
        dv[c] += ax[c];
      }
    }
    const CT f1 = (CT)a.fac1, fe = (CT)(a.fac1 * a.eps);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      acc[c] *= f1;
      dv[c] *= fe;
    }
  }
  if (MODE == 0) {
    // components >= dim are written (as 0) when the caller passes them,
    // like the reference's num_outputs = 3 in a 2D run
    for (int c = 0; c < 3; ++c) {
      if (a.do_mom && a.acc[c] != nullptr) a.acc[c][i] = (T)acc[c];
      if (a.do_xsph && a.dv[c] != nullptr) a.dv[c][i] = (T)dv[c];
    }
    return;
  }
  // Integrator updates in the reference's calc order: mom, gravity, step,
  // xsph; every property update is `current + k*dt` (integrator.py:505) or
  // `p0 + 0.5*dt*(k1+k2)` (:616).  Non-LocalReal particles have k = 0.
  const T dt = (T)a.dt;
  const T hdt = (T)(0.5 * a.dt);
  const T vel[3] = {ud, vd, wd};
  double mv2 = 0.0;
  for (int c = 0; c < dim; ++c) {
    const T x0 = a.s0[c][i];
    const T u0 = a.s0[3 + c][i];
    const T km = (T)acc[c];
    const T kx = (T)dv[c];
    const T kg = (local && a.has_gravity) ? (T)a.g[c] : T(0);
    const T ks = local ? vel[c] : T(0);  // PositionStepping: k = u at eval time
    T xn, un;
    if (MODE == 1 || MODE == 3) {
      un = u0 + km * dt;
      if (a.has_gravity) un = un + kg * dt;
      xn = x0 + ks * dt;
      if (a.do_xsph) xn = xn + kx * dt;
      if (MODE == 1) {
        a.k1[c][i] = km;
        a.k1[3 + c][i] = kx;
      }
    } else {
      const T k1m = a.k1[c][i], k1x = a.k1[3 + c][i];
      un = u0 + hdt * (k1m + km);
      if (a.has_gravity) un = un + hdt * (kg + kg);
      xn = x0 + hdt * (u0 + ks);
      if (a.do_xsph) xn = xn + hdt * (k1x + kx);
    }
    a.out[c][i] = xn;
    a.out[3 + c][i] = un;
    const double mv = (double)xn - (double)x0;
    mv2 += mv * mv;
  }
  if (a.disp_out != nullptr) {
    // wave max, then one atomic per wave (non-negative floats order as ints)
    float disp = move_bound(mv2);
    for (int off = 32; off > 0; off >>= 1) disp = fmaxf(disp, __shfl_xor(disp, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax((int*)a.disp_out, __float_as_int(disp));
  }
}

template <typename T>
static int32_t fill_views(PassArgs<T>& a, const sph_view* dst, int32_t nsrc,
                          const sph_view* srcs, const char* who) {
  if (nsrc < 0 || nsrc > SPH_MAX_SRC || (nsrc > 0 && srcs == nullptr)) {
    set_error("%s: %d sources (max %d)", who, nsrc, SPH_MAX_SRC);
    return SPH_EINVAL;
  }
  a.dst = to_view<T>(*dst);
  a.nsrc = nsrc;
  for (int s = 0; s < nsrc; ++s) {
    if (srcs[s].dtype != dst->dtype) {
      set_error("%s: source %d dtype differs from the destination", who, s);
      return SPH_EUNSUPPORTED;
    }
    a.src[s] = to_view<T>(srcs[s]);
    a.flags[s] = SPH_SRC_MOMENTUM | SPH_SRC_XSPH;
  }
  return SPH_OK;
}

static int32_t check_common(const sph_grid* g, const sph_params* prm, const sph_view* dst,
                            int32_t nsrc, const sph_view* srcs, bool vel, const char* who) {
  SPH_TRY(check_grid(g, who));
  if (prm == nullptr || prm->dim < 1 || prm->dim > 3) {
    set_error("%s: dimension must be 1, 2 or 3", who);
    return SPH_EUNSUPPORTED;
  }
  if (!(prm->radius_scale > 0)) {
    set_error("%s: radius_scale must be positive", who);
    return SPH_EINVAL;
  }
  SPH_TRY(check_view(dst, who, vel, vel, true));
  for (int s = 0; s < nsrc; ++s) {
    SPH_TRY(check_view(&srcs[s], who, vel, true, true));
    if (srcs[s].n > 0 && !srcs[s].m) {
      set_error("%s: source %d has no mass", who, s);
      return SPH_EINVAL;
    }
  }
  return SPH_OK;
}


template <typename T, typename CT>
static void launch_density(hipStream_t st, const Grid& G, const PassArgs<T>& a, bool tait) {
  const int nb = nblocks_for(a.dst.n);
  if (tait)
    hipLaunchKernelGGL((density_kernel<T, CT, true>), dim3(nb), dim3(kBlock), 0, st, G, a);
  else
    hipLaunchKernelGGL((density_kernel<T, CT, false>), dim3(nb), dim3(kBlock), 0, st, G, a);
}

template <typename T, typename CT>
static void launch_momentum(hipStream_t st, const Grid& G, const PassArgs<T>& a, int mode) {
  const int nb = nblocks_for(a.dst.n);
  switch (mode) {
    case 0:
      hipLaunchKernelGGL((momentum_kernel<T, CT, 0>), dim3(nb), dim3(kBlock), 0, st, G, a);
      break;
    case 1:
      hipLaunchKernelGGL((momentum_kernel<T, CT, 1>), dim3(nb), dim3(kBlock), 0, st, G, a);
      break;
    case 2:
      hipLaunchKernelGGL((momentum_kernel<T, CT, 2>), dim3(nb), dim3(kBlock), 0, st, G, a);
      break;
    default:
      hipLaunchKernelGGL((momentum_kernel<T, CT, 3>), dim3(nb), dim3(kBlock), 0, st, G, a);
      break;
  }
}

template <typename T>
static int32_t density_impl(hipStream_t st, const sph_grid* g, const sph_params* prm,
                            const sph_view* dst, int32_t nsrc, const sph_view* srcs,
                            double rmax, const float* disp, void* rho, void* p, void* cs,
                            const sph_nlist* nl, const sph_src_cache* cache) {
  PassArgs<T> a{};
  SPH_TRY(fill_views<T>(a, dst, nsrc, srcs, "sph_density"));
  set_params(a, prm);
  a.cache_src = -1;
  if (cache != nullptr) {
    a.cache_sums = cache->sums;
    a.cache_src = cache->src;
    a.cache_mode = cache->mode;
  }
  a.disp = disp;
  a.rho = (T*)rho;
  a.p = (T*)p;
  a.cs = (T*)cs;
  const bool do_tait = p != nullptr && cs != nullptr;
  Grid G = make_grid(*g, rmax);
  if (dst->n == 0) return SPH_OK;
  SPH_TRY(check_period(g, rmax, "sph_density"));
  if (packed_inputs(a, false) && !a.symmetric && !G.periodic && a.excl_self == 0) {
    if (nl != nullptr) {
      a.nl_idx = nl->idx;
      a.nl_count = nl->count;
      a.nl_max = nl->max_per;
      a.nl_mode = nl->mode;
      a.nl_skin = nl->skin;
    }
    return fast_density(st, *g, rmax, reinterpret_cast<const PassArgs<float>&>(a), do_tait,
                        prm->pair_prec);
  }
  if (nl != nullptr) {
    set_error("sph_density: neighbour lists are written by the packed FP32 path only (no "
              "FP64 storage, symmetric pairs, periodic axes or exclude_self)");
    return SPH_EINVAL;
  }
  if (sizeof(T) == 4 && prm->pair_prec == SPH_PAIR_F64)
    launch_density<T, double>(st, G, a, do_tait);
  else
    launch_density<T, T>(st, G, a, do_tait);
  return launch_status("sph_density");
}

template <typename T>
static int32_t momentum_impl(hipStream_t st, int mode, const sph_grid* g, const sph_params* prm,
                             const sph_view* dst, int32_t nsrc, const sph_view* srcs,
                             const uint32_t* flags, double rmax, const float* disp,
                             void* const* acc, void* const* dv, double dt,
                             const void* const* s0, void* const* k1, void* const* out,
                             float* disp_out, const sph_nlist* nl) {
  const char* who = mode == 0 ? "sph_momentum_xsph" : "sph_wcsph_stage";
  PassArgs<T> a{};
  SPH_TRY(fill_views<T>(a, dst, nsrc, srcs, who));
  set_params(a, prm);
  for (int s = 0; s < nsrc; ++s) a.flags[s] = flags ? flags[s] : SPH_SRC_MOMENTUM;
  a.disp = disp;
  if (mode == 0) {
    // plain pass: a NULL acc / dv (or a NULL first component) skips that sum
    a.do_mom = acc != nullptr && acc[0] != nullptr;
    a.do_xsph = dv != nullptr && dv[0] != nullptr;
    for (int c = 0; c < 3; ++c) {
      a.acc[c] = a.do_mom ? (T*)acc[c] : nullptr;
      a.dv[c] = a.do_xsph ? (T*)dv[c] : nullptr;
    }
    for (int c = 1; c < prm->dim; ++c) {
      if ((a.do_mom && !a.acc[c]) || (a.do_xsph && !a.dv[c])) {
        set_error("%s: output component %d is NULL", who, c);
        return SPH_EINVAL;
      }
    }
  } else {
    // stage: momentum always, XSPH when the scheme has it
    a.do_mom = 1;
    a.do_xsph = prm->has_xsph ? 1 : 0;
  }
  a.dt = dt;
  a.disp_out = disp_out;
  if (mode != 0) {
    if (!s0 || !out || (mode != 3 && !k1)) {
      set_error("%s: state0 / k1 / out arrays are required", who);
      return SPH_EINVAL;
    }
    for (int c = 0; c < 6; ++c) {
      const int comp = c % 3;
      const bool need = comp < prm->dim;
      a.s0[c] = (const T*)s0[c];
      a.out[c] = (T*)out[c];
      a.k1[c] = k1 ? (T*)k1[c] : nullptr;
      if (need && (!a.s0[c] || !a.out[c] || (mode != 3 && !a.k1[c]))) {
        set_error("%s: component %d of state0/k1/out is NULL", who, c);
        return SPH_EINVAL;
      }
      if ((mode == 1 || mode == 3) && need && dst->n > 0 &&
          ((const void*)a.out[c] == (const void*)a.dst.x ||
           (const void*)a.out[c] == (const void*)a.dst.y ||
           (const void*)a.out[c] == (const void*)a.dst.z ||
           (const void*)a.out[c] == (const void*)a.dst.u ||
           (const void*)a.out[c] == (const void*)a.dst.v ||
           (const void*)a.out[c] == (const void*)a.dst.w)) {
        set_error("%s: stage %d output may not alias the evaluated state", who, mode);
        return SPH_EINVAL;
      }
    }
  }
  Grid G = make_grid(*g, rmax);
  if (dst->n == 0) return SPH_OK;
  SPH_TRY(check_period(g, rmax, who));
  if (packed_inputs(a, true) && !a.symmetric && !G.periodic) {
    if (nl != nullptr) {
      a.nl_idx = nl->idx;
      a.nl_count = nl->count;
      a.nl_max = nl->max_per;
      a.nl_mode = 1;
      a.nl_skin = nl->skin;
    }
    return fast_momentum(st, *g, rmax, reinterpret_cast<const PassArgs<float>&>(a), mode,
                         prm->pair_prec);
  }
  if (nl != nullptr) {
    set_error("%s: neighbour lists are read by the packed FP32 path only (no FP64 storage, "
              "symmetric pairs or periodic axes)", who);
    return SPH_EINVAL;
  }
  if (sizeof(T) == 4 && prm->pair_prec == SPH_PAIR_F64)
    launch_momentum<T, double>(st, G, a, mode);
  else
    launch_momentum<T, T>(st, G, a, mode);
  return launch_status(who);
}

template <typename T>
static int32_t density_rate_impl(hipStream_t st, const sph_grid* g, const sph_params* prm,
                                 const sph_view* dst, int32_t nsrc, const sph_view* srcs,
                                 double rmax, const float* disp, void* rate) {
  PassArgs<T> a{};
  SPH_TRY(fill_views<T>(a, dst, nsrc, srcs, "sph_density_rate"));
  set_params(a, prm);
  a.disp = disp;
  a.rate = (T*)rate;
  Grid G = make_grid(*g, rmax);
  if (dst->n == 0) return SPH_OK;
  SPH_TRY(check_period(g, rmax, "sph_density_rate"));
  const int nb = nblocks_for(a.dst.n);
  if (sizeof(T) == 4 && prm->pair_prec == SPH_PAIR_F64)
    hipLaunchKernelGGL((density_rate_kernel<T, double>), dim3(nb), dim3(kBlock), 0, st, G, a);
  else
    hipLaunchKernelGGL((density_rate_kernel<T, T>), dim3(nb), dim3(kBlock), 0, st, G, a);
  return launch_status("sph_density_rate");
}

}  // namespace sph

using namespace sph;

extern "C" {

int32_t sph_density_rate(void* stream, const sph_grid* g, const sph_params* prm,
                         const sph_view* dst, int32_t nsrc, const sph_view* srcs, double rmax,
                         const float* disp_dev, void* rate_dev) {
  SPH_TRY(check_common(g, prm, dst, nsrc, srcs, true, "sph_density_rate"));
  if (dst->n > 0 && rate_dev == nullptr) {
    set_error("sph_density_rate: NULL output");
    return SPH_EINVAL;
  }
  if (dst->dtype == SPH_F32)
    return density_rate_impl<float>(as_stream(stream), g, prm, dst, nsrc, srcs, rmax, disp_dev,
                                     rate_dev);
  return density_rate_impl<double>(as_stream(stream), g, prm, dst, nsrc, srcs, rmax, disp_dev,
                                   rate_dev);
}

static int32_t density_entry(void* stream, const sph_grid* g, const sph_params* prm,
                             const sph_view* dst, int32_t nsrc, const sph_view* srcs,
                             double rmax, const float* disp_dev, void* rho_dev, void* p_dev,
                             void* cs_dev, const sph_nlist* nlist_out,
                             const sph_src_cache* cache) {
  SPH_TRY(check_common(g, prm, dst, nsrc, srcs, false, "sph_density"));
  if (cache != nullptr &&
      (nlist_out != nullptr || cache->src < 0 || cache->src >= nsrc ||
       (cache->mode != SPH_CACHE_FILL && cache->mode != SPH_CACHE_USE) ||
       (dst->n > 0 && cache->sums == nullptr))) {
    set_error("sph_density: bad source cache (a valid source, mode 1 or 2, sums, no list)");
    return SPH_EINVAL;
  }
  if (dst->n > 0 && rho_dev == nullptr) {
    set_error("sph_density: NULL rho output");
    return SPH_EINVAL;
  }
  if ((p_dev == nullptr) != (cs_dev == nullptr)) {
    set_error("sph_density: p and cs outputs must both be given or both be NULL");
    return SPH_EINVAL;
  }
  if (nlist_out != nullptr && (!nlist_out->idx || !nlist_out->count || nlist_out->max_per < 0)) {
    set_error("sph_density: neighbour list needs idx, count and max_per >= 0");
    return SPH_EINVAL;
  }
  for (int s = 0; s < 32; ++s) {
    if (!((prm->exclude_self >> s) & 1)) continue;
    if (s >= nsrc || srcs[s].x != dst->x || srcs[s].n != dst->n) {
      set_error("sph_density: exclude_self names source %d, which is not the destination", s);
      return SPH_EINVAL;
    }
  }
  if (dst->dtype == SPH_F32)
    return density_impl<float>(as_stream(stream), g, prm, dst, nsrc, srcs, rmax, disp_dev,
                               rho_dev, p_dev, cs_dev, nlist_out, cache);
  return density_impl<double>(as_stream(stream), g, prm, dst, nsrc, srcs, rmax, disp_dev,
                              rho_dev, p_dev, cs_dev, nlist_out, cache);
}

int32_t sph_density(void* stream, const sph_grid* g, const sph_params* prm, const sph_view* dst,
                    int32_t nsrc, const sph_view* srcs, double rmax, const float* disp_dev,
                    void* rho_dev, void* p_dev, void* cs_dev, const sph_nlist* nlist_out) {
  return density_entry(stream, g, prm, dst, nsrc, srcs, rmax, disp_dev, rho_dev, p_dev, cs_dev,
                       nlist_out, nullptr);
}

int32_t sph_density_cached(void* stream, const sph_grid* g, const sph_params* prm,
                           const sph_view* dst, int32_t nsrc, const sph_view* srcs, double rmax,
                           const float* disp_dev, void* rho_dev, void* p_dev, void* cs_dev,
                           const sph_src_cache* cache) {
  return density_entry(stream, g, prm, dst, nsrc, srcs, rmax, disp_dev, rho_dev, p_dev, cs_dev,
                       nullptr, cache);
}

int32_t sph_density_multi(void* stream, const sph_grid* g, const sph_params* prm, int32_t ndst,
                          const sph_view* dsts, int32_t nsrc, const sph_view* srcs, double rmax,
                          const float* disp_dev, void* const* rho_dev, void* const* p_dev,
                          void* const* cs_dev) {
  if (ndst < 0 || ndst > 2 || (ndst > 0 && (!dsts || !rho_dev))) {
    set_error("sph_density_multi: ndst must be 0, 1 or 2 with dsts and rho arrays");
    return SPH_EINVAL;
  }
  auto out = [](void* const* a, int k) { return a ? a[k] : nullptr; };
  for (int k = 0; k < ndst; ++k) {
    SPH_TRY(check_common(g, prm, &dsts[k], nsrc, srcs, false, "sph_density_multi"));
    if (dsts[k].n > 0 && rho_dev[k] == nullptr) {
      set_error("sph_density_multi: NULL rho output of destination %d", k);
      return SPH_EINVAL;
    }
    if ((out(p_dev, k) == nullptr) != (out(cs_dev, k) == nullptr)) {
      set_error("sph_density_multi: p and cs of destination %d must both be given or NULL", k);
      return SPH_EINVAL;
    }
  }
  if (prm->exclude_self != 0) {
    set_error("sph_density_multi: exclude_self is served by sph_density");
    return SPH_EINVAL;
  }
  // one launch for both destinations when both take the FP32-pair packed
  // scan with the same epilogue; otherwise one sph_density call each
  if (ndst == 2 && dsts[0].dtype == SPH_F32 && dsts[1].dtype == SPH_F32 &&
      prm->pair_prec != SPH_PAIR_F64 &&
      (out(p_dev, 0) != nullptr) == (out(p_dev, 1) != nullptr)) {
    PassArgs<float> a[2];
    bool fast = true;
    for (int k = 0; k < 2; ++k) {
      a[k] = PassArgs<float>{};
      SPH_TRY(fill_views<float>(a[k], &dsts[k], nsrc, srcs, "sph_density_multi"));
      set_params(a[k], prm);
      a[k].cache_src = -1;
      a[k].disp = disp_dev;
      a[k].rho = (float*)rho_dev[k];
      a[k].p = (float*)out(p_dev, k);
      a[k].cs = (float*)out(cs_dev, k);
      fast = fast && packed_inputs(a[k], false) && !a[k].symmetric;
    }
    const Grid G = make_grid(*g, rmax);
    if (fast && !G.periodic) {
      SPH_TRY(check_period(g, rmax, "sph_density_multi"));
      return fast_density_pair(as_stream(stream), *g, rmax, a[0], a[1], out(p_dev, 0) != nullptr);
    }
  }
  for (int k = 0; k < ndst; ++k)
    SPH_TRY(density_entry(stream, g, prm, &dsts[k], nsrc, srcs, rmax, disp_dev, rho_dev[k],
                          out(p_dev, k), out(cs_dev, k), nullptr, nullptr));
  return SPH_OK;
}

int32_t sph_tait(void* stream, const sph_params* prm, const sph_view* a, void* p_dev,
                 void* cs_dev) {
  if (prm == nullptr || a == nullptr || a->n < 0) {
    set_error("sph_tait: bad arguments");
    return SPH_EINVAL;
  }
  if (a->n == 0) return SPH_OK;
  if (!a->rho || !p_dev || !cs_dev) {
    set_error("sph_tait: rho, p and cs are required");
    return SPH_EINVAL;
  }
  const double B = prm->co * prm->co * prm->ro / prm->gamma;
  hipStream_t st = as_stream(stream);
  const int nb = nblocks_for(a->n);
  if (a->dtype == SPH_F32)
    hipLaunchKernelGGL(tait_kernel<float>, dim3(nb), dim3(kBlock), 0, st, (int)a->n,
                       (const float*)a->rho, a->tag, prm->co, prm->ro, prm->gamma, B,
                       (float*)p_dev, (float*)cs_dev);
  else if (a->dtype == SPH_F64)
    hipLaunchKernelGGL(tait_kernel<double>, dim3(nb), dim3(kBlock), 0, st, (int)a->n,
                       (const double*)a->rho, a->tag, prm->co, prm->ro, prm->gamma, B,
                       (double*)p_dev, (double*)cs_dev);
  else {
    set_error("sph_tait: dtype unsupported");
    return SPH_EUNSUPPORTED;
  }
  return launch_status("sph_tait");
}

int32_t sph_momentum_xsph(void* stream, const sph_grid* g, const sph_params* prm,
                          const sph_view* dst, int32_t nsrc, const sph_view* srcs,
                          const uint32_t* src_flags_host, double rmax, const float* disp_dev,
                          void* const* acc_dev, void* const* dv_dev, const sph_nlist* nlist_in) {
  SPH_TRY(check_common(g, prm, dst, nsrc, srcs, true, "sph_momentum_xsph"));
  if (dst->n > 0 && (!dst->p || !dst->cs)) {
    set_error("sph_momentum_xsph: destination p and cs are required");
    return SPH_EINVAL;
  }
  if (dst->dtype == SPH_F32)
    return momentum_impl<float>(as_stream(stream), 0, g, prm, dst, nsrc, srcs, src_flags_host,
                                rmax, disp_dev, acc_dev, dv_dev, 0.0, nullptr, nullptr,
                                nullptr, nullptr, nlist_in);
  return momentum_impl<double>(as_stream(stream), 0, g, prm, dst, nsrc, srcs, src_flags_host,
                               rmax, disp_dev, acc_dev, dv_dev, 0.0, nullptr, nullptr, nullptr,
                               nullptr, nlist_in);
}

int32_t sph_wcsph_stage(void* stream, int32_t stage, const sph_grid* g, const sph_params* prm,
                        const sph_view* dst, int32_t nsrc, const sph_view* srcs,
                        const uint32_t* src_flags_host, double rmax, const float* disp_dev,
                        double dt, const void* const* state0_dev, void* const* k1_dev,
                        void* const* out_dev, float* disp_out_dev, const sph_nlist* nlist_in) {
  SPH_TRY(check_common(g, prm, dst, nsrc, srcs, true, "sph_wcsph_stage"));
  if (stage < 0 || stage > 2) {
    set_error("sph_wcsph_stage: stage must be 0 (Euler), 1 or 2 (RK2)");
    return SPH_EINVAL;
  }
  if (dst->n > 0 && (!dst->p || !dst->cs)) {
    set_error("sph_wcsph_stage: destination p and cs are required");
    return SPH_EINVAL;
  }
  const int mode = stage == 0 ? 3 : stage;
  if (dst->dtype == SPH_F32)
    return momentum_impl<float>(as_stream(stream), mode, g, prm, dst, nsrc, srcs,
                                src_flags_host, rmax, disp_dev, nullptr, nullptr, dt, state0_dev,
                                k1_dev, out_dev, disp_out_dev, nlist_in);
  return momentum_impl<double>(as_stream(stream), mode, g, prm, dst, nsrc, srcs, src_flags_host,
                               rmax, disp_dev, nullptr, nullptr, dt, state0_dev, k1_dev, out_dev,
                               disp_out_dev, nlist_in);
}

}  // extern "C"
