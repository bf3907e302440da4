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
  // neighbour list written by the density pass, read by the momentum pass
  int* nl_idx;
  int* nl_count;
  int nl_max;
};

// Tait equation of state in FP64 (eos_funcs.pyx:105-116).
__device__ __forceinline__ void tait(double rho, double co, double ro, double gamma, double B,
                                     double& p, double& cs) {
  const double ratio = rho / ro;
  const double gamma2 = 0.5 * (gamma - 1.0);
  p = (pow(ratio, gamma) - 1.0) * B;
  cs = pow(ratio, gamma2) * co;
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
    for_each_row(g, S.cstart, cx, cy, cz, ring, [&](int j0, int j1) {
      for (int j = j0; j < j1; ++j) {
        CT d2;
        if (!Cutoff<T, CT>::test(S.x[j], S.y[j], S.z[j], xd, yd, zd, r2f, r2d, d2)) continue;
        const CT hab = CT(0.5) * ((CT)hd + (CT)S.h[j]);
        const CT inv = CT(1) / hab;
        const CT q = sqrt(d2) * inv;
        nr += (CT)S.m[j] * (w_shape<CT>(q) * h_dim_inv<CT>(inv, dim));
      }
    });
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
      for_each_row(g, S.cstart, cx, cy, cz, ring, [&](int j0, int j1) {
        for (int j = j0; j < j1; ++j) {
          CT d2;
          if (!Cutoff<T, CT>::test(S.x[j], S.y[j], S.z[j], xd, yd, zd, r2f, r2d, d2)) continue;
          const CT hab = CT(0.5) * (ha + (CT)S.h[j]);
          const CT inv = CT(1) / hab;
          const CT rr = sqrt(d2);
          const CT q = rr * inv;
          const CT fac = h_dim_inv<CT>(inv, dim);
          const CT rx = (CT)xd - (CT)S.x[j], ry = (CT)yd - (CT)S.y[j],
                   rz = (CT)zd - (CT)S.z[j];
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
  float disp = 0.f;
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
    disp = fmaxf(disp, fabsf((float)(xn - x0)));
  }
  if (a.disp_out != nullptr) {
    // wave max, then one atomic per wave (non-negative floats order as ints)
    for (int off = 32; off > 0; off >>= 1) disp = fmaxf(disp, __shfl_xor(disp, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax((int*)a.disp_out, __float_as_int(disp));
  }
}

// ---------------------------------------------------------------------------
// Fast path: FP32 storage with packed float4 inputs.
//
// One lane per destination, as above, but the pair arithmetic is taken out
// of the candidate loop.  Candidates are screened from the packed (x,y,z,h)
// copy (one 16-byte load each); an accepted source index goes into the
// lane's queue in LDS.  When any lane's queue is full, and at the end of
// each source array, the wave drains all queues together: every lane runs
// the pair arithmetic on its own queued sources, so the expensive part runs
// with nearly all lanes active instead of the ~15% that accept a given
// candidate.  The wave-uniform control (ballots) keeps every lane in the
// loops; lanes past the end of the array only idle.
// ---------------------------------------------------------------------------
constexpr int kQueue = 32;
constexpr int kUnroll = 4;
__device__ constexpr float4 far4 = {3.0e38f, 3.0e38f, 3.0e38f, 0.f};

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
  return v;
}

__device__ __forceinline__ bool accept32(const float4& s, const float4& d, float r2f, double r2d) {
  const float dx = s.x - d.x, dy = s.y - d.y, dz = s.z - d.z;
  const float d2f = dx * dx + dy * dy + dz * dz;
  return Within32::test(s.x, s.y, s.z, d.x, d.y, d.z, d2f, r2f, r2d);
}

template <typename CT>
__device__ __forceinline__ CT pair_d2(const float4& d, const float4& s, CT& rx, CT& ry, CT& rz) {
  rx = (CT)d.x - (CT)s.x;
  ry = (CT)d.y - (CT)s.y;
  rz = (CT)d.z - (CT)s.z;
  return rx * rx + ry * ry + rz * rz;
}

// Row range of grid row (y, z) between x0 and x1 in one source array; empty
// when the row lies outside the grid.
__device__ __forceinline__ void row_range(const Grid& g, const int* __restrict__ cstart, int cx,
                                          int y, int z, int ring, int& j0, int& j1) {
  if (y < 0 || y >= g.dims[1] || z < 0 || z >= g.dims[2]) {
    j0 = j1 = 0;
    return;
  }
  const int x0 = max(cx - ring, 0), x1 = min(cx + ring, g.dims[0] - 1);
  const int base = (z * g.dims[1] + y) * g.dims[0];
  j0 = cstart[base + x0];
  j1 = cstart[base + x1 + 1];
}

template <typename CT, bool TAIT>
__global__ __launch_bounds__(kBlock) void density_fast(Grid g, PassArgs<float> a) {
  __shared__ int qbuf[kBlock / 64][kQueue][64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int* q = &qbuf[wid][0][lane];
  const int bid = xcd_block(blockIdx.x, gridDim.x);
  const int i = bid * kBlock + threadIdx.x;
  const bool valid = i < a.dst.n;
  const bool local = valid && (a.dst.tag == nullptr || a.dst.tag[i] == 0);
  const float4 pd = valid ? a.dst.pos4[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  const double r = (double)pd.w * a.rs;
  const double r2d = r * r;
  const float r2f = (float)r2d;
  int cx = 0, cy = 0, cz = 0;
  if (local) cell_xyz(a.dst.cell[i], g, cx, cy, cz);
  const int ring = ring_of(g, a.disp);
  const int dim = a.dim;
  const CT hd = (CT)pd.w;
  CT rho = CT(0);
  const int64_t ndst = a.dst.n;
  const bool write_list = a.nl_idx != nullptr && local;
  int lpos = 0;  // entries of this destination in the neighbour list
  for (int s = 0; s < a.nsrc; ++s) {
    const View<float> S = a.src[s];
    CT nr = CT(0);
    int cnt = 0, cnt_s = 0;
    auto drain = [&]() {
      const int m = wave_max(cnt);
      for (int k = 0; k < m; ++k) {
        if (k < cnt) {
          const int j = q[k * 64];
          const float4 ps = S.pos4[j];
          CT rx, ry, rz;
          const CT d2 = pair_d2<CT>(pd, ps, rx, ry, rz);
          const CT hab = CT(0.5) * (hd + (CT)ps.w);
          const CT inv = CT(1) / hab;
          const CT qq = sqrt(d2) * inv;
          nr += (CT)S.vel4[j].w * (w_shape<CT>(qq) * h_dim_inv<CT>(inv, dim));
          if (write_list && lpos + k < a.nl_max) a.nl_idx[(int64_t)(lpos + k) * ndst + i] = j;
        }
      }
      lpos += cnt;
      cnt_s += cnt;
      cnt = 0;
    };
    for (int dz = -ring; dz <= ring; ++dz) {
      for (int dy = -ring; dy <= ring; ++dy) {
        int j0 = 0, j1 = 0;
        if (local) row_range(g, S.cstart, cx, cy + dy, cz + dz, ring, j0, j1);
        // four candidates per iteration: their loads are independent and
        // in flight together
        while (__any(j0 < j1)) {
          if (__any(cnt > kQueue - kUnroll)) drain();
          float4 c[kUnroll];
#pragma unroll
          for (int u = 0; u < kUnroll; ++u)
            c[u] = (j0 + u < j1) ? S.pos4[j0 + u] : far4;
#pragma unroll
          for (int u = 0; u < kUnroll; ++u) {
            if (j0 + u < j1 && accept32(c[u], pd, r2f, r2d)) {
              q[cnt * 64] = j0 + u;
              ++cnt;
            }
          }
          j0 += kUnroll;
        }
      }
    }
    drain();
    rho += nr;
    if (a.nl_count != nullptr && valid) a.nl_count[(int64_t)s * ndst + i] = local ? cnt_s : 0;
  }
  if (!valid) return;
  if (!local) {
    a.rho[i] = 0.f;
    if (TAIT) {
      a.p[i] = 0.f;
      a.cs[i] = 0.f;
    }
    if (a.dst.aux4 != nullptr) a.dst.aux4[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    return;
  }
  rho *= (CT)a.fac1;
  const float rhof = (float)rho;
  a.rho[i] = rhof;
  if (TAIT) {
    double p, cs;
    tait((double)rhof, a.co, a.ro, a.gamma, a.B, p, cs);
    a.p[i] = (float)p;
    a.cs[i] = (float)cs;
    if (a.dst.aux4 != nullptr) {
      const float pf = (float)p;
      a.dst.aux4[i] = make_float4(rhof, pf, (float)cs,
                                  (float)((double)pf / ((double)rhof * (double)rhof)));
    }
  }
}

template <typename CT, int MODE>
__global__ __launch_bounds__(kBlock) void momentum_fast(Grid g, PassArgs<float> a) {
  __shared__ int qbuf[kBlock / 64][kQueue][64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int* q = &qbuf[wid][0][lane];
  const int bid = xcd_block(blockIdx.x, gridDim.x);
  const int i = bid * kBlock + threadIdx.x;
  const bool valid = i < a.dst.n;
  const bool local = valid && (a.dst.tag == nullptr || a.dst.tag[i] == 0);
  const int dim = a.dim;
  const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
  const float4 pd = valid ? a.dst.pos4[i] : zero4;
  const float4 vd = valid ? a.dst.vel4[i] : zero4;
  const float4 ad = valid ? a.dst.aux4[i] : zero4;
  CT acc[3] = {CT(0), CT(0), CT(0)};
  CT dv[3] = {CT(0), CT(0), CT(0)};
  const double r = (double)pd.w * a.rs;
  const double r2d = r * r;
  const float r2f = (float)r2d;
  int cx = 0, cy = 0, cz = 0;
  if (local) cell_xyz(a.dst.cell[i], g, cx, cy, cz);
  const int ring = ring_of(g, a.disp);
  const CT ha = (CT)pd.w, rhoa = (CT)ad.x, ca = (CT)ad.z;
  const CT Aa = sizeof(CT) == 8 ? (CT)ad.y / (rhoa * rhoa) : (CT)ad.w;
  const CT alpha = (CT)a.alpha, beta = (CT)a.beta, eta2 = (CT)(a.eta * a.eta);
  // neighbour list of the density pass: used unless it overflowed
  const int64_t ndst = a.dst.n;
  const bool have_list = a.nl_idx != nullptr;
  bool use_list = false;
  if (have_list && local) {
    int total = 0;
    for (int s = 0; s < a.nsrc; ++s) total += a.nl_count[(int64_t)s * ndst + i];
    use_list = total <= a.nl_max;
  }
  const bool scan = local && !use_list;
  int lpos = 0;
  for (int s = 0; s < a.nsrc; ++s) {
    const View<float> S = a.src[s];
    const bool do_m = (a.flags[s] & SPH_SRC_MOMENTUM) != 0 && a.do_mom;
    const bool do_x = (a.flags[s] & SPH_SRC_XSPH) != 0 && a.do_xsph;
    const int cnt_s = (have_list && local) ? a.nl_count[(int64_t)s * ndst + i] : 0;
    if (!do_m && !do_x) {
      lpos += cnt_s;
      continue;
    }
    CT am[3] = {CT(0), CT(0), CT(0)};
    CT ax[3] = {CT(0), CT(0), CT(0)};
    auto pair = [&](int j) {
      const float4 ps = S.pos4[j];
      const float4 vs = S.vel4[j];
      const float4 as = S.aux4[j];
      CT rx, ry, rz;
      const CT d2 = pair_d2<CT>(pd, ps, rx, ry, rz);
      const CT hab = CT(0.5) * (ha + (CT)ps.w);
      const CT inv = CT(1) / hab;
      const CT rr = sqrt(d2);
      const CT qq = rr * inv;
      const CT fac = h_dim_inv<CT>(inv, dim);
      const CT mb = (CT)vs.w;
      const CT rhob = (CT)as.x;
      if (do_m) {
        const CT vx = (CT)vd.x - (CT)vs.x, vy = (CT)vd.y - (CT)vs.y, vz = (CT)vd.z - (CT)vs.z;
        const CT dot = vx * rx + vy * ry + vz * rz;
        const CT Ab = sizeof(CT) == 8 ? (CT)as.y / (rhob * rhob) : (CT)as.w;
        CT tmp = Aa + Ab;
        if (dot < CT(0)) {
          const CT cab = CT(0.5) * (ca + (CT)as.z);
          const CT rhoab = CT(0.5) * (rhoa + rhob);
          const CT mu = hab * dot / (d2 + eta2 * hab * hab);
          tmp += (-alpha * cab * mu + beta * mu * mu) / rhoab;
        }
        const CT gf = -mb * tmp * (grad_shape<CT>(qq, rr, hab, inv) * fac);
        am[0] += gf * rx;
        am[1] += gf * ry;
        am[2] += gf * rz;
      }
      if (do_x) {
        const CT rhoab = CT(0.5) * (rhob + rhoa);
        const CT temp = mb * (w_shape<CT>(qq) * fac) / rhoab;
        ax[0] += temp * ((CT)vs.x - (CT)vd.x);
        ax[1] += temp * ((CT)vs.y - (CT)vd.y);
        ax[2] += temp * ((CT)vs.z - (CT)vd.z);
      }
    };
    if (have_list) {
      const int m = wave_max(use_list ? cnt_s : 0);
      for (int k = 0; k < m; ++k) {
        if (use_list && k < cnt_s) pair(a.nl_idx[(int64_t)(lpos + k) * ndst + i]);
      }
      lpos += cnt_s;
    }
    if (__any(scan)) {
      int cnt = 0;
      auto drain = [&]() {
        const int m = wave_max(cnt);
        for (int k = 0; k < m; ++k) {
          if (k < cnt) pair(q[k * 64]);
        }
        cnt = 0;
      };
      for (int dz = -ring; dz <= ring; ++dz) {
        for (int dy = -ring; dy <= ring; ++dy) {
          int j0 = 0, j1 = 0;
          if (scan) row_range(g, S.cstart, cx, cy + dy, cz + dz, ring, j0, j1);
          // four candidates per iteration: their loads are independent and
          // in flight together
          while (__any(j0 < j1)) {
            if (__any(cnt > kQueue - kUnroll)) drain();
            float4 c[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; ++u)
              c[u] = (j0 + u < j1) ? S.pos4[j0 + u] : far4;
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
              if (j0 + u < j1 && accept32(c[u], pd, r2f, r2d)) {
                q[cnt * 64] = j0 + u;
                ++cnt;
              }
            }
            j0 += kUnroll;
          }
        }
      }
      drain();
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      acc[c] += am[c];
      dv[c] += ax[c];
    }
  }
  if (!valid) return;
  {
    const CT f1 = (CT)a.fac1, fe = (CT)(a.fac1 * a.eps);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      acc[c] *= f1;
      dv[c] *= fe;
    }
  }
  if (MODE == 0) {
    for (int c = 0; c < 3; ++c) {
      if (a.do_mom && a.acc[c] != nullptr) a.acc[c][i] = (float)acc[c];
      if (a.do_xsph && a.dv[c] != nullptr) a.dv[c][i] = (float)dv[c];
    }
    return;
  }
  const float dt = (float)a.dt;
  const float hdt = (float)(0.5 * a.dt);
  const float vel[3] = {vd.x, vd.y, vd.z};
  float disp = 0.f;
  for (int c = 0; c < dim; ++c) {
    const float x0 = a.s0[c][i];
    const float u0 = a.s0[3 + c][i];
    const float km = local ? (float)acc[c] : 0.f;
    const float kx = local ? (float)dv[c] : 0.f;
    const float kg = (local && a.has_gravity) ? (float)a.g[c] : 0.f;
    const float ks = local ? vel[c] : 0.f;
    float xn, un;
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
      const float k1m = a.k1[c][i], k1x = a.k1[3 + c][i];
      un = u0 + hdt * (k1m + km);
      if (a.has_gravity) un = un + hdt * (kg + kg);
      xn = x0 + hdt * (u0 + ks);
      if (a.do_xsph) xn = xn + hdt * (k1x + kx);
    }
    a.out[c][i] = xn;
    a.out[3 + c][i] = un;
    disp = fmaxf(disp, fabsf(xn - x0));
  }
  if (a.disp_out != nullptr) {
    for (int off = 32; off > 0; off >>= 1) disp = fmaxf(disp, __shfl_xor(disp, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax((int*)a.disp_out, __float_as_int(disp));
  }
}

__global__ __launch_bounds__(kBlock) void pack_kernel(View<float> a, float4* __restrict__ pos4,
                                                      float4* __restrict__ vel4) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.n) return;
  if (pos4 != nullptr) pos4[i] = make_float4(a.x[i], a.y[i], a.z[i], a.h[i]);
  if (vel4 != nullptr) vel4[i] = make_float4(a.u[i], a.v[i], a.w[i], a.m[i]);
}

// True when the destination and every source carry packed FP32 inputs.
template <typename T>
static bool packed_inputs(const PassArgs<T>& a, bool need_aux) {
  if (sizeof(T) != 4) return false;
  if (!a.dst.pos4 || !a.dst.vel4 || (need_aux && !a.dst.aux4)) return false;
  for (int s = 0; s < a.nsrc; ++s)
    if (!a.src[s].pos4 || !a.src[s].vel4 || (need_aux && !a.src[s].aux4)) return false;
  return true;
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

template <typename T>
static void set_params(PassArgs<T>& a, const sph_params* prm) {
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

template <typename CT>
static void launch_momentum_fast(hipStream_t st, const Grid& G, const PassArgs<float>& a,
                                 int mode) {
  const int nb = nblocks_for(a.dst.n);
  switch (mode) {
    case 0:
      hipLaunchKernelGGL((momentum_fast<CT, 0>), dim3(nb), dim3(kBlock), 0, st, G, a);
      break;
    case 1:
      hipLaunchKernelGGL((momentum_fast<CT, 1>), dim3(nb), dim3(kBlock), 0, st, G, a);
      break;
    case 2:
      hipLaunchKernelGGL((momentum_fast<CT, 2>), dim3(nb), dim3(kBlock), 0, st, G, a);
      break;
    default:
      hipLaunchKernelGGL((momentum_fast<CT, 3>), dim3(nb), dim3(kBlock), 0, st, G, a);
      break;
  }
}

template <typename T>
static int32_t density_impl(hipStream_t st, const sph_grid* g, const sph_params* prm,
                            const sph_view* dst, int32_t nsrc, const sph_view* srcs,
                            double rmax, const float* disp, void* rho, void* p, void* cs,
                            const sph_nlist* nl) {
  PassArgs<T> a{};
  SPH_TRY(fill_views<T>(a, dst, nsrc, srcs, "sph_density"));
  set_params(a, prm);
  a.disp = disp;
  a.rho = (T*)rho;
  a.p = (T*)p;
  a.cs = (T*)cs;
  const bool do_tait = p != nullptr && cs != nullptr;
  Grid G = make_grid(*g, rmax);
  if (dst->n == 0) return SPH_OK;
  if (nl != nullptr && packed_inputs(a, false)) {
    a.nl_idx = nl->idx;
    a.nl_count = nl->count;
    a.nl_max = nl->max_per;
  }
  if (packed_inputs(a, false)) {
    const int nb = nblocks_for(a.dst.n);
    const PassArgs<float>& af = reinterpret_cast<const PassArgs<float>&>(a);
    if (prm->pair_prec == SPH_PAIR_F64) {
      if (do_tait)
        hipLaunchKernelGGL((density_fast<double, true>), dim3(nb), dim3(kBlock), 0, st, G, af);
      else
        hipLaunchKernelGGL((density_fast<double, false>), dim3(nb), dim3(kBlock), 0, st, G, af);
    } else {
      if (do_tait)
        hipLaunchKernelGGL((density_fast<float, true>), dim3(nb), dim3(kBlock), 0, st, G, af);
      else
        hipLaunchKernelGGL((density_fast<float, false>), dim3(nb), dim3(kBlock), 0, st, G, af);
    }
    return launch_status("sph_density");
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
  if (packed_inputs(a, true)) {
    if (nl != nullptr) {
      a.nl_idx = nl->idx;
      a.nl_count = nl->count;
      a.nl_max = nl->max_per;
    }
    const PassArgs<float>& af = reinterpret_cast<const PassArgs<float>&>(a);
    if (prm->pair_prec == SPH_PAIR_F64)
      launch_momentum_fast<double>(st, G, af, mode);
    else
      launch_momentum_fast<float>(st, G, af, mode);
    return launch_status(who);
  }
  if (sizeof(T) == 4 && prm->pair_prec == SPH_PAIR_F64)
    launch_momentum<T, double>(st, G, a, mode);
  else
    launch_momentum<T, T>(st, G, a, mode);
  return launch_status(who);
}

}  // namespace sph

using namespace sph;

extern "C" {

int32_t sph_density(void* stream, const sph_grid* g, const sph_params* prm, const sph_view* dst,
                    int32_t nsrc, const sph_view* srcs, double rmax, const float* disp_dev,
                    void* rho_dev, void* p_dev, void* cs_dev, const sph_nlist* nlist_out) {
  SPH_TRY(check_common(g, prm, dst, nsrc, srcs, false, "sph_density"));
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
  if (dst->dtype == SPH_F32)
    return density_impl<float>(as_stream(stream), g, prm, dst, nsrc, srcs, rmax, disp_dev,
                               rho_dev, p_dev, cs_dev, nlist_out);
  return density_impl<double>(as_stream(stream), g, prm, dst, nsrc, srcs, rmax, disp_dev,
                              rho_dev, p_dev, cs_dev, nlist_out);
}

int32_t sph_pack(void* stream, const sph_view* a, void* pos4_dev, void* vel4_dev) {
  SPH_TRY(check_view(a, "sph_pack", vel4_dev != nullptr, false, false));
  if (a->dtype != SPH_F32) {
    set_error("sph_pack: packed copies exist for FP32 storage only");
    return SPH_EUNSUPPORTED;
  }
  if (vel4_dev != nullptr && a->n > 0 && !a->m) {
    set_error("sph_pack: m is required for vel4");
    return SPH_EINVAL;
  }
  if (a->n == 0) return SPH_OK;
  hipLaunchKernelGGL(pack_kernel, dim3(nblocks_for(a->n)), dim3(kBlock), 0, as_stream(stream),
                     to_view<float>(*a), (float4*)pos4_dev, (float4*)vel4_dev);
  return launch_status("sph_pack");
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
