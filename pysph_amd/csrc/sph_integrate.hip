// Integrator array updates of the generic (per-calc) path and the FP64
// conservation reductions.
//
// Reference: Integrator.step `current + k*dt` (integrator.py:505),
// EulerIntegrator.final_step `p0 + k1*dt` (:550), RK2Integrator.final_step
// `p0 + 0.5*dt*(k1 + k2)` (:616), set_initial_arrays / reset_current_arrays
// copies (:339-392).  Each op of one launch updates one property array.
#include <math.h>

#include "sph_internal.h"

namespace sph {

constexpr int kMaxOps = 12;

template <typename T>
struct UpdArgs {
  T* out[kMaxOps];
  T* a[kMaxOps];
  const T* b[kMaxOps];
  const T* c[kMaxOps];
  int nops;
  int n;
  T s;
  const int64_t* tag;
};

template <typename T, int MODE>
__global__ __launch_bounds__(kBlock) void update_kernel(UpdArgs<T> u) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= u.n) return;
  for (int k = 0; k < u.nops; ++k) {
    T r;
    if (MODE == 0) {
      // SPHFunction.eval: accumulate, 0 where tag != LocalReal
      r = (u.tag == nullptr || u.tag[i] == 0) ? u.a[k][i] + u.b[k][i] * u.s : T(0);
    } else if (MODE == 2) {
      r = u.a[k][i] + u.b[k][i] * u.s;
    } else if (MODE == 1) {
      r = u.a[k][i] + u.s * (u.b[k][i] + u.c[k][i]);
    } else if (MODE == 5) {
      r = (u.a[k][i] + T(2) * u.b[k][i]) + T(2) * u.c[k][i];
    } else if (MODE == 7) {
      r = T(2) * u.a[k][i] - u.b[k][i];
    } else if (MODE == 8) {
      r = u.a[k][i] + u.s * (u.b[k][i] - u.c[k][i]);
    } else if (MODE == 3) {
      r = (u.tag == nullptr || u.tag[i] == 0) ? u.a[k][i] : T(0);
    } else {
      r = (u.tag == nullptr || u.tag[i] == 0) ? u.s : T(0);
    }
    u.out[k][i] = r;
    if (MODE == 1 || MODE == 2) u.a[k][i] = r;
  }
}

template <typename T>
static int32_t update_impl(hipStream_t st, int64_t n, int mode, double s, int nops,
                           const sph_update_op* ops, const int64_t* tag) {
  for (int k0 = 0; k0 < nops; k0 += kMaxOps) {
    UpdArgs<T> u;
    u.nops = nops - k0 < kMaxOps ? nops - k0 : kMaxOps;
    u.n = (int)n;
    u.s = (T)s;
    u.tag = tag;
    for (int k = 0; k < u.nops; ++k) {
      const sph_update_op& op = ops[k0 + k];
      if (!op.out || (mode != 4 && !op.a) || (mode != 3 && mode != 4 && !op.b) ||
          ((mode == 1 || mode == 5 || mode == 8) && !op.c)) {
        set_error("sph_update: op %d has a NULL operand", k0 + k);
        return SPH_EINVAL;
      }
      u.out[k] = (T*)op.out;
      u.a[k] = (T*)op.a;
      u.b[k] = (const T*)op.b;
      u.c[k] = (const T*)op.c;
    }
    if (n == 0) continue;
    const int nb = nblocks_for(n);
    switch (mode) {
      case 0:
        hipLaunchKernelGGL((update_kernel<T, 0>), dim3(nb), dim3(kBlock), 0, st, u);
        break;
      case 1:
        hipLaunchKernelGGL((update_kernel<T, 1>), dim3(nb), dim3(kBlock), 0, st, u);
        break;
      case 2:
        hipLaunchKernelGGL((update_kernel<T, 2>), dim3(nb), dim3(kBlock), 0, st, u);
        break;
      case 3:
        hipLaunchKernelGGL((update_kernel<T, 3>), dim3(nb), dim3(kBlock), 0, st, u);
        break;
      case 5:
        hipLaunchKernelGGL((update_kernel<T, 5>), dim3(nb), dim3(kBlock), 0, st, u);
        break;
      case 7:
        hipLaunchKernelGGL((update_kernel<T, 7>), dim3(nb), dim3(kBlock), 0, st, u);
        break;
      case 8:
        hipLaunchKernelGGL((update_kernel<T, 8>), dim3(nb), dim3(kBlock), 0, st, u);
        break;
      default:
        hipLaunchKernelGGL((update_kernel<T, 4>), dim3(nb), dim3(kBlock), 0, st, u);
        break;
    }
  }
  return launch_status("sph_update");
}

// ---------------------------------------------------------------------------
// MoveCircleX / MoveCircleY (external_force.pyx:99-162): the velocity of
// unit-speed motion on a circle about the origin, from the angle
// arccos(x / |p|).  Non-LocalReal particles get 0 (sph_func.pyx:211-234).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void circle_kernel(View<T> a, int axis,
                                                        T* __restrict__ o0,
                                                        T* __restrict__ o1) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.n) return;
  double r0 = 0.0, r1 = 0.0;
  if (a.tag == nullptr || a.tag[i] == 0) {
    const double x = (double)a.x[i], y = (double)a.y[i], z = (double)a.z[i];
    const double len = sqrt(x * x + y * y + z * z);
    const double angle = acos(x / len);
    if (axis == 0) {
      r0 = -sin(angle);
      if (y < 0) r0 *= -1.0;
    } else {
      r1 = cos(angle);
    }
  }
  if (o0) o0[i] = (T)r0;
  if (o1) o1[i] = (T)r1;
}

// ---------------------------------------------------------------------------
// FP64 conservation sums: deterministic two-level reduction (per-block
// partials in a fixed order, then one block).
// ---------------------------------------------------------------------------
constexpr int kConsBlocks = 512;

template <typename T>
__global__ __launch_bounds__(kBlock) void cons_partial(View<T> a, double g0, double g1,
                                                       double g2, double ro, double gamma,
                                                       double B, double* __restrict__ part) {
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < a.n;
       i += (int64_t)gridDim.x * kBlock) {
    if (a.tag != nullptr && a.tag[i] != 0) continue;
    const double m = (double)a.m[i];
    const double u = (double)a.u[i], v = (double)a.v[i], w = (double)a.w[i];
    acc[0] += m;
    acc[1] += m * u;
    acc[2] += m * v;
    acc[3] += m * w;
    acc[4] += 0.5 * m * (u * u + v * v + w * w);
    acc[5] -= m * (g0 * (double)a.x[i] + g1 * (double)a.y[i] + g2 * (double)a.z[i]);
    const double rho = (double)a.rho[i];
    const double e = B * (pow(rho / ro, gamma - 1.0) / ((gamma - 1.0) * ro) + 1.0 / rho -
                          gamma / ((gamma - 1.0) * ro));
    acc[6] += m * e;
  }
  __shared__ double red[7][kBlock];
  for (int k = 0; k < 7; ++k) red[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int k = 0; k < 7; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 7) part[blockIdx.x * 8 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void cons_final(const double* __restrict__ part, int nb, double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  for (int k = 0; k < 7; ++k) {
    double s = 0;
    for (int b = 0; b < nb; ++b) s += part[b * 8 + k];
    out[k] += s;
  }
  out[7] = out[4] + out[5] + out[6];
}

}  // namespace sph

using namespace sph;

namespace sph {

constexpr int kMaxWrapExtra = 6;

template <typename T>
struct WrapArgs {
  T* pos[3];
  T* extra[kMaxWrapExtra][3];
  int nextra;
  int n;
  T lo[3], hi[3], per[3];
};

// Periodic wrap of the stored coordinates (the binning point rule of
// cell.pyx:392-398): one shift per axis, from the position, applied to the
// position and to every extra copy.
template <typename T>
__global__ __launch_bounds__(kBlock) void wrap_kernel(WrapArgs<T> w) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= w.n) return;
  for (int a = 0; a < 3; ++a) {
    if (w.pos[a] == nullptr) continue;
    const T x = w.pos[a][i];
    T sh = T(0);
    if (x >= w.hi[a]) sh = -w.per[a];
    else if (x < w.lo[a]) sh = w.per[a];
    if (sh == T(0)) continue;
    w.pos[a][i] = x + sh;
    for (int k = 0; k < w.nextra; ++k)
      if (w.extra[k][a] != nullptr) w.extra[k][a][i] = w.extra[k][a][i] + sh;
  }
}

template <typename T>
static int32_t wrap_impl(hipStream_t st, const sph_grid* g, int64_t n, void* const* pos,
                         int32_t nextra, void* const* extra) {
  WrapArgs<T> w{};
  w.n = (int)n;
  w.nextra = nextra;
  for (int a = 0; a < 3; ++a) {
    const bool per = g->period[a] > 0.0;
    w.pos[a] = per ? (T*)pos[a] : nullptr;
    if (per && w.pos[a] == nullptr) {
      set_error("sph_periodic_wrap: periodic axis %d has no coordinate array", a);
      return SPH_EINVAL;
    }
    w.lo[a] = (T)g->pmin[a];
    w.hi[a] = (T)(g->pmin[a] + g->period[a]);
    w.per[a] = (T)g->period[a];
    for (int k = 0; k < nextra; ++k) w.extra[k][a] = per ? (T*)extra[3 * k + a] : nullptr;
  }
  if (n == 0) return SPH_OK;
  hipLaunchKernelGGL(wrap_kernel<T>, dim3(nblocks_for(n)), dim3(kBlock), 0, st, w);
  return launch_status("sph_periodic_wrap");
}

}  // namespace sph

using namespace sph;

extern "C" {

int32_t sph_periodic_wrap(void* stream, const sph_grid* g, int32_t dtype, int64_t n,
                          void* const* pos_dev_host, int32_t nextra,
                          void* const* extra_dev_host) {
  if (g == nullptr || n < 0 || n > (int64_t)0x7fffffff || nextra < 0 ||
      nextra > kMaxWrapExtra || (n > 0 && pos_dev_host == nullptr) ||
      (nextra > 0 && extra_dev_host == nullptr)) {
    set_error("sph_periodic_wrap: bad arguments (at most %d extra arrays)", kMaxWrapExtra);
    return SPH_EINVAL;
  }
  SPH_TRY(check_grid(g, "sph_periodic_wrap"));
  if (dtype == SPH_F32)
    return wrap_impl<float>(as_stream(stream), g, n, pos_dev_host, nextra, extra_dev_host);
  if (dtype == SPH_F64)
    return wrap_impl<double>(as_stream(stream), g, n, pos_dev_host, nextra, extra_dev_host);
  set_error("sph_periodic_wrap: dtype unsupported");
  return SPH_EUNSUPPORTED;
}

int32_t sph_update(void* stream, int32_t dtype, int64_t n, int32_t mode, double s, int32_t nops,
                   const sph_update_op* ops_host, const int64_t* tag_dev) {
  if (n < 0 || nops < 0 || (nops > 0 && !ops_host) || mode < 0 || mode > 8 || mode == 6) {
    set_error("sph_update: bad arguments");
    return SPH_EINVAL;
  }
  if (dtype == SPH_F32) return update_impl<float>(as_stream(stream), n, mode, s, nops, ops_host, tag_dev);
  if (dtype == SPH_F64) return update_impl<double>(as_stream(stream), n, mode, s, nops, ops_host, tag_dev);
  set_error("sph_update: dtype unsupported");
  return SPH_EUNSUPPORTED;
}

int32_t sph_move_circle(void* stream, const sph_view* a, int32_t axis, void* out0_dev,
                        void* out1_dev) {
  SPH_TRY(check_view(a, "sph_move_circle", false, false, false));
  if (axis < 0 || axis > 1) {
    set_error("sph_move_circle: axis must be 0 (x) or 1 (y)");
    return SPH_EINVAL;
  }
  if (a->n == 0) return SPH_OK;
  hipStream_t st = as_stream(stream);
  const int nb = nblocks_for(a->n);
  if (a->dtype == SPH_F32)
    hipLaunchKernelGGL(circle_kernel<float>, dim3(nb), dim3(kBlock), 0, st, to_view<float>(*a),
                       (int)axis, (float*)out0_dev, (float*)out1_dev);
  else
    hipLaunchKernelGGL(circle_kernel<double>, dim3(nb), dim3(kBlock), 0, st,
                       to_view<double>(*a), (int)axis, (double*)out0_dev, (double*)out1_dev);
  return launch_status("sph_move_circle");
}

int32_t sph_conservation_f64(void* stream, const sph_params* prm, const sph_view* a,
                             double* out_dev) {
  if (prm == nullptr || out_dev == nullptr) {
    set_error("sph_conservation_f64: bad arguments");
    return SPH_EINVAL;
  }
  SPH_TRY(check_view(a, "sph_conservation_f64", true, true, false));
  hipStream_t st = as_stream(stream);
  if (a->n == 0) return SPH_OK;
  double* part = nullptr;
  hipError_t e = hipMallocAsync((void**)&part, (size_t)kConsBlocks * 8 * sizeof(double), st);
  if (e != hipSuccess) {
    set_error("sph_conservation_f64: %s", hipGetErrorString(e));
    return (int32_t)e;
  }
  const double B = prm->co * prm->co * prm->ro / prm->gamma;
  if (a->dtype == SPH_F32)
    hipLaunchKernelGGL(cons_partial<float>, dim3(kConsBlocks), dim3(kBlock), 0, st,
                       to_view<float>(*a), prm->g[0], prm->g[1], prm->g[2], prm->ro, prm->gamma,
                       B, part);
  else
    hipLaunchKernelGGL(cons_partial<double>, dim3(kConsBlocks), dim3(kBlock), 0, st,
                       to_view<double>(*a), prm->g[0], prm->g[1], prm->g[2], prm->ro, prm->gamma,
                       B, part);
  hipLaunchKernelGGL(cons_final, dim3(1), dim3(64), 0, st, part, kConsBlocks, out_dev);
  (void)hipFreeAsync(part, st);
  return launch_status("sph_conservation_f64");
}

}  // extern "C"
