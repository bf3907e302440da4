// Packed FP32 fast path: the single-array kernels and their launchers (the
// device building blocks and the algorithm notes are in sph_fast.h).
#include "sph_fast.h"

namespace sph {


template <typename CT, bool TAIT, bool BUILD, int DIM>
__global__ __launch_bounds__(kBlock)
#if SPH_DENS_WAVES > 0
__attribute__((amdgpu_waves_per_eu(SPH_DENS_WAVES)))
#endif
void density_scan(Grid g, PassArgs<float> a) {
  __shared__ __align__(16) int qbuf[kBlock / 64][kDensQueue][64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int bid = xcd_block(blockIdx.x, gridDim.x);
  density_scan_body<CT, TAIT, BUILD, DIM>(g, a, bid * kBlock + threadIdx.x,
                                          &qbuf[wid][0][lane]);
}

// density_scan (FP32 pairs, no list) with the stencil rows of sources of
// uniform h staged in LDS (sph_params.lds_rows, opt-in).
template <bool TAIT, int DIM>
__global__ __launch_bounds__(kBlock)
#if SPH_DENS_WAVES > 0
__attribute__((amdgpu_waves_per_eu(SPH_DENS_WAVES)))
#endif
void density_scan_lds(Grid g, PassArgs<float> a) {
  __shared__ __align__(16) int qbuf[kBlock / 64][kDensQueue][64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int bid = xcd_block(blockIdx.x, gridDim.x);
  density_scan_body<float, TAIT, false, DIM, true>(g, a, bid * kBlock + threadIdx.x,
                                                   &qbuf[wid][0][lane]);
}

// The density passes of two destination arrays over the same sources in
// one launch: blocks [0, nb0) serve a0's destinations, the others a1's.
// One grid fills the GPU while the longer pass drains, which two launches
// would need a second stream for.
template <typename CT, bool TAIT, int DIM, bool LDS>
__device__ __forceinline__ void density_scan2_run(const Grid& g, const PassArgs<float>& a0,
                                                  const PassArgs<float>& a1, int nb0) {
  __shared__ __align__(16) int qbuf[kBlock / 64][kDensQueue][64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int bid = xcd_block(blockIdx.x, gridDim.x);
  // a block-uniform branch with one body per argument set: selecting a
  // reference to either kernel argument at run time would copy them to
  // scratch memory and read every field from there
  if (bid < nb0)
    density_scan_body<CT, TAIT, false, DIM, LDS>(g, a0, bid * kBlock + threadIdx.x,
                                                 &qbuf[wid][0][lane]);
  else
    density_scan_body<CT, TAIT, false, DIM, LDS>(g, a1, (bid - nb0) * kBlock + threadIdx.x,
                                                 &qbuf[wid][0][lane]);
}

template <typename CT, bool TAIT, int DIM>
__global__ __launch_bounds__(kBlock)
#if SPH_DENS_WAVES > 0
__attribute__((amdgpu_waves_per_eu(SPH_DENS_WAVES)))
#endif
void density_scan2(Grid g, PassArgs<float> a0, PassArgs<float> a1, int nb0) {
  density_scan2_run<CT, TAIT, DIM, false>(g, a0, a1, nb0);
}

// density_scan2 with the stencil rows staged in LDS (opt-in).
template <bool TAIT, int DIM>
__global__ __launch_bounds__(kBlock)
#if SPH_DENS_WAVES > 0
__attribute__((amdgpu_waves_per_eu(SPH_DENS_WAVES)))
#endif
void density_scan2_lds(Grid g, PassArgs<float> a0, PassArgs<float> a1, int nb0) {
  density_scan2_run<float, TAIT, DIM, true>(g, a0, a1, nb0);
}

// Density from a list (exact test on every entry); lanes whose list is
// invalid or overflowed scan the cells.
template <typename CT, bool TAIT>
__global__ __launch_bounds__(kBlock) void density_list(Grid g, PassArgs<float> a) {
  const int bid = xcd_block(blockIdx.x, gridDim.x);
  const int i = bid * kBlock + threadIdx.x;
  const bool valid = i < a.dst.n;
  const bool local = valid && (a.dst.tag == nullptr || a.dst.tag[i] == 0);
  const float4 pd = valid ? a.dst.pos4[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  const double r = (double)pd.w * a.rs;
  const double r2d = r * r;
  const float r2f = (float)r2d;
  const int dim = a.dim;
  const CT hd = (CT)pd.w;
  const int64_t ndst = a.dst.n;
  int total = 0;
  if (local)
    for (int s = 0; s < a.nsrc; ++s) total += a.nl_count[(int64_t)s * ndst + i];
  const bool use = local && total <= a.nl_max && list_valid(a.disp, a.nl_skin, r);
  const bool scan = local && !use;
  int cx = 0, cy = 0, cz = 0;
  if (scan) cell_xyz(a.dst.cell[i], g, cx, cy, cz);
  const int ring = ring_of(g, a.disp);
  const Cull cu = make_cull(pd, r + (a.disp ? (double)*a.disp : 0.0), g);
  CT rho = CT(0);
  int lpos = 0;
  for (int s = 0; s < a.nsrc; ++s) {
    const View<float> S = a.src[s];
    const int cnt_s = local ? a.nl_count[(int64_t)s * ndst + i] : 0;
    CT nr = CT(0);
    const int* lst = a.nl_idx + (int64_t)lpos * ndst + i;
    const int m = wave_max(use ? cnt_s : 0);
    for (int k = 0; k < m; k += 2) {
      const bool e0 = use && k < cnt_s, e1 = use && k + 1 < cnt_s;
      const int j0 = e0 ? lst[(int64_t)k * ndst] : 0;
      const int j1 = e1 ? lst[(int64_t)(k + 1) * ndst] : 0;
      const float4 p0 = e0 ? S.pos4[j0] : kFar4;
      const float4 p1 = e1 ? S.pos4[j1] : kFar4;
      const float m0 = e0 ? S.vel4[j0].w : 0.f, m1 = e1 ? S.vel4[j1].w : 0.f;
      if (e0 && accept_exact(p0, pd, r2f, r2d)) nr += w_term<CT>(pd, p0, hd, (CT)m0, dim);
      if (e1 && accept_exact(p1, pd, r2f, r2d)) nr += w_term<CT>(pd, p1, hd, (CT)m1, dim);
    }
    lpos += cnt_s;
    if (__any(scan)) {
      for (int dz = -ring; dz <= ring; ++dz) {
        for (int dy = -ring; dy <= ring; ++dy) {
          int j0 = 0, j1 = 0;
          if (scan) row_range_cull(g, S.cstart, cx, cy + dy, cz + dz, ring, cu, j0, j1);
          for (int j = j0; j < j1; ++j) {
            const float4 ps = S.pos4[j];
            if (accept_exact(ps, pd, r2f, r2d)) nr += w_term<CT>(pd, ps, hd, (CT)S.vel4[j].w, dim);
          }
        }
      }
    }
    rho += nr;
  }
  if (!valid) return;
  density_epilogue<TAIT>(a, i, local, (double)(rho * (CT)a.fac1));
}

// Six waves per SIMD: the compiler keeps the momentum scan within 80 VGPRs
// (it takes 88 unconstrained, five waves), without spilling; measured 7%
// faster on the 3D dam break.
template <typename CT, int MODE, int DIM>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kMomWaves)))
void momentum_scan(Grid g, PassArgs<float> a) {
  __shared__ int qbuf[kBlock / 64][kQueue][64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int bid = xcd_block(blockIdx.x, gridDim.x);
  momentum_scan_body<CT, MODE, DIM>(g, a, bid * kBlock + threadIdx.x, &qbuf[wid][0][lane]);
}

// List pass: exact test on every entry, two entries per iteration; lanes
// whose list is invalid or overflowed scan the cells with inline pairs.
template <typename CT, int MODE>
__global__ __launch_bounds__(kBlock) void momentum_list(Grid g, PassArgs<float> a) {
  const int bid = xcd_block(blockIdx.x, gridDim.x);
  const int i = bid * kBlock + threadIdx.x;
  const bool valid = i < a.dst.n;
  const bool local = valid && (a.dst.tag == nullptr || a.dst.tag[i] == 0);
  const MomDst<CT> D = mom_dst<CT>(a, i, valid);
  const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
  const double r = (double)D.pd.w * a.rs;
  const double r2d = r * r;
  const float r2f = (float)r2d;
  const int64_t ndst = a.dst.n;
  int total = 0;
  if (local)
    for (int s = 0; s < a.nsrc; ++s) total += a.nl_count[(int64_t)s * ndst + i];
  const bool use = local && total <= a.nl_max && list_valid(a.disp, a.nl_skin, r);
  const bool scan = local && !use;
  // an exact list read at the positions it was built at needs no re-test
  const bool retest = a.nl_skin != 0.f || a.disp != nullptr;
  int cx = 0, cy = 0, cz = 0;
  if (scan) cell_xyz(a.dst.cell[i], g, cx, cy, cz);
  const int ring = ring_of(g, a.disp);
  const Cull cu = make_cull(D.pd, r + (a.disp ? (double)*a.disp : 0.0), g);
  CT acc[3] = {CT(0), CT(0), CT(0)};
  CT dv[3] = {CT(0), CT(0), CT(0)};
  int lpos = 0;
  for (int s = 0; s < a.nsrc; ++s) {
    const View<float> S = a.src[s];
    const bool do_m = (a.flags[s] & SPH_SRC_MOMENTUM) != 0 && a.do_mom;
    const bool do_x = (a.flags[s] & SPH_SRC_XSPH) != 0 && a.do_xsph;
    const int cnt_s = local ? a.nl_count[(int64_t)s * ndst + i] : 0;
    if (!do_m && !do_x) {
      lpos += cnt_s;
      continue;
    }
    CT am[3] = {CT(0), CT(0), CT(0)};
    CT ax[3] = {CT(0), CT(0), CT(0)};
    const int* lst = a.nl_idx + (int64_t)lpos * ndst + i;
    const int m = wave_max(use ? cnt_s : 0);
    for (int k = 0; k < m; k += 2) {
      const bool e0 = use && k < cnt_s, e1 = use && k + 1 < cnt_s;
      const int j0 = e0 ? lst[(int64_t)k * ndst] : 0;
      const int j1 = e1 ? lst[(int64_t)(k + 1) * ndst] : 0;
      // all six loads issued before the tests: their latencies overlap
      const float4 p0 = e0 ? S.pos4[j0] : kFar4;
      const float4 p1 = e1 ? S.pos4[j1] : kFar4;
      const float4 v0 = e0 ? S.vel4[j0] : zero4, w0 = e0 ? S.aux4[j0] : zero4;
      const float4 v1 = e1 ? S.vel4[j1] : zero4, w1 = e1 ? S.aux4[j1] : zero4;
      if (e0 && (!retest || accept_exact(p0, D.pd, r2f, r2d)))
        mom_pair<CT>(D, p0, v0, w0, do_m, do_x, am, ax);
      if (e1 && (!retest || accept_exact(p1, D.pd, r2f, r2d)))
        mom_pair<CT>(D, p1, v1, w1, do_m, do_x, am, ax);
    }
    lpos += cnt_s;
    if (__any(scan)) {
      for (int dz = -ring; dz <= ring; ++dz) {
        for (int dy = -ring; dy <= ring; ++dy) {
          int j0 = 0, j1 = 0;
          if (scan) row_range_cull(g, S.cstart, cx, cy + dy, cz + dz, ring, cu, j0, j1);
          for (int j = j0; j < j1; ++j) {
            const float4 ps = S.pos4[j];
            if (accept_exact(ps, D.pd, r2f, r2d))
              mom_pair<CT>(D, ps, S.vel4[j], S.aux4[j], do_m, do_x, am, ax);
          }
        }
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      acc[c] += am[c];
      dv[c] += ax[c];
    }
  }
  if (!valid) return;
  mom_epilogue<CT, MODE>(a, i, local, acc, dv, D.vd);
}

// ---------------------------------------------------------------------------
// packing and launchers
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void pack_kernel(View<float> a, float4* __restrict__ pos4,
                                                      float4* __restrict__ vel4,
                                                      float* __restrict__ pos3,
                                                      float4* __restrict__ pm4,
                                                      int* __restrict__ hrange,
                                                      int* __restrict__ mrange) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const bool valid = i < a.n;
  if (valid) {
    const float x = a.x[i], y = a.y[i], z = a.z[i];
    if (pos4 != nullptr) pos4[i] = make_float4(x, y, z, a.h[i]);
    if (vel4 != nullptr) vel4[i] = make_float4(a.u[i], a.v[i], a.w[i], a.m[i]);
    if (pm4 != nullptr) pm4[i] = make_float4(x, y, z, a.m[i]);
    if (pos3 != nullptr) {
      pos3[3 * (int64_t)i] = x;
      pos3[3 * (int64_t)i + 1] = y;
      pos3[3 * (int64_t)i + 2] = z;
    }
  }
  // uniformity of h: hrange starts as (h_0, h_0) (hrange_init); a particle
  // whose h differs stores its own h into hrange[1] (any such store makes
  // the pair unequal), so uniform h costs no store at all
  if (hrange != nullptr && valid) {
    const int hb = __float_as_int(a.h[i]);
    if (hb != __float_as_int(a.h[0])) hrange[1] = hb;
  }
  // the same for the mass (mrange)
  if (mrange != nullptr && valid) {
    const int mb = __float_as_int(a.m[i]);
    if (mb != __float_as_int(a.m[0])) mrange[1] = mb;
  }
}

// (h_0, h_0) and (m_0, m_0) into the uniformity pairs (either may be NULL)
__global__ void range_init(const float* __restrict__ h, int* __restrict__ hrange,
                           const float* __restrict__ m, int* __restrict__ mrange) {
  if (hrange != nullptr) {
    hrange[0] = __float_as_int(h[0]);
    hrange[1] = __float_as_int(h[0]);
  }
  if (mrange != nullptr) {
    mrange[0] = __float_as_int(m[0]);
    mrange[1] = __float_as_int(m[0]);
  }
}

// Compile-time dimension of the FP32-pair scan kernels (0: from the
// arguments).
template <typename CT>
static int scan_dim(int dim) {
  return sizeof(CT) == 4 && (dim == 2 || dim == 3) ? dim : 0;
}

template <typename CT, int DIM>
static void density_scan_launch(hipStream_t st, int nb, const Grid& G, const PassArgs<float>& a,
                                bool tait, bool build) {
#define SPH_DENS(K) hipLaunchKernelGGL(K, dim3(nb), dim3(kBlock), 0, st, G, a)
  if (build) {
    if (tait) SPH_DENS((density_scan<CT, true, true, DIM>));
    else SPH_DENS((density_scan<CT, false, true, DIM>));
  } else if (sizeof(CT) == 4 && a.lds_rows) {
    // stencil rows staged in LDS (opt-in; FP32 pairs)
    if (tait) SPH_DENS((density_scan_lds<true, DIM>));
    else SPH_DENS((density_scan_lds<false, DIM>));
  } else {
    if (tait) SPH_DENS((density_scan<CT, true, false, DIM>));
    else SPH_DENS((density_scan<CT, false, false, DIM>));
  }
#undef SPH_DENS
}

template <typename CT>
static void density_launch(hipStream_t st, const Grid& G, const PassArgs<float>& a, bool tait) {
  const int nb = nblocks_for(a.dst.n);
  const bool list = a.nl_idx != nullptr;
  const bool build = list && a.nl_mode == 0;
  if (list && !build) {
    if (tait)
      hipLaunchKernelGGL((density_list<CT, true>), dim3(nb), dim3(kBlock), 0, st, G, a);
    else
      hipLaunchKernelGGL((density_list<CT, false>), dim3(nb), dim3(kBlock), 0, st, G, a);
    return;
  }
  switch (scan_dim<CT>(a.dim)) {
    case 2: density_scan_launch<CT, 2>(st, nb, G, a, tait, build); break;
    case 3: density_scan_launch<CT, 3>(st, nb, G, a, tait, build); break;
    default: density_scan_launch<CT, 0>(st, nb, G, a, tait, build); break;
  }
}

int32_t fast_density(hipStream_t st, const sph_grid& g, double rmax, PassArgs<float> a,
                     bool tait, int pair_prec) {
  const bool build = a.nl_idx != nullptr && a.nl_mode == 0;
  if (a.nl_idx != nullptr && (a.nl_skin < 0.f || a.nl_max < 0)) {
    set_error("sph_density: negative list skin or capacity");
    return SPH_EINVAL;
  }
  // the ring of a list build covers the superset radius
  const Grid G = make_grid(g, build ? rmax * (1.0 + (double)a.nl_skin) : rmax);
  if (pair_prec == SPH_PAIR_F64)
    density_launch<double>(st, G, a, tait);
  else
    density_launch<float>(st, G, a, tait);
  return launch_status("sph_density");
}

template <int DIM>
static void density_pair_launch(hipStream_t st, const Grid& G, const PassArgs<float>& a0,
                                const PassArgs<float>& a1, bool tait) {
  const int nb0 = nblocks_for(a0.dst.n), nb1 = nblocks_for(a1.dst.n);
  if (nb0 + nb1 == 0) return;
  if (a0.lds_rows) {
    // stencil rows staged in LDS (opt-in)
    if (tait)
      hipLaunchKernelGGL((density_scan2_lds<true, DIM>), dim3(nb0 + nb1), dim3(kBlock), 0, st,
                         G, a0, a1, nb0);
    else
      hipLaunchKernelGGL((density_scan2_lds<false, DIM>), dim3(nb0 + nb1), dim3(kBlock), 0, st,
                         G, a0, a1, nb0);
  } else if (tait) {
    hipLaunchKernelGGL((density_scan2<float, true, DIM>), dim3(nb0 + nb1), dim3(kBlock), 0, st,
                       G, a0, a1, nb0);
  } else {
    hipLaunchKernelGGL((density_scan2<float, false, DIM>), dim3(nb0 + nb1), dim3(kBlock), 0,
                       st, G, a0, a1, nb0);
  }
}

int32_t fast_density_pair(hipStream_t st, const sph_grid& g, double rmax, PassArgs<float> a0,
                          PassArgs<float> a1, bool tait) {
  const Grid G = make_grid(g, rmax);
  switch (scan_dim<float>(a0.dim)) {
    case 2: density_pair_launch<2>(st, G, a0, a1, tait); break;
    case 3: density_pair_launch<3>(st, G, a0, a1, tait); break;
    default: density_pair_launch<0>(st, G, a0, a1, tait); break;
  }
  return launch_status("sph_density_multi");
}

template <typename CT, int MODE>
static void momentum_launch_mode(hipStream_t st, int nb, const Grid& G,
                                 const PassArgs<float>& a) {
  if (a.nl_idx != nullptr) {
    hipLaunchKernelGGL((momentum_list<CT, MODE>), dim3(nb), dim3(kBlock), 0, st, G, a);
    return;
  }
  switch (scan_dim<CT>(a.dim)) {
    case 2:
      hipLaunchKernelGGL((momentum_scan<CT, MODE, 2>), dim3(nb), dim3(kBlock), 0, st, G, a);
      break;
    case 3:
      hipLaunchKernelGGL((momentum_scan<CT, MODE, 3>), dim3(nb), dim3(kBlock), 0, st, G, a);
      break;
    default:
      hipLaunchKernelGGL((momentum_scan<CT, MODE, 0>), dim3(nb), dim3(kBlock), 0, st, G, a);
      break;
  }
}

template <typename CT>
static void momentum_launch(hipStream_t st, const Grid& G, const PassArgs<float>& a, int mode) {
  const int nb = nblocks_for(a.dst.n);
  switch (mode) {
    case 0: momentum_launch_mode<CT, 0>(st, nb, G, a); break;
    case 1: momentum_launch_mode<CT, 1>(st, nb, G, a); break;
    case 2: momentum_launch_mode<CT, 2>(st, nb, G, a); break;
    default: momentum_launch_mode<CT, 3>(st, nb, G, a); break;
  }
}

int32_t fast_momentum(hipStream_t st, const sph_grid& g, double rmax, PassArgs<float> a,
                      int mode, int pair_prec) {
  if (a.nl_idx != nullptr && (a.nl_count == nullptr || a.nl_max < 0 || a.nl_skin < 0.f)) {
    set_error("momentum pass: incomplete neighbour list");
    return SPH_EINVAL;
  }
  const Grid G = make_grid(g, rmax);
  if (pair_prec == SPH_PAIR_F64)
    momentum_launch<double>(st, G, a, mode);
  else
    momentum_launch<float>(st, G, a, mode);
  return launch_status(mode == 0 ? "sph_momentum_xsph" : "sph_wcsph_stage");
}

}  // namespace sph

using namespace sph;

extern "C" int32_t sph_pack(void* stream, const sph_view* a, void* pos4_dev, void* vel4_dev,
                            void* pos3_dev, void* pm4_dev, int32_t* hrange_dev,
                            int32_t* mrange_dev) {
  SPH_TRY(check_view(a, "sph_pack", vel4_dev != nullptr, false, false));
  if (a->dtype != SPH_F32) {
    set_error("sph_pack: packed copies exist for FP32 storage only");
    return SPH_EUNSUPPORTED;
  }
  if ((vel4_dev != nullptr || pm4_dev != nullptr || mrange_dev != nullptr) && a->n > 0 &&
      !a->m) {
    set_error("sph_pack: m is required for vel4, pm4 and mrange");
    return SPH_EINVAL;
  }
  if (a->n == 0) return SPH_OK;
  hipStream_t st = as_stream(stream);
  if (hrange_dev != nullptr || mrange_dev != nullptr)
    hipLaunchKernelGGL(range_init, dim3(1), dim3(1), 0, st, (const float*)a->h, hrange_dev,
                       (const float*)a->m, mrange_dev);
  hipLaunchKernelGGL(pack_kernel, dim3(nblocks_for(a->n)), dim3(kBlock), 0, st,
                     to_view<float>(*a), (float4*)pos4_dev, (float4*)vel4_dev, (float*)pos3_dev,
                     (float4*)pm4_dev, (int*)hrange_dev, (int*)mrange_dev);
  return launch_status("sph_pack");
}
