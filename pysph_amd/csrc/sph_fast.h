// Device building blocks of the packed FP32 neighbour passes, shared by the
// single-array kernels (sph_fast.hip) and the batched multi-scene kernels
// (sph_batch.hip).
//
// Packed FP32 fast path of the neighbour passes.
//
// Inputs are float4 copies per particle: pos4 = (x,y,z,h), vel4 = (u,v,w,m),
// aux4 = (rho,cs,p/rho^2,p) (the FP32 pair arithmetic reads the first three,
// one 12-byte load), optionally pm4 = (x,y,z,m) with the array's h range
// (a density source of uniform h: one load per pair), and the (x,y,z) copy
// pos3 that the candidate screen reads.  One lane per destination particle,
// the particles sorted by cell so that the lanes of a wave share stencil
// rows.
//
// Cell scans: a lane screens the candidates of its stencil rows (four per
// step, three 16-byte loads from pos3) and queues the accepted source
// indices in LDS; when a queue is full, and after each source array, the
// wave drains all queues in one wave-uniform loop, so the pair arithmetic
// runs with ~70% of lanes active instead of the ~40% that accept any one
// candidate.
//
// Neighbour lists: a density scan can record the accepted sources of every
// destination.  Without skin the list holds exactly the neighbours and the
// momentum pass of the same stage reads it without testing again.  With a
// skin the scan records every source within radius*(1+skin) (a superset):
// later passes apply the exact test to each entry, as long as no particle
// has moved more than skin*radius/2 (checked on the device; lanes where it
// fails search the cells again with a widened ring).
//
// Reference formulas: SPHRho (density_funcs.pyx:38-84), TaitEquation
// (eos_funcs.pyx:105-116), MomentumEquation (pressure_funcs.pyx:154-245),
// XSPHCorrection (xsph_funcs.pyx:31-88), cubic spline (kernels.pyx:345-449),
// integrator updates (integrator.py:472-514, :597-619).
#pragma once
#include <math.h>

#include "sph_passes.h"

namespace sph {

// LDS queue depth per lane and candidate loads in flight per lane; the
// momentum scans' 24 entries keep six 256-lane blocks within the 160 KiB of
// LDS, the occupancy their 80 VGPRs allow
#ifndef SPH_QUEUE
#define SPH_QUEUE 24
#endif
// LDS queue depth of the density scans (sources of varied h, list builds)
#ifndef SPH_DENS_QUEUE
#define SPH_DENS_QUEUE 16
#endif
#ifndef SPH_UNROLL
#define SPH_UNROLL 4
#endif
constexpr int kQueue = SPH_QUEUE;
constexpr int kDensQueue = SPH_DENS_QUEUE;
// occupancy target of the momentum scan kernels (waves per SIMD)
constexpr int kMomWaves = 6;
constexpr int kUnroll = SPH_UNROLL;
#ifndef SPH_DENS_UNROLL
#define SPH_DENS_UNROLL SPH_UNROLL
#endif
constexpr int kDensUnroll = SPH_DENS_UNROLL;  // candidates per screen step, density
// density sources of uniform h are summed without the LDS queue
#ifndef SPH_DENS_DIRECT
#define SPH_DENS_DIRECT 1
#endif
constexpr bool kDensDirect = SPH_DENS_DIRECT != 0;
// queue-free density sources of uniform mass are read from the (x,y,z) copy
#ifndef SPH_DENS_XYZ
#define SPH_DENS_XYZ 1
#endif
constexpr bool kDensXyz = SPH_DENS_XYZ != 0;
// a density pass over two sources of uniform h sums them row by row
// together (one source, or more than two, take the source-by-source loop)
constexpr int kDirectSrc = 2;
// occupancy target of the density scan kernels (waves per SIMD, 0: none):
// seven waves hold them within 72 VGPRs without spilling
#ifndef SPH_DENS_WAVES
#define SPH_DENS_WAVES 7
#endif
// momentum pairs of a source of uniform h use per-destination constants
#ifndef SPH_MOM_UNI
#define SPH_MOM_UNI 1
#endif
constexpr bool kMomUni = SPH_MOM_UNI != 0;
__device__ constexpr float4 kFar4 = {3.0e38f, 3.0e38f, 3.0e38f, 0.f};

// float4 gather at a non-negative index as a 32-bit byte offset from the
// array base, so the load takes the scalar-base + 32-bit-offset form
// (packed arrays hold at most kPackedMaxN = 2^27 particles).
__device__ __forceinline__ float4 ld4(const float4* __restrict__ base, int j) {
  return *(const float4*)((const char*)base + (unsigned)j * 16u);
}

// compile-time flag for generic lambdas
template <bool B>
struct BoolC {
  static constexpr bool value = B;
};

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
  return v;
}

// Exact neighbour test (FP32 screen, FP64 decision in the band).
__device__ __forceinline__ bool accept_exact(const float4& s, const float4& d, float r2f,
                                             double r2d) {
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

// Reciprocal and square root: hardware 1-ulp instructions for FP32 pair
// arithmetic, IEEE operations for FP64.
template <typename CT>
__device__ __forceinline__ CT frcp(CT x) {
  return CT(1) / x;
}
template <>
__device__ __forceinline__ float frcp<float>(float x) {
  return __builtin_amdgcn_rcpf(x);
}
template <typename CT>
__device__ __forceinline__ CT fsqrt(CT x) {
  return sqrt(x);
}
template <>
__device__ __forceinline__ float fsqrt<float>(float x) {
  return __builtin_amdgcn_sqrtf(x);
}

// Branch-free cubic spline shape W(q) (kernels.pyx:367-374).
template <typename CT>
__device__ __forceinline__ CT w_sel(CT q) {
  const CT t = CT(2) - q;
  const CT outer = CT(0.25) * t * t * t;
  const CT inner = CT(1) - CT(1.5) * (q * q) * (CT(1) - CT(0.5) * q);
  return q > CT(2) ? CT(0) : (q > CT(1) ? outer : inner);
}

// Branch-free gradient factor (kernels.pyx:403-412): grad W = r_ab * g.
template <typename CT>
__device__ __forceinline__ CT g_sel(CT q, CT r, CT h, CT inv) {
  const CT t = CT(2) - q;
  const CT outer = CT(-0.75) * t * t * frcp<CT>(h * r);
  const CT inner = CT(3) * (CT(0.75) * q - CT(1)) * (inv * inv);
  return q > CT(2) ? CT(0) : (q >= CT(1) ? outer : (q > CT(1e-14) ? inner : CT(0)));
}

// ---------------------------------------------------------------------------
// stencil rows with geometric culling
// ---------------------------------------------------------------------------
struct Cull {
  float xd, yd, zd, rc2, r;
};

// Distance from v to the extent of cell c along one axis; border cells
// reach to infinity because the binning clamps outside particles into them.
__device__ __forceinline__ float axis_gap(float v, int c, int n, int lo, float cs) {
  float gap = 0.f;
  if (c > 0) gap = fmaxf(gap, (float)(c + lo) * cs - v);
  if (c < n - 1) gap = fmaxf(gap, v - (float)(c + lo + 1) * cs);
  return gap;
}

// rc is widened for FP32 rounding of positions and cell boundaries.
__device__ __forceinline__ Cull make_cull(const float4& pd, double rc, const Grid& g) {
  Cull c;
  c.xd = pd.x;
  c.yd = pd.y;
  c.zd = pd.z;
  const float scale = fabsf(pd.x) + fabsf(pd.y) + fabsf(pd.z) + g.cs_f;
  const float r = (float)rc * 1.00001f + 8e-7f * scale;
  c.rc2 = r * r;
  c.r = r;
  return c;
}

// Stencil offsets [a, b] (within +-ring) along one axis whose cells can
// hold a point within the cull radius of v; border cells reach to infinity.
// The margin of the cull radius covers the FP32 rounding of the bounds, as
// in row_range_cull.
__device__ __forceinline__ void axis_span(float v, float r, int c, int n, int lo, float inv_cs,
                                          int ring, int& a, int& b) {
  const float last = (float)(n - 1);
  const float fa = fminf(fmaxf(floorf((v - r) * inv_cs) - (float)lo, 0.f), last);
  const float fb = fminf(fmaxf(floorf((v + r) * inv_cs) - (float)lo, 0.f), last);
  if (!(fa <= fb)) {  // NaN positions keep the ring
    a = -ring;
    b = ring;
    return;
  }
  a = max((int)fa - c, -ring);
  b = min((int)fb - c, ring);
}

// Particle range of the cells of grid row (y, z) within the stencil ring of
// cell column cx whose extent can hold a point within the cull radius.
// The cell-start indices [k0, k1] (read cstart[k0] and cstart[k1]) of that
// range; false when the row holds no candidate cell.
__device__ __forceinline__ bool row_cells_cull(const Grid& g, int cx, int y, int z, int ring,
                                               const Cull& cu, int& k0, int& k1) {
  if (y < 0 || y >= g.dims[1] || z < 0 || z >= g.dims[2]) return false;
  const float gy = axis_gap(cu.yd, y, g.dims[1], g.lo[1], g.cs_f);
  const float gz = axis_gap(cu.zd, z, g.dims[2], g.lo[2], g.cs_f);
  const float w2 = cu.rc2 - gy * gy - gz * gz;
  if (!(w2 >= 0.f)) return false;
  const float w = __builtin_amdgcn_sqrtf(w2);
  const float last = (float)(g.dims[0] - 1);
  const float fa = fminf(fmaxf(floorf((cu.xd - w) * g.inv_cs_f) - (float)g.lo[0], 0.f), last);
  const float fb = fminf(fmaxf(floorf((cu.xd + w) * g.inv_cs_f) - (float)g.lo[0], 0.f), last);
  int x0 = max(cx - ring, 0), x1 = min(cx + ring, g.dims[0] - 1);
  if (fa > (float)x1 || fb < (float)x0) return false;  // NaN positions keep the ring
  if (fa > (float)x0) x0 = (int)fa;
  if (fb < (float)x1) x1 = (int)fb;
  const int base = (z * g.dims[1] + y) * g.dims[0];
  k0 = base + x0;
  k1 = base + x1 + 1;
  return true;
}

__device__ __forceinline__ void row_range_cull(const Grid& g, const int* __restrict__ cstart,
                                               int cx, int y, int z, int ring, const Cull& cu,
                                               int& j0, int& j1) {
  j0 = j1 = 0;
  if (y < 0 || y >= g.dims[1] || z < 0 || z >= g.dims[2]) return;
  const float gy = axis_gap(cu.yd, y, g.dims[1], g.lo[1], g.cs_f);
  const float gz = axis_gap(cu.zd, z, g.dims[2], g.lo[2], g.cs_f);
  const float w2 = cu.rc2 - gy * gy - gz * gz;
  if (!(w2 >= 0.f)) return;
  // 1-ulp hardware square root: the cull radius carries a 1e-5 margin
  const float w = __builtin_amdgcn_sqrtf(w2);
  // cells of [x - w, x + w]; the border cells reach to infinity (particles
  // outside the grid are clamped into them), so clamp the range into the grid
  const float last = (float)(g.dims[0] - 1);
  const float fa = fminf(fmaxf(floorf((cu.xd - w) * g.inv_cs_f) - (float)g.lo[0], 0.f), last);
  const float fb = fminf(fmaxf(floorf((cu.xd + w) * g.inv_cs_f) - (float)g.lo[0], 0.f), last);
  int x0 = max(cx - ring, 0), x1 = min(cx + ring, g.dims[0] - 1);
  if (fa > (float)x1 || fb < (float)x0) return;  // NaN positions keep the ring
  if (fa > (float)x0) x0 = (int)fa;
  if (fb < (float)x1) x1 = (int)fb;
  const int base = (z * g.dims[1] + y) * g.dims[0];
  j0 = cstart[base + x0];
  j1 = cstart[base + x1 + 1];
}

// Raw buffer resource over a source's packed positions, the (x,y,z) copy
// when the source has one (n * 12 bytes), else pos4 (n * 16 bytes); gfx9
// resource word 3.  A load at or past the end returns zeros instead of
// touching memory, so candidate loads need no index clamp.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t cand_rsrc(const View<float>& S) {
  if (S.pos3 != nullptr)
    return __builtin_amdgcn_make_buffer_rsrc((void*)S.pos3, (short)0,
                                             (int)((unsigned)S.n * 12u), 0x00020000);
  return __builtin_amdgcn_make_buffer_rsrc((void*)S.pos4, (short)0,
                                           (int)((unsigned)S.n * 16u), 0x00020000);
}

// (x, y, z) of candidate slot u of a block starting at byte offset off.
__device__ __forceinline__ float3 ld_xyz(__amdgpu_buffer_rsrc_t rs, int off) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b96(rs, off, 0, 0);
  return make_float3(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]));
}

// Screen the candidates j0 .. j0+kUnroll-1 (those below j1) of one stencil
// row against destination pd and push the accepted ones into the lane's
// LDS queue q (entries 64 apart), branch-free: every lane loads kUnroll
// candidates through the bounds-checked buffer view `rs` (cand_rsrc: one
// base offset, the slots as immediate offsets, three 16-byte loads per four
// candidates from the (x,y,z) copy, else one 12-byte load each from pos4;
// slots past the row end are loaded but never accepted), writes every
// candidate into the next queue slot and
// advances its count only on acceptance.  The caller keeps
// cnt <= kQueue - kUnroll on entry.  The exact test (Within32) runs its FP64
// branch only when a candidate of some lane falls in the +-2e-6 band; with
// `super` the test is the list-build superset d2 < rl2f instead.
template <int U = kUnroll>
__device__ __forceinline__ void screen_block(const View<float>& S, __amdgpu_buffer_rsrc_t rs,
                                             int j0, int j1, const float4& pd, float r2lo,
                                             float r2hi, double r2d, bool super, float rl2f,
                                             int* q, int& cnt) {
  float3 c[U];
  if (S.pos3 != nullptr && U % 4 == 0) {
    // (x,y,z) copy: four candidates are 48 contiguous bytes, three loads
    int off = (int)((unsigned)j0 * 12u);
    // keep off opaque, so that the slots become immediate load offsets
    // (an empty register constraint: no instruction is emitted)
    asm("" : "+v"(off));
#pragma unroll
    for (int g = 0; g < U / 4; ++g) {
      float w[12];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b128(rs, off + 48 * g + 16 * k, 0, 0);
#pragma unroll
        for (int e = 0; e < 4; ++e) w[4 * k + e] = __uint_as_float(v[e]);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) c[4 * g + u] = make_float3(w[3 * u], w[3 * u + 1], w[3 * u + 2]);
    }
  } else {
    int off = (int)((unsigned)j0 * 16u);
    asm("" : "+v"(off));
#pragma unroll
    for (int u = 0; u < U; ++u) c[u] = ld_xyz(rs, off + 16 * u);
  }
  const int rem = j1 - j0;
  bool ok[U], bnd[U];
  bool band = false;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const float dx = c[u].x - pd.x, dy = c[u].y - pd.y, dz = c[u].z - pd.z;
    const float d2f = dx * dx + dy * dy + dz * dz;
    const bool in = u < rem;
    if (super) {
      ok[u] = in && d2f < rl2f;
      bnd[u] = false;
    } else {
      ok[u] = in && d2f < r2lo;
      bnd[u] = in && !(d2f < r2lo) && !(d2f > r2hi);
    }
    band = band || bnd[u];
  }
  if (__any(band)) {
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (bnd[u])
        ok[u] = exact_within((double)c[u].x, (double)c[u].y, (double)c[u].z, (double)pd.x,
                             (double)pd.y, (double)pd.z, r2d);
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    q[cnt * 64] = j0 + u;
    cnt += ok[u] ? 1 : 0;
  }
}

// 4^(1/3): with v = 4^(1/3) max(1 - q, 0), v^3 = 4 max(1 - q, 0)^3.
constexpr float kCbrt4 = 1.5874010519681994f;

// records a wave stages for density_lds_row: its 4 KB LDS region (the
// queue area the queue-free density path does not use)
constexpr int kStage = 256;

// Summation density over U candidate slots of one stencil row of a source
// whose h is uniform: no queue.  `off` is the byte offset of the first slot
// in the source's candidate records, loaded through the bounds-checked
// buffer view `rs` (one base offset, the slots as immediate offsets), and
// `rem` the bytes left in the lane's range, so slot u is a candidate when
// u * stride < rem.  With P3 the records are the 12-byte (x, y, z) copy
// (pos3: four slots in three 16-byte loads, two in a 16-byte and an 8-byte
// load) and every source particle has the mass `mb` (the source's mass is
// uniform); otherwise they are the 16-byte (x, y, z, m) records of pm4.
// Both give the same products, so the same bits; the 12-byte records move a
// quarter fewer bytes through the vector memory pipeline, which the pass
// keeps busy.  With FROM_LDS the records are (x, y, z, m) entries of the
// wave's staged row `stage` (density_lds_row), `off` and `rem` in records
// instead of bytes.  Every lane screens its
// slots like screen_block (FP32 test, FP64 inside the +-2e-6 band) and adds
// m 4w(q) of every accepted one, in row order (the order a queue would
// drain them in).  4w is four times the cubic spline's shape
// (kernels.pyx:367-374), 4 - 6 q^2 + 3 q^3 for q <= 1 and (2 - q)^3 for
// 1 < q < 2, written branch-free as t^3 - v^3 with t = max(2 - q, 0) and
// v = 4^(1/3) max(1 - q, 0) (t^3 - 4 (1 - q)^3 is the inner polynomial),
// q = |r_a - r_b| / h_ab from the screened squared distance; `ninv` =
// -1/h_ab and `nkinv` = -4^(1/3)/h_ab.  The caller scales the row sums by
// fac / 4 (the normalisation of a source of uniform h is the same for every
// pair), which saves the per-pair multiplications by the constants.
// Evaluating w for every slot costs about what queueing the accepted ones
// costs, and it removes the drain loop with its dependent index read,
// gather and memory wait per pair.  FP32 pair arithmetic only.
template <int U, bool P3, bool FROM_LDS = false>
__device__ __forceinline__ void density_direct_block(__amdgpu_buffer_rsrc_t rs, int off, int rem,
                                                     const float4& pd, float r2lo, float r2hi,
                                                     double r2d, float ninv, float nkinv,
                                                     float mb, float& nr,
                                                     const float4* stage = nullptr) {
  static_assert(!P3 || U % 4 == 0 || U == 2, "pos3 blocks: groups of four slots, or two");
  constexpr int kStride = FROM_LDS ? 1 : (P3 ? 12 : 16);
  float4 c[U];
  if constexpr (FROM_LDS) {
#pragma unroll
    for (int u = 0; u < U; ++u) c[u] = stage[(off + u) & (kStage - 1)];
  } else if constexpr (P3 && U == 2) {
    // two slots: 24 bytes, one 16-byte and one 8-byte load
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0);
    const auto t = __builtin_amdgcn_raw_buffer_load_b64(rs, off + 16, 0, 0);
    c[0] = make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]), mb);
    c[1] = make_float4(__uint_as_float(v[3]), __uint_as_float(t[0]), __uint_as_float(t[1]), mb);
  } else if constexpr (P3) {
#pragma unroll
    for (int g = 0; g < U / 4; ++g) {
      float w[12];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b128(rs, off + 48 * g + 16 * k, 0, 0);
#pragma unroll
        for (int e = 0; e < 4; ++e) w[4 * k + e] = __uint_as_float(v[e]);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        c[4 * g + u] = make_float4(w[3 * u], w[3 * u + 1], w[3 * u + 2], mb);
    }
  } else {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const auto v = __builtin_amdgcn_raw_buffer_load_b128(rs, off + 16 * u, 0, 0);
      c[u] = make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]),
                         __uint_as_float(v[3]));
    }
  }
  bool ok[U], bnd[U];
  float d2[U];
  bool band = false;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const float dx = c[u].x - pd.x, dy = c[u].y - pd.y, dz = c[u].z - pd.z;
    d2[u] = dx * dx + dy * dy + dz * dz;
    const bool in = kStride * u < rem;
    const bool lo = d2[u] < r2lo;
    const bool hi = d2[u] <= r2hi;
    ok[u] = in && lo;
    // lo implies hi (r2lo < r2hi); a NaN distance is neither
    bnd[u] = in && (hi != lo);
    band = band || bnd[u];
  }
  if (__any(band)) {
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (bnd[u])
        ok[u] = exact_within((double)c[u].x, (double)c[u].y, (double)c[u].z, (double)pd.x,
                             (double)pd.y, (double)pd.z, r2d);
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const float s = __builtin_amdgcn_sqrtf(d2[u]);
    // t = max(2 - q, 0) makes w exactly 0 from q = 2 on
    const float t = fmaxf(fmaf(ninv, s, 2.f), 0.f);
    const float v = fmaxf(fmaf(nkinv, s, kCbrt4), 0.f);
    const float w4 = fmaf(-(v * v), v, t * t * t);
    nr = fmaf(ok[u] ? c[u].w : 0.f, w4, nr);
  }
}

// The candidate records of a density source of uniform h for
// density_direct_block: the (x, y, z) copy when the source also has a
// uniform mass (returned in mb), else the (x, y, z, m) records.
struct DirectSrc {
  __amdgpu_buffer_rsrc_t rs;
  float mb;
  bool p3;
};

__device__ __forceinline__ DirectSrc direct_src(const View<float>& S) {
  DirectSrc d;
  d.p3 = kDensXyz && S.pos3 != nullptr && S.mrange != nullptr && S.mrange[0] == S.mrange[1];
  d.mb = d.p3 ? __int_as_float(S.mrange[0]) : 0.f;
  d.rs = d.p3 ? __builtin_amdgcn_make_buffer_rsrc((void*)S.pos3, (short)0,
                                                  (int)((unsigned)S.n * 12u), 0x00020000)
              : __builtin_amdgcn_make_buffer_rsrc((void*)S.pm4, (short)0,
                                                  (int)((unsigned)S.n * 16u), 0x00020000);
  return d;
}

// Every candidate block of one stencil row [j0, j1) of a direct source.
// From the (x, y, z) copy: blocks of four slots while some lane of the wave
// has four candidates left, then blocks of two, so that fewer slots past
// the lanes' row ends are loaded and screened (the wave runs each block for
// all its lanes; 2^24 cube: 8 % fewer slots).  From the (x, y, z, m)
// records: blocks of four (the second block shape costs that kernel more
// registers than it saves).
__device__ __forceinline__ void density_direct_row(const DirectSrc& d, int j0, int j1,
                                                   const float4& pd, float r2lo, float r2hi,
                                                   double r2d, float inv, float& nr) {
  // byte offsets; off is kept opaque, so that the slots become immediate
  // load offsets (an empty register constraint: no instruction is emitted)
  if (d.p3) {
    int off = (int)((unsigned)j0 * 12u);
    const int end = (int)((unsigned)j1 * 12u);
    asm("" : "+v"(off));
    while (__any(end - off >= 48)) {
      density_direct_block<4, true>(d.rs, off, end - off, pd, r2lo, r2hi, r2d, -inv,
                                    -kCbrt4 * inv, d.mb, nr);
      off += 48;
    }
    while (__any(off < end)) {
      density_direct_block<2, true>(d.rs, off, end - off, pd, r2lo, r2hi, r2d, -inv,
                                    -kCbrt4 * inv, d.mb, nr);
      off += 24;
    }
  } else {
    int off = (int)((unsigned)j0 * 16u);
    const int end = (int)((unsigned)j1 * 16u);
    asm("" : "+v"(off));
    do {
      density_direct_block<4, false>(d.rs, off, end - off, pd, r2lo, r2hi, r2d, -inv,
                                     -kCbrt4 * inv, 0.f, nr);
      off += 64;
    } while (__any(off < end));
  }
}

// Minimum / maximum over the wave (DPP row shifts within rows of 16 lanes,
// then row broadcasts; lane 63 ends with the result).
__device__ __forceinline__ int wave_min_dpp(int v) {
  v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, 0x111, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, 0x112, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, 0x114, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, 0x118, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, 0x142, 0xa, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, 0x143, 0xc, 0xf, false));
  return __builtin_amdgcn_readlane(v, 63);
}
__device__ __forceinline__ int wave_max_dpp(int v) {
  v = max(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x111, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x112, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x114, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x118, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x142, 0xa, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x143, 0xc, 0xf, false));
  return __builtin_amdgcn_readlane(v, 63);
}

static_assert(kStage * 16 == kDensQueue * 64 * 4, "one wave's LDS region");

// One stencil row of a direct source through LDS: the wave loads the union
// of its lanes' ranges [lo, hi) of (x, y, z, m) records into its region with
// coalesced 16-byte loads, then every lane screens its own range from LDS
// in the same slot order as density_direct_row (the same sums).  False,
// with nothing done, when the union exceeds the region or a lane's range is
// not inside it.
__device__ __forceinline__ bool density_lds_row(const DirectSrc& d,
                                                const float4* __restrict__ pm4, float4* stage,
                                                int j0, int j1, const float4& pd, float r2lo,
                                                float r2hi, double r2d, float inv, float& nr) {
  const bool has = j0 < j1;
  const int lo = wave_min_dpp(has ? j0 : INT_MAX);
  const int hi = wave_max_dpp(has ? j1 : INT_MIN);
  if (lo >= hi || hi - lo > kStage) return false;
  const int nu = hi - lo;
  if (!__all(!has || (j0 >= lo && j1 <= hi))) return false;
  const int lane = threadIdx.x & 63;
  for (int k = lane; k < nu; k += 64) stage[k] = pm4[lo + k];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  int o = has ? j0 - lo : 0;
  int rem = has ? j1 - j0 : 0;
  while (__any(rem >= 4)) {
    density_direct_block<4, false, true>(d.rs, o, rem, pd, r2lo, r2hi, r2d, -inv, -kCbrt4 * inv,
                                         0.f, nr, stage);
    o += 4;
    rem -= 4;
  }
  while (__any(rem > 0)) {
    density_direct_block<2, false, true>(d.rs, o, rem, pd, r2lo, r2hi, r2d, -inv, -kCbrt4 * inv,
                                         0.f, nr, stage);
    o += 2;
    rem -= 2;
  }
  // every lane's reads are done before the region is written again
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  return true;
}

// Whether a list built with the given skin still holds every neighbour of a
// destination of search radius r after a displacement of at most disp.
__device__ __forceinline__ bool list_valid(const float* disp, float skin, double r) {
  if (disp == nullptr) return true;
  return 2.0 * (double)(*disp) <= (double)skin * r;
}

// ---------------------------------------------------------------------------
// density
// ---------------------------------------------------------------------------
// m W(r_a - r_b, (h_a + h_b) / 2); ps carries the source position.
template <typename CT>
__device__ __forceinline__ CT w_term_hb(const float4& pd, const float4& ps, CT hd, CT hb, CT m,
                                        int dim) {
  CT rx, ry, rz;
  const CT d2 = pair_d2<CT>(pd, ps, rx, ry, rz);
  const CT hab = CT(0.5) * (hd + hb);
  const CT inv = frcp<CT>(hab);
  const CT qq = fsqrt<CT>(d2) * inv;
  return m * (w_sel<CT>(qq) * h_dim_inv<CT>(inv, dim));
}

template <typename CT>
__device__ __forceinline__ CT w_term(const float4& pd, const float4& ps, CT hd, CT m, int dim) {
  return w_term_hb<CT>(pd, ps, hd, (CT)ps.w, m, dim);
}

template <bool TAIT>
__device__ __forceinline__ void density_epilogue(const PassArgs<float>& a, int i, bool local,
                                                 double rho) {
  if (!local) {
    a.rho[i] = 0.f;
    if (TAIT) {
      a.p[i] = 0.f;
      a.cs[i] = 0.f;
    }
    if (a.dst.aux4 != nullptr) a.dst.aux4[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    return;
  }
  const float rhof = (float)rho;
  a.rho[i] = rhof;
  if (TAIT) {
    double p, cs;
    tait((double)rhof, a.co, a.ro, a.gamma, a.B, p, cs);
    const float pf = (float)p;
    a.p[i] = pf;
    a.cs[i] = (float)cs;
    if (a.dst.aux4 != nullptr)
      a.dst.aux4[i] = make_float4(rhof, (float)cs,
                                  (float)((double)pf / ((double)rhof * (double)rhof)), pf);
  }
}

// Cell scan of destination i (all lanes of the wave call it together; q
// is the lane's LDS queue, kDensQueue ints 64 apart); BUILD also records the
// list (superset radius).
// DIM > 0 fixes the dimension at compile time (the kernels of dimensions 2
// and 3); DIM = 0 reads it from the arguments.
template <typename CT, bool TAIT, bool BUILD, int DIM = 0, bool LDS = false>
__device__ __forceinline__ void density_scan_body(const Grid& g, const PassArgs<float>& a,
                                                  int i, int* q) {
  const bool valid = i < a.dst.n;
  const bool local = valid && (a.dst.tag == nullptr || a.dst.tag[i] == 0);
  const float4 pd = valid ? a.dst.pos4[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  const double r = (double)pd.w * a.rs;
  const double r2d = r * r;
  const float r2f = (float)r2d;
  // a list built without skin holds exactly the neighbours
  const bool exact_list = !BUILD || a.nl_skin == 0.f;
  const double rl = BUILD ? r * (1.0 + (double)a.nl_skin) : r;
  const float rl2f = (float)(rl * rl) * 1.00001f;
  const float r2lo = r2f * 0.999998f, r2hi = r2f * 1.000002f;
  const bool write_list = BUILD && local;
  int cx = 0, cy = 0, cz = 0;
  if (local) cell_xyz(a.dst.cell[i], g, cx, cy, cz);
  const int ring = ring_of(g, a.disp);
  const double dd = a.disp ? (double)*a.disp : 0.0;
  const Cull cu = make_cull(pd, rl + dd, g);
  int dylo, dyhi, dzlo, dzhi;
  axis_span(cu.yd, cu.r, cy, g.dims[1], g.lo[1], g.inv_cs_f, ring, dylo, dyhi);
  axis_span(cu.zd, cu.r, cz, g.dims[2], g.lo[2], g.inv_cs_f, ring, dzlo, dzhi);
  const int dim = DIM ? DIM : a.dim;
  const CT hd = (CT)pd.w;
  const int64_t ndst = a.dst.n;
  CT rho = CT(0);
  int lpos = 0;
  if constexpr (kDensDirect && !BUILD && sizeof(CT) == 4) {
    // every source of uniform h: the stencil rows outside, the sources
    // inside, so that a row's culled cell range is computed once for all
    // sources; each source keeps its own sum, added in source order (the
    // same bits as summing source by source)
    bool all_uni = a.nsrc >= 2 && a.nsrc <= kDirectSrc;
    for (int s = 0; s < a.nsrc; ++s) {
      const View<float> S = a.src[s];
      all_uni = all_uni && S.pm4 != nullptr && S.hrange != nullptr &&
                S.hrange[0] == S.hrange[1];
    }
    if (all_uni) {
      float inv[kDirectSrc], fac[kDirectSrc], nrs[kDirectSrc];
      bool scan[kDirectSrc];
      DirectSrc ds[kDirectSrc];
#pragma unroll
      for (int s = 0; s < kDirectSrc; ++s) {
        ds[s] = direct_src(a.src[s < a.nsrc ? s : 0]);
        nrs[s] = 0.f;
        scan[s] = s < a.nsrc && !(s == a.cache_src && a.cache_mode == 2);
        const float hb = s < a.nsrc ? __int_as_float(a.src[s].hrange[0]) : 1.f;
        const float hab = 0.5f * ((float)hd + hb);
        inv[s] = frcp<float>(hab);
        fac[s] = h_dim_inv<float>(inv[s], dim);
      }
      for (int dz = -ring; dz <= ring; ++dz) {
        const bool zok = local && dz >= dzlo && dz <= dzhi;
        if (!__any(zok)) continue;
        for (int dy = -ring; dy <= ring; ++dy) {
          const bool yok = zok && dy >= dylo && dy <= dyhi;
          if (!__any(yok)) continue;
          int k0 = 0, k1 = 0;
          const bool row = yok && row_cells_cull(g, cx, cy + dy, cz + dz, ring, cu, k0, k1);
#pragma unroll
          for (int s = 0; s < kDirectSrc; ++s) {
            if (!scan[s]) continue;
            const View<float> S = a.src[s];
            int j0 = 0, j1 = 0;
            if (row) {
              j0 = S.cstart[k0];
              j1 = S.cstart[k1];
            }
            if (!__any(j0 < j1)) continue;
            if (!(LDS && density_lds_row(ds[s], S.pm4, (float4*)(q - (threadIdx.x & 63)), j0,
                                         j1, pd, r2lo, r2hi, r2d, inv[s], nrs[s])))
              density_direct_row(ds[s], j0, j1, pd, r2lo, r2hi, r2d, inv[s], nrs[s]);
          }
        }
      }
#pragma unroll
      for (int s = 0; s < kDirectSrc; ++s) {
        if (s >= a.nsrc) continue;
        CT nr = (CT)(nrs[s] * (0.25f * fac[s]));
        if (s == a.cache_src && a.cache_mode == 2) {
          if (local) nr = (CT)a.cache_sums[i];
        } else if (s == a.cache_src && a.cache_mode == 1 && valid) {
          a.cache_sums[i] = (double)nr;
        }
        rho += nr;
      }
      if (!valid) return;
      density_epilogue<TAIT>(a, i, local, (double)(rho * (CT)a.fac1));
      return;
    }
  }
  for (int s = 0; s < a.nsrc; ++s) {
    const View<float> S = a.src[s];
    const __amdgpu_buffer_rsrc_t rs = cand_rsrc(S);
    CT nr = CT(0);
    if (!BUILD && s == a.cache_src && a.cache_mode == 2) {
      // stored partial sum of a static source (exact round trip via double)
      if (local) nr = (CT)a.cache_sums[i];
      rho += nr;
      continue;
    }
    int cnt = 0, cnt_s = 0;
    // a source of uniform h with an (x,y,z,m) copy: one 16-byte record per
    // pair, h_b from the source's h-uniformity pair (the same value, the
    // same bits)
    const bool uni_h = exact_list && S.pm4 != nullptr && S.hrange != nullptr &&
                       S.hrange[0] == S.hrange[1];
    const CT h_src = uni_h ? (CT)__int_as_float(S.hrange[0]) : CT(0);
    if (kDensDirect && !BUILD && uni_h && sizeof(CT) == 4) {
      // queue-free summation (density_direct_block): h_ab, 1/h_ab and the
      // kernel normalisation are fixed per destination and source
      const float hab = 0.5f * ((float)hd + (float)h_src);
      const float inv = frcp<float>(hab);
      const float fac = h_dim_inv<float>(inv, dim);
      float nrf = 0.f;
      const DirectSrc d = direct_src(S);
      for (int dz = -ring; dz <= ring; ++dz) {
        const bool zok = local && dz >= dzlo && dz <= dzhi;
        if (!__any(zok)) continue;
        for (int dy = -ring; dy <= ring; ++dy) {
          const bool yok = zok && dy >= dylo && dy <= dyhi;
          if (!__any(yok)) continue;
          int j0 = 0, j1 = 0;
          if (yok) row_range_cull(g, S.cstart, cx, cy + dy, cz + dz, ring, cu, j0, j1);
          if (__any(j0 < j1)) {
            if (!(LDS && density_lds_row(d, S.pm4, (float4*)(q - (threadIdx.x & 63)), j0, j1,
                                         pd, r2lo, r2hi, r2d, inv, nrf)))
              density_direct_row(d, j0, j1, pd, r2lo, r2hi, r2d, inv, nrf);
          }
        }
      }
      nr = (CT)(nrf * (0.25f * fac));
      if (s == a.cache_src && a.cache_mode == 1 && valid) a.cache_sums[i] = (double)nr;
      rho += nr;
      continue;
    }
    auto drain = [&]() {
      if (uni_h) {
        for (int k = 0; __any(k < cnt); ++k) {
          if (k < cnt) {
            const int j = q[k * 64];
            const float4 pm = ld4(S.pm4, j);
            nr += w_term_hb<CT>(pd, pm, hd, h_src, (CT)pm.w, dim);
            if (write_list && lpos + k < a.nl_max) a.nl_idx[(int64_t)(lpos + k) * ndst + i] = j;
          }
        }
      } else {
        for (int k = 0; __any(k < cnt); ++k) {
          if (k < cnt) {
            const int j = q[k * 64];
            const float4 ps = ld4(S.pos4, j);
            if (exact_list || accept_exact(ps, pd, r2f, r2d))
              nr += w_term<CT>(pd, ps, hd, (CT)ld4(S.vel4, j).w, dim);
            if (write_list && lpos + k < a.nl_max) a.nl_idx[(int64_t)(lpos + k) * ndst + i] = j;
          }
        }
      }
      lpos += cnt;
      cnt_s += cnt;
      cnt = 0;
    };
    for (int dz = -ring; dz <= ring; ++dz) {
      const bool zok = local && dz >= dzlo && dz <= dzhi;
      if (!__any(zok)) continue;
      for (int dy = -ring; dy <= ring; ++dy) {
        const bool yok = zok && dy >= dylo && dy <= dyhi;
        if (!__any(yok)) continue;
        int j0 = 0, j1 = 0;
        if (yok) row_range_cull(g, S.cstart, cx, cy + dy, cz + dz, ring, cu, j0, j1);
        while (__any(j0 < j1)) {
          if (__any(cnt > kDensQueue - kDensUnroll)) drain();
          screen_block<kDensUnroll>(S, rs, j0, j1, pd, r2lo, r2hi, r2d, !exact_list, rl2f, q,
                                    cnt);
          j0 += kDensUnroll;
        }
      }
    }
    drain();
    if (!BUILD && s == a.cache_src && a.cache_mode == 1 && valid) a.cache_sums[i] = (double)nr;
    rho += nr;
    if (BUILD && valid) a.nl_count[(int64_t)s * ndst + i] = local ? cnt_s : 0;
  }
  if (!valid) return;
  density_epilogue<TAIT>(a, i, local, (double)(rho * (CT)a.fac1));
}
// ---------------------------------------------------------------------------
// momentum + XSPH (+ integrator update)
// ---------------------------------------------------------------------------
template <typename CT>
struct MomDst {
  CT ha, rhoa, ca, Aa, alpha, beta, eta2;
  float4 pd, vd;
  int dim;
};

// One pair of the momentum (DM) and XSPH (DX) sums.  The roles are fixed
// at compile time, so no branch separates a pair's gathers: the compiler
// would otherwise sink the loads that only one role uses behind that
// role's branch, and wait for memory twice per pair.
template <typename CT, bool DM, bool DX>
__device__ __forceinline__ void mom_pair_r(const MomDst<CT>& D, const float4& ps,
                                           const float4& vs, const float4& as, CT* am, CT* ax) {
  CT rx, ry, rz;
  const CT d2 = pair_d2<CT>(D.pd, ps, rx, ry, rz);
  const CT hab = CT(0.5) * (D.ha + (CT)ps.w);
  const CT inv = frcp<CT>(hab);
  const CT rr = fsqrt<CT>(d2);
  const CT qq = rr * inv;
  const CT fac = h_dim_inv<CT>(inv, D.dim);
  const CT mb = (CT)vs.w;
  const CT rhob = (CT)as.x;
  const CT irhoab = frcp<CT>(CT(0.5) * (D.rhoa + rhob));
  // the velocity difference serves both sums (v_b - v_a = -(v_a - v_b)
  // exactly), and every source field is read before the branches, so the
  // three gathers of a pair go out together
  const CT vx = (CT)D.vd.x - (CT)vs.x, vy = (CT)D.vd.y - (CT)vs.y, vz = (CT)D.vd.z - (CT)vs.z;
  const CT Ab = sizeof(CT) == 8 ? (CT)as.w / (rhob * rhob) : (CT)as.z;
  const CT cab = CT(0.5) * (D.ca + (CT)as.y);
  if (DM) {
    const CT dot = vx * rx + vy * ry + vz * rz;
    const CT mu = hab * dot * frcp<CT>(d2 + D.eta2 * hab * hab);
    const CT piab = (-D.alpha * cab * mu + D.beta * mu * mu) * irhoab;
    const CT tmp = D.Aa + Ab + (dot < CT(0) ? piab : CT(0));
    const CT gf = -mb * tmp * (g_sel<CT>(qq, rr, hab, inv) * fac);
    am[0] += gf * rx;
    am[1] += gf * ry;
    am[2] += gf * rz;
  }
  if (DX) {
    const CT temp = mb * (w_sel<CT>(qq) * fac) * irhoab;
    ax[0] += temp * -vx;
    ax[1] += temp * -vy;
    ax[2] += temp * -vz;
  }
}

// Constants of the pairs of one destination with a source whose h is
// uniform (h_b fixed): h_ab, 1/h_ab, the normalisation 1/h_ab^dim and the
// gradient factors folded with it.
struct MomUni {
  float hab, inv, fac, eta2hh, kout, kin;
};

__device__ __forceinline__ MomUni mom_uni(float ha, float hb, float eta2, int dim) {
  MomUni u;
  u.hab = 0.5f * (ha + hb);
  u.inv = frcp<float>(u.hab);
  u.fac = h_dim_inv<float>(u.inv, dim);
  u.eta2hh = eta2 * u.hab * u.hab;
  u.kout = -0.75f * u.inv * u.fac;
  u.kin = 3.0f * (u.inv * u.inv) * u.fac;
  return u;
}

// mom_pair_r<float, DM, DX> for a source of uniform h: the per-pair
// reciprocal of h_ab and the normalisation come from MomUni, and one
// reciprocal square root gives both |r_ab| and the 1/|r_ab| of the outer
// gradient branch (kernels.pyx:403-412), so a pair costs three
// quarter-rate transcendental instructions instead of five.
template <bool DM, bool DX>
__device__ __forceinline__ void mom_pair_u(const MomDst<float>& D, const MomUni& U,
                                           const float4& ps, const float4& vs,
                                           const float4& as, float* am, float* ax) {
  float rx, ry, rz;
  const float d2 = pair_d2<float>(D.pd, ps, rx, ry, rz);
  const float rs = __builtin_amdgcn_rsqf(d2);
  const float rr = d2 > 0.f ? d2 * rs : 0.f;  // the self pair: r = 0, not 0 * inf
  const float qq = rr * U.inv;
  const float mb = vs.w;
  const float rhob = as.x;
  const float irhoab = frcp<float>(0.5f * (D.rhoa + rhob));
  const float vx = D.vd.x - vs.x, vy = D.vd.y - vs.y, vz = D.vd.z - vs.z;
  const float Ab = as.z;
  const float cab = 0.5f * (D.ca + as.y);
  if (DM) {
    const float dot = vx * rx + vy * ry + vz * rz;
    const float mu = U.hab * dot * frcp<float>(d2 + U.eta2hh);
    const float piab = (-D.alpha * cab * mu + D.beta * mu * mu) * irhoab;
    const float tmp = D.Aa + Ab + (dot < 0.f ? piab : 0.f);
    const float t = 2.f - qq;
    const float outer = U.kout * (t * t) * rs;
    const float inner = U.kin * (0.75f * qq - 1.f);
    const float gfac = qq > 2.f ? 0.f : (qq >= 1.f ? outer : (qq > 1e-14f ? inner : 0.f));
    const float gf = -mb * tmp * gfac;
    am[0] += gf * rx;
    am[1] += gf * ry;
    am[2] += gf * rz;
  }
  if (DX) {
    const float temp = mb * (w_sel<float>(qq) * U.fac) * irhoab;
    ax[0] += temp * -vx;
    ax[1] += temp * -vy;
    ax[2] += temp * -vz;
  }
}

// The same pair with the roles chosen at run time (list passes).
template <typename CT>
__device__ __forceinline__ void mom_pair(const MomDst<CT>& D, const float4& ps, const float4& vs,
                                         const float4& as, bool do_m, bool do_x, CT* am,
                                         CT* ax) {
  if (do_m && do_x)
    mom_pair_r<CT, true, true>(D, ps, vs, as, am, ax);
  else if (do_m)
    mom_pair_r<CT, true, false>(D, ps, vs, as, am, ax);
  else if (do_x)
    mom_pair_r<CT, false, true>(D, ps, vs, as, am, ax);
}

template <typename CT, int DIM = 0>
__device__ __forceinline__ MomDst<CT> mom_dst(const PassArgs<float>& a, int i, bool valid) {
  const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
  MomDst<CT> D;
  D.pd = valid ? a.dst.pos4[i] : zero4;
  D.vd = valid ? a.dst.vel4[i] : zero4;
  const float4 ad = valid ? a.dst.aux4[i] : zero4;
  D.dim = DIM ? DIM : a.dim;
  D.ha = (CT)D.pd.w;
  D.rhoa = (CT)ad.x;
  D.ca = (CT)ad.y;
  D.Aa = sizeof(CT) == 8 ? (CT)ad.w / (D.rhoa * D.rhoa) : (CT)ad.z;
  D.alpha = (CT)a.alpha;
  D.beta = (CT)a.beta;
  D.eta2 = (CT)(a.eta * a.eta);
  return D;
}

// Outputs of the plain pass (MODE 0) or the integrator update of one
// particle in the reference's per-calc order mom, gravity, step, xsph
// (MODE 1 RK2 predictor, 2 RK2 final, 3 Euler); see momentum_kernel.
template <typename CT, int MODE>
__device__ __forceinline__ void mom_epilogue(const PassArgs<float>& a, int i, bool local,
                                             CT* acc, CT* dv, const float4& vd) {
  const CT f1 = (CT)a.fac1, fe = (CT)(a.fac1 * a.eps);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    acc[c] *= f1;
    dv[c] *= fe;
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
  double mv2 = 0.0;
  for (int c = 0; c < a.dim; ++c) {
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
    const double mv = (double)xn - (double)x0;  // exact for two floats this close
    mv2 += mv * mv;
  }
  if (a.disp_out != nullptr) {
    float disp = move_bound(mv2);
    for (int off = 32; off > 0; off >>= 1) disp = fmaxf(disp, __shfl_xor(disp, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax((int*)a.disp_out, __float_as_int(disp));
  }
}

// Cell scan with LDS queues (no list) for destination i; q as in
// density_scan_body.
template <typename CT, int MODE, int DIM = 0>
__device__ __forceinline__ void momentum_scan_body(const Grid& g, const PassArgs<float>& a,
                                                   int i, int* q) {
  const bool valid = i < a.dst.n;
  const bool local = valid && (a.dst.tag == nullptr || a.dst.tag[i] == 0);
  const MomDst<CT> D = mom_dst<CT, DIM>(a, i, valid);
  const double r = (double)D.pd.w * a.rs;
  const double r2d = r * r;
  const float r2f = (float)r2d;
  const float r2lo = r2f * 0.999998f, r2hi = r2f * 1.000002f;
  int cx = 0, cy = 0, cz = 0;
  if (local) cell_xyz(a.dst.cell[i], g, cx, cy, cz);
  const int ring = ring_of(g, a.disp);
  const Cull cu = make_cull(D.pd, r + (a.disp ? (double)*a.disp : 0.0), g);
  int dylo, dyhi, dzlo, dzhi;
  axis_span(cu.yd, cu.r, cy, g.dims[1], g.lo[1], g.inv_cs_f, ring, dylo, dyhi);
  axis_span(cu.zd, cu.r, cz, g.dims[2], g.lo[2], g.inv_cs_f, ring, dzlo, dzhi);
  CT acc[3] = {CT(0), CT(0), CT(0)};
  CT dv[3] = {CT(0), CT(0), CT(0)};
  for (int s = 0; s < a.nsrc; ++s) {
    const View<float> S = a.src[s];
    const __amdgpu_buffer_rsrc_t rs = cand_rsrc(S);
    const bool do_m = (a.flags[s] & SPH_SRC_MOMENTUM) != 0 && a.do_mom;
    const bool do_x = (a.flags[s] & SPH_SRC_XSPH) != 0 && a.do_xsph;
    if (!do_m && !do_x) continue;
    CT am[3] = {CT(0), CT(0), CT(0)};
    CT ax[3] = {CT(0), CT(0), CT(0)};
    int cnt = 0;
    // a source of uniform h (FP32 pairs): per-destination pair constants
    const bool uni = kMomUni && sizeof(CT) == 4 && S.hrange != nullptr &&
                     S.hrange[0] == S.hrange[1];
    // one drain loop per role set (the roles are uniform over a source)
    auto drain_r = [&](auto dm, auto dx) {
      constexpr bool M = decltype(dm)::value, X = decltype(dx)::value;
      if constexpr (sizeof(CT) == 4) {
        if (uni) {
          // formed per drain, so that they hold no registers while screening
          const MomUni U = mom_uni((float)D.ha, __int_as_float(S.hrange[0]), (float)D.eta2,
                                   D.dim);
          for (int k = 0; __any(k < cnt); ++k) {
            if (k < cnt) {
              const int j = q[k * 64];
              mom_pair_u<M, X>(reinterpret_cast<const MomDst<float>&>(D), U, ld4(S.pos4, j),
                               ld4(S.vel4, j), ld4(S.aux4, j), (float*)am, (float*)ax);
            }
          }
          cnt = 0;
          return;
        }
      }
      for (int k = 0; __any(k < cnt); ++k) {
        if (k < cnt) {
          const int j = q[k * 64];
          mom_pair_r<CT, M, X>(D, ld4(S.pos4, j), ld4(S.vel4, j), ld4(S.aux4, j), am, ax);
        }
      }
      cnt = 0;
    };
    auto drain = [&]() {
      if (do_m && do_x)
        drain_r(BoolC<true>{}, BoolC<true>{});
      else if (do_m)
        drain_r(BoolC<true>{}, BoolC<false>{});
      else
        drain_r(BoolC<false>{}, BoolC<true>{});
    };
    for (int dz = -ring; dz <= ring; ++dz) {
      const bool zok = local && dz >= dzlo && dz <= dzhi;
      if (!__any(zok)) continue;
      for (int dy = -ring; dy <= ring; ++dy) {
        const bool yok = zok && dy >= dylo && dy <= dyhi;
        if (!__any(yok)) continue;
        int j0 = 0, j1 = 0;
        if (yok) row_range_cull(g, S.cstart, cx, cy + dy, cz + dz, ring, cu, j0, j1);
        while (__any(j0 < j1)) {
          if (__any(cnt > kQueue - kUnroll)) drain();
          screen_block(S, rs, j0, j1, D.pd, r2lo, r2hi, r2d, false, 0.f, q, cnt);
          j0 += kUnroll;
        }
      }
    }
    drain();
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      acc[c] += am[c];
      dv[c] += ax[c];
    }
  }
  if (!valid) return;
  mom_epilogue<CT, MODE>(a, i, local, acc, dv, D.vd);
}

}  // namespace sph
