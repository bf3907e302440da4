// Batched multi-scene WCSPH: hundreds of independent scenes of uneven size
// stepped together with one launch per phase.
//
// Layout: all scenes' fluid particles form one concatenated set (two copies:
// the binning at the start of an RK2 step permutes fa into fb, the step
// evaluates both stages in that order and writes the final state back into
// fa), all boundary particles another.  Scene s owns the cells [cell_base_s,
// cell_base_s + cells_s) of one global cell space, so one counting sort bins
// every scene and keeps each scene contiguous and cell-ordered.
//
// Work: a tile is 64 consecutive particles of one scene, processed by one
// wave, so that the scene (its grid and scheme constants) is wave-uniform and
// lives in scalar registers.  The per-particle code is the single-array fast
// path's (density_scan_body, momentum_scan_body) with the scene's arguments,
// which keeps every scene's results bit-identical to stepping it alone.
// Schedules (a plain grid, or persistent waves walking per-XCD segments of
// the tile list statically or through atomic counters) only move tiles
// between waves.
#include <algorithm>
#include <vector>

#include <string.h>

#include "sph_fast.h"

namespace sph {

struct SceneArgs {
  Grid g;
  PassArgs<float> dens[2][2];  // [stage - 1][0 fluid dst, 1 boundary dst]
  PassArgs<float> stage[2];    // [stage - 1]
};

struct Tile {
  int scene, start, count, pad;
};

struct Batch {
  int nscenes = 0, dim = 2, cus = 256;
  int64_t nf = 0, nb = 0, ncells = 0;
  sph_batch_buffers buf{};
  SceneArgs* d_args = nullptr;
  BatchBinGrid* d_bgrid = nullptr;
  Tile* d_tiles[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // [set][order]
  int* d_seg[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};     // 9 segment starts
  int ntiles[2] = {0, 0};
  int* d_counters = nullptr;  // kXcds per launch site
  float* d_disp = nullptr;    // largest predictor displacement of the step
  // (x, y, z, m) copies of the fluid and boundary sets (written with pos4 /
  // vel4) and, per scene, the h-uniformity pair of each set ([scene][set]
  // -> 2 words: equal FP32 bits = uniform h), so that uniform-h scenes take
  // the same density and momentum pair paths as a single array does
  float* d_fpm4 = nullptr;
  float* d_bpm4 = nullptr;
  int* d_hrange = nullptr;
};

constexpr int kSites = 6;

__device__ __forceinline__ int site_of_density(int phase, int set) { return phase * 2 + set; }

// Calls f(t) for every tile t this wave owns under the schedule.  Every
// branch is wave-uniform; every wave leaves the loop once its XCD segment
// is exhausted.
template <typename F>
__device__ __forceinline__ void for_each_tile(int mode, int ntiles, const int* __restrict__ seg,
                                              int grab, int* counters, F&& f) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  constexpr int wpb = kBlock / 64;
  if (mode == 0) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * wpb + wid;
    if (t < ntiles) f(t);
    return;
  }
  const int xcd = blockIdx.x % kXcds;
  const int lb = blockIdx.x / kXcds;
  const int nbx = gridDim.x / kXcds;
  const int x0 = seg[xcd], x1 = seg[xcd + 1];
  if (mode == 1) {
    for (int t = x0 + lb * wpb + wid; t < x1; t += nbx * wpb) f(t);
    return;
  }
  for (;;) {
    int base = 0;
    if (lane == 0) base = atomicAdd(&counters[xcd], grab);
    base = x0 + __shfl(base, 0, 64);
    if (base >= x1) break;
    const int e = min(base + grab, x1);
    for (int t = base; t < e; ++t) f(t);
  }
}

template <int DIM>
__global__ __launch_bounds__(kBlock) void batch_density_kernel(
    const SceneArgs* __restrict__ args, int phase, int set, const Tile* __restrict__ tiles,
    int ntiles, const int* __restrict__ seg, int mode, int grab, int* counters) {
  __shared__ __align__(16) int qbuf[kBlock / 64][kDensQueue][64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int* q = &qbuf[wid][0][lane];
  for_each_tile(mode, ntiles, seg, grab, counters, [&](int t) {
    const Tile T = tiles[t];
    const int s = __builtin_amdgcn_readfirstlane(T.scene);
    const SceneArgs& A = args[s];
    density_scan_body<float, true, false, DIM>(A.g, A.dens[phase][set],
                                               __builtin_amdgcn_readfirstlane(T.start) + lane, q);
  });
}

template <int MODE, int DIM>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kMomWaves)))
void batch_stage_kernel(
    const SceneArgs* __restrict__ args, const Tile* __restrict__ tiles, int ntiles,
    const int* __restrict__ seg, int mode, int grab, int* counters) {
  __shared__ int qbuf[kBlock / 64][kQueue][64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int* q = &qbuf[wid][0][lane];
  for_each_tile(mode, ntiles, seg, grab, counters, [&](int t) {
    const Tile T = tiles[t];
    const int s = __builtin_amdgcn_readfirstlane(T.scene);
    const SceneArgs& A = args[s];
    momentum_scan_body<float, MODE, DIM>(A.g, A.stage[MODE - 1],
                                         __builtin_amdgcn_readfirstlane(T.start) + lane, q);
  });
}

// Three 32-bit arrays copied element-wise (dst[i] = src[i]).
__global__ __launch_bounds__(kBlock) void copy3_kernel(
    int64_t n, const uint32_t* __restrict__ a0, uint32_t* __restrict__ b0,
    const uint32_t* __restrict__ a1, uint32_t* __restrict__ b1, const uint32_t* __restrict__ a2,
    uint32_t* __restrict__ b2) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  b0[i] = a0[i];
  b1[i] = a1[i];
  b2[i] = a2[i];
}

// pos4 = (x, y, z, h), vel4 = (u, v, w, m), pm4 = (x, y, z, m) from separate
// component arrays.
__global__ __launch_bounds__(kBlock) void batch_pack_kernel(
    int64_t n, const float* __restrict__ x, const float* __restrict__ y,
    const float* __restrict__ z, const float* __restrict__ h, const float* __restrict__ u,
    const float* __restrict__ v, const float* __restrict__ w, const float* __restrict__ m,
    float4* __restrict__ pos4, float4* __restrict__ vel4, float4* __restrict__ pm4) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float xi = x[i], yi = y[i], zi = z[i], mi = m[i];
  pos4[i] = make_float4(xi, yi, zi, h[i]);
  vel4[i] = make_float4(u[i], v[i], w[i], mi);
  pm4[i] = make_float4(xi, yi, zi, mi);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static int64_t scene_cells(const sph_batch_scene& s) {
  return (int64_t)s.grid.dims[0] * s.grid.dims[1] * s.grid.dims[2];
}

static View<float> set_view(const sph_batch_set& S, int64_t off, int64_t n, const float* pos4,
                            const float* vel4, float* aux4, const int32_t* cell,
                            const int32_t* cstart, const float* pm4 = nullptr,
                            const int* hrange = nullptr) {
  sph_view v{};
  v.n = n;
  v.dtype = SPH_F32;
  v.x = S.x + off; v.y = S.y + off; v.z = S.z + off;
  v.u = S.u + off; v.v = S.v + off; v.w = S.w + off;
  v.h = S.h + off; v.m = S.m + off; v.rho = S.rho + off;
  v.p = S.p + off; v.cs = S.cs + off;
  v.tag = nullptr;
  v.cell = cell ? cell + off : nullptr;
  v.cell_start = cstart;
  v.pos4 = pos4 ? (void*)(pos4 + 4 * off) : nullptr;
  v.vel4 = vel4 ? (void*)(vel4 + 4 * off) : nullptr;
  v.aux4 = aux4 ? (void*)(aux4 + 4 * off) : nullptr;
  v.pm4 = pm4 ? (void*)(pm4 + 4 * off) : nullptr;
  v.hrange = hrange;
  return to_view<float>(v);
}

static int32_t check_set(const sph_batch_set& S, int64_t n, const char* name) {
  if (n == 0) return SPH_OK;
  if (!S.x || !S.y || !S.z || !S.u || !S.v || !S.w || !S.h || !S.m || !S.rho || !S.p || !S.cs ||
      !S.uid) {
    set_error("sph_batch_create: set %s has a NULL field", name);
    return SPH_EINVAL;
  }
  return SPH_OK;
}

template <typename T>
static int32_t upload(T** dst, const std::vector<T>& v) {
  const size_t bytes = std::max<size_t>(v.size(), 1) * sizeof(T);
  hipError_t e = hipMalloc((void**)dst, bytes);
  if (e == hipSuccess && !v.empty())
    e = hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    set_error("sph_batch_create: %s", hipGetErrorString(e));
    return (int32_t)e;
  }
  return SPH_OK;
}

static void free_batch(Batch* b) {
  if (!b) return;
  (void)hipFree(b->d_args);
  (void)hipFree(b->d_bgrid);
  for (int s = 0; s < 2; ++s)
    for (int o = 0; o < 2; ++o) {
      (void)hipFree(b->d_tiles[s][o]);
      (void)hipFree(b->d_seg[s][o]);
    }
  (void)hipFree(b->d_counters);
  (void)hipFree(b->d_disp);
  (void)hipFree(b->d_fpm4);
  (void)hipFree(b->d_bpm4);
  (void)hipFree(b->d_hrange);
  delete b;
}

// Tile lists of one destination set in both orders, with the 9 XCD segment
// starts of each.
static void make_tiles(const std::vector<sph_batch_scene>& sc, int set,
                       std::vector<Tile> out[2], std::vector<int> seg[2]) {
  std::vector<Tile> nat;
  std::vector<double> cost;
  for (int s = 0; s < (int)sc.size(); ++s) {
    const int64_t n = set == 0 ? sc[s].nf : sc[s].nb;
    for (int64_t k = 0; k < n; k += 64) {
      Tile t{s, (int)k, (int)std::min<int64_t>(64, n - k), 0};
      nat.push_back(t);
      cost.push_back(std::max(sc[s].cost, 1e-12) * t.count);
    }
  }
  const int nt = (int)nat.size();
  // order 0: scene order, equal tile counts per XCD
  out[0] = nat;
  seg[0].resize(kXcds + 1);
  for (int x = 0; x <= kXcds; ++x) seg[0][x] = (int)((int64_t)nt * x / kXcds);
  // order 1: equal estimated cost per XCD, costliest tiles first within each
  std::vector<double> pre(nt + 1, 0.0);
  for (int t = 0; t < nt; ++t) pre[t + 1] = pre[t] + cost[t];
  seg[1].resize(kXcds + 1);
  seg[1][0] = 0;
  for (int x = 1; x < kXcds; ++x) {
    const double target = pre[nt] * x / kXcds;
    int t = (int)(std::lower_bound(pre.begin(), pre.end(), target) - pre.begin());
    seg[1][x] = std::max(seg[1][x - 1], std::min(t, nt));
  }
  seg[1][kXcds] = nt;
  out[1] = nat;
  for (int x = 0; x < kXcds; ++x) {
    std::vector<int> idx;
    for (int t = seg[1][x]; t < seg[1][x + 1]; ++t) idx.push_back(t);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    for (size_t k = 0; k < idx.size(); ++k) out[1][seg[1][x] + k] = nat[idx[k]];
  }
}

static Batch* as_batch(void* h) { return reinterpret_cast<Batch*>(h); }

static int32_t launch_blocks(const Batch* b, const sph_batch_sched* s, int ntiles) {
  if (s->mode == 0) return std::max(1, (ntiles + 3) / 4);
  const int per = std::max(1, std::min(16, (int)s->blocks_per_cu));
  return std::max(kXcds, (b->cus * per / kXcds) * kXcds);
}

static int32_t check_sched(const sph_batch_sched* s, const char* who) {
  if (s == nullptr || s->mode < 0 || s->mode > 2 || s->order < 0 || s->order > 1 ||
      (s->mode == 2 && s->grab < 1) || (s->mode != 0 && s->blocks_per_cu < 1)) {
    set_error("%s: invalid schedule", who);
    return SPH_EINVAL;
  }
  return SPH_OK;
}

}  // namespace sph

using namespace sph;

extern "C" {

int64_t sph_batch_ncells(int32_t nscenes, const sph_batch_scene* scenes) {
  int64_t c = 0;
  for (int s = 0; s < nscenes; ++s) c += scene_cells(scenes[s]);
  return c;
}

int32_t sph_batch_create(int32_t nscenes, const sph_batch_scene* scenes,
                         const sph_batch_buffers* buffers, void** handle) {
  if (nscenes <= 0 || !scenes || !buffers || !handle) {
    set_error("sph_batch_create: bad arguments");
    return SPH_EINVAL;
  }
  *handle = nullptr;
  const sph_batch_buffers& B = *buffers;
  std::vector<sph_batch_scene> sc(scenes, scenes + nscenes);
  const int dim = sc[0].prm.dim;
  int64_t ncells = 0, f_end = 0, b_end = 0;
  for (int s = 0; s < nscenes; ++s) {
    const sph_batch_scene& S = sc[s];
    SPH_TRY(check_grid(&S.grid, "sph_batch_create"));
    if (S.prm.dim != dim || dim < 2 || dim > 3) {
      set_error("sph_batch_create: scene %d: all scenes must share dim 2 or 3", s);
      return SPH_EUNSUPPORTED;
    }
    if (S.f0 != f_end || S.nf < 0 || S.b0 != b_end || S.nb < 0) {
      set_error("sph_batch_create: scene %d: slices must tile the sets in scene order", s);
      return SPH_EINVAL;
    }
    if (!(S.rmax > 0) || !(S.prm.radius_scale > 0) || !(S.dt > 0)) {
      set_error("sph_batch_create: scene %d: rmax, radius_scale and dt must be positive", s);
      return SPH_EINVAL;
    }
    if (S.grid.period[0] > 0 || S.grid.period[1] > 0 || S.grid.period[2] > 0) {
      set_error("sph_batch_create: scene %d: periodic grids are not supported", s);
      return SPH_EUNSUPPORTED;
    }
    if (S.prm.symmetric) {
      set_error("sph_batch_create: scene %d: variable-h symmetric pairs are not supported", s);
      return SPH_EUNSUPPORTED;
    }
    if (S.prm.exclude_self) {
      set_error("sph_batch_create: scene %d: exclude_self is not supported", s);
      return SPH_EUNSUPPORTED;
    }
    f_end += S.nf;
    b_end += S.nb;
    ncells += scene_cells(S);
  }
  if (f_end != B.nf_total || b_end != B.nb_total) {
    set_error("sph_batch_create: slices cover %lld fluid / %lld boundary particles, buffers "
              "hold %lld / %lld", (long long)f_end, (long long)b_end, (long long)B.nf_total,
              (long long)B.nb_total);
    return SPH_EINVAL;
  }
  if (ncells > (int64_t)INT32_MAX - 2 || B.nf_total > INT32_MAX || B.nb_total > INT32_MAX) {
    set_error("sph_batch_create: more than 2^31 cells or particles");
    return SPH_EINVAL;
  }
  SPH_TRY(check_set(B.fa, B.nf_total, "fa"));
  SPH_TRY(check_set(B.fb, B.nf_total, "fb"));
  SPH_TRY(check_set(B.bnd, B.nb_total, "bnd"));
  SPH_TRY(check_set(B.bnd_in, B.nb_total, "bnd_in"));
  for (int c = 0; c < dim; ++c)
    for (float* const* arr : {B.x1a, B.x1b}) {
      if (B.nf_total > 0 && (!arr[c] || !arr[3 + c])) {
        set_error("sph_batch_create: X1 component %d is NULL", c);
        return SPH_EINVAL;
      }
    }
  for (int c = 0; c < dim; ++c)
    for (float* const* arr : {B.k1a, B.k1b}) {
      if (B.nf_total > 0 && (!arr[c] || !arr[3 + c])) {
        set_error("sph_batch_create: k1 component %d is NULL", c);
        return SPH_EINVAL;
      }
    }
  if ((B.nf_total > 0 && (!B.fpos4 || !B.fvel4 || !B.faux4 || !B.fscene || !B.fperm ||
                          !B.fcell || !B.fcell_start)) ||
      (B.nb_total > 0 && (!B.bpos4 || !B.bvel4 || !B.baux4 || !B.bscene || !B.bperm ||
                          !B.bcell || !B.bcell_start)) ||
      !B.ws || B.ws_bytes < sph_bin_workspace_bytes(std::max(B.nf_total, B.nb_total), ncells)) {
    set_error("sph_batch_create: missing packed, binning or workspace buffers");
    return SPH_EINVAL;
  }
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);

  Batch* b = new Batch();
  b->nscenes = nscenes;
  b->dim = dim;
  b->cus = cus > 0 ? cus : 256;
  b->nf = B.nf_total;
  b->nb = B.nb_total;
  b->ncells = ncells;
  b->buf = B;

  if (upload(&b->d_disp, std::vector<float>(1, 0.f)) != SPH_OK) {
    free_batch(b);
    return SPH_EINVAL;
  }
  {
    // h never changes during a run: the h-uniformity pairs are fixed here
    std::vector<float> hf((size_t)B.nf_total), hb((size_t)B.nb_total);
    hipError_t e = hipSuccess;
    if (B.nf_total > 0)
      e = hipMemcpy(hf.data(), B.fa.h, hf.size() * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && B.nb_total > 0)
      e = hipMemcpy(hb.data(), B.bnd_in.h, hb.size() * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      set_error("sph_batch_create: %s", hipGetErrorString(e));
      free_batch(b);
      return (int32_t)e;
    }
    std::vector<int> hr((size_t)nscenes * 4, 0);
    auto pair = [](const float* h, int64_t n, int* out) {
      if (n <= 0) return;
      int b0, b1;
      memcpy(&b0, &h[0], 4);
      b1 = b0;
      for (int64_t i = 1; i < n; ++i) {
        int bi;
        memcpy(&bi, &h[i], 4);
        if (bi != b0) {
          b1 = bi;
          break;
        }
      }
      out[0] = b0;
      out[1] = b1;
    };
    for (int s = 0; s < nscenes; ++s) {
      pair(hf.data() + sc[s].f0, sc[s].nf, &hr[4 * s]);
      pair(hb.data() + sc[s].b0, sc[s].nb, &hr[4 * s + 2]);
    }
    if (upload(&b->d_hrange, hr) != SPH_OK ||
        upload(&b->d_fpm4, std::vector<float>((size_t)std::max<int64_t>(B.nf_total, 1) * 4)) !=
            SPH_OK ||
        upload(&b->d_bpm4, std::vector<float>((size_t)std::max<int64_t>(B.nb_total, 1) * 4)) !=
            SPH_OK) {
      free_batch(b);
      return SPH_EINVAL;
    }
  }
  std::vector<SceneArgs> args(nscenes);
  std::vector<BatchBinGrid> bgrid(nscenes);
  int64_t base = 0;
  for (int s = 0; s < nscenes; ++s) {
    const sph_batch_scene& S = sc[s];
    BatchBinGrid& bg = bgrid[s];
    bg.cs = S.grid.cell_size;
    for (int d = 0; d < 3; ++d) {
      bg.lo[d] = S.grid.lo[d];
      bg.dims[d] = S.grid.dims[d];
    }
    bg.cell_base = (int)base;
    SceneArgs& A = args[s];
    A.g = make_grid(S.grid, S.rmax);
    A.g.cell_base = (int)base;
    const int32_t* fcs = B.fcell_start + base;
    const int32_t* bcs = S.nb > 0 ? B.bcell_start + base : nullptr;
    base += scene_cells(S);
    // both stages evaluate the fluid in the order of the step's binning
    // (fb); stage 2 reads the predictor X1 through the packed copies and
    // widens its stencil by the predictor's displacement, and the step's
    // results (density of X1, final state) go to fa
    const sph_batch_set* fset[2] = {&B.fb, &B.fb};
    for (int ph = 0; ph < 2; ++ph) {
      const View<float> fsrc = set_view(*fset[ph], 0, B.nf_total, B.fpos4, B.fvel4, B.faux4,
                                        B.fcell, fcs, b->d_fpm4, b->d_hrange + 4 * s);
      const View<float> bsrc = set_view(B.bnd, 0, B.nb_total, B.bpos4, B.bvel4, B.baux4,
                                        B.bcell, bcs, b->d_bpm4, b->d_hrange + 4 * s + 2);
      for (int set = 0; set < 2; ++set) {
        PassArgs<float>& a = A.dens[ph][set];
        a = PassArgs<float>{};
        set_params(a, &S.prm);
        if (set == 0)
          a.dst = set_view(*fset[ph], S.f0, S.nf, B.fpos4, B.fvel4, B.faux4, B.fcell, fcs);
        else
          a.dst = set_view(B.bnd, S.b0, S.nb, B.bpos4, B.bvel4, B.baux4, B.bcell, bcs);
        a.nsrc = S.nb > 0 ? 2 : 1;
        a.src[0] = fsrc;
        a.flags[0] = SPH_SRC_MOMENTUM | SPH_SRC_XSPH;
        if (S.nb > 0) {
          a.src[1] = bsrc;
          a.flags[1] = SPH_SRC_MOMENTUM | SPH_SRC_XSPH;
        }
        const sph_batch_set& D = set == 0 ? (ph == 0 ? B.fb : B.fa) : B.bnd;
        const int64_t off = set == 0 ? S.f0 : S.b0;
        a.rho = D.rho + off;
        a.p = D.p + off;
        a.cs = D.cs + off;
        if (ph == 1) a.disp = b->d_disp;
      }
      // stage ph+1: momentum of the fluid
      PassArgs<float>& m = A.stage[ph];
      m = PassArgs<float>{};
      set_params(m, &S.prm);
      m.dst = set_view(*fset[ph], S.f0, S.nf, B.fpos4, B.fvel4, B.faux4, B.fcell, fcs);
      m.nsrc = S.nb > 0 ? 2 : 1;
      m.src[0] = fsrc;
      m.flags[0] = SPH_SRC_MOMENTUM | (S.prm.has_xsph ? SPH_SRC_XSPH : 0u);
      if (S.nb > 0) {
        m.src[1] = bsrc;
        m.flags[1] = SPH_SRC_MOMENTUM;
      }
      m.do_mom = 1;
      m.do_xsph = S.prm.has_xsph ? 1 : 0;
      m.dt = S.dt;
      const sph_batch_set& X = B.fb;
      const float* xs[6] = {X.x, X.y, X.z, X.u, X.v, X.w};
      float* fin[6] = {B.fa.x, B.fa.y, B.fa.z, B.fa.u, B.fa.v, B.fa.w};
      for (int c = 0; c < 6; ++c) {
        const bool need = (c % 3) < dim;
        m.s0[c] = need ? xs[c] + S.f0 : nullptr;
        m.k1[c] = need ? B.k1a[c] + S.f0 : nullptr;
        m.out[c] = need ? (ph == 0 ? B.x1a[c] : fin[c]) + S.f0 : nullptr;
      }
      if (ph == 0)
        m.disp_out = b->d_disp;  // stage 1 records the predictor's displacement
      else
        m.disp = b->d_disp;      // stage 2 widens its stencil by it
    }
  }
  int32_t rc = upload(&b->d_args, args);
  if (rc == SPH_OK) rc = upload(&b->d_bgrid, bgrid);
  for (int set = 0; set < 2 && rc == SPH_OK; ++set) {
    std::vector<Tile> tl[2];
    std::vector<int> sg[2];
    make_tiles(sc, set, tl, sg);
    b->ntiles[set] = (int)tl[0].size();
    for (int o = 0; o < 2 && rc == SPH_OK; ++o) {
      rc = upload(&b->d_tiles[set][o], tl[o]);
      if (rc == SPH_OK) rc = upload(&b->d_seg[set][o], sg[o]);
    }
  }
  if (rc == SPH_OK) rc = upload(&b->d_counters, std::vector<int>(kSites * kXcds, 0));
  if (rc != SPH_OK) {
    free_batch(b);
    return rc;
  }
  *handle = b;
  return SPH_OK;
}

int32_t sph_batch_destroy(void* handle) {
  free_batch(as_batch(handle));
  return SPH_OK;
}

int32_t sph_batch_bin(void* handle, void* stream, int32_t which) {
  Batch* b = as_batch(handle);
  if (!b || which < 0 || which > 2) {
    set_error("sph_batch_bin: bad arguments");
    return SPH_EINVAL;
  }
  hipStream_t st = as_stream(stream);
  const sph_batch_buffers& B = b->buf;
  const int dim = b->dim;
  if (which == 2) {
    if (b->nb == 0) return SPH_OK;
    const sph_batch_set& S = B.bnd_in;
    SPH_TRY(batch_bin(st, b->nb, b->ncells, S.x, S.y, S.z, B.bscene, b->d_bgrid, B.bperm,
                      B.bcell, B.bcell_start, B.ws));
    const void* src[12] = {S.x, S.y, S.z, S.u, S.v, S.w, S.h, S.m, S.rho, S.p, S.cs, S.uid};
    const sph_batch_set& D = B.bnd;
    void* dst[12] = {D.x, D.y, D.z, D.u, D.v, D.w, D.h, D.m, D.rho, D.p, D.cs, D.uid};
    const int32_t esz[12] = {4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4};
    return permute_fields(st, b->nb, B.bperm, 12, src, dst, esz);
  }
  if (b->nf == 0) return SPH_OK;
  const sph_batch_set& S = which == 0 ? B.fa : B.fb;
  const sph_batch_set& D = which == 0 ? B.fb : B.fa;
  const float* px = S.x;
  const float* py = S.y;
  const float* pz = S.z;
  if (which == 1) {  // bin by X1; components >= dim are those of X
    px = B.x1a[0];
    py = B.x1a[1];
    if (dim > 2) pz = B.x1a[2];
  }
  SPH_TRY(batch_bin(st, b->nf, b->ncells, px, py, pz, B.fscene, b->d_bgrid, B.fperm, B.fcell,
                    B.fcell_start, B.ws));
  std::vector<const void*> src = {S.x, S.y, S.z, S.u, S.v, S.w, S.h, S.m, S.rho, S.p, S.cs,
                                  S.uid};
  std::vector<void*> dst = {D.x, D.y, D.z, D.u, D.v, D.w, D.h, D.m, D.rho, D.p, D.cs, D.uid};
  if (which == 1) {
    for (int c = 0; c < 6; ++c) {
      if ((c % 3) >= dim) continue;
      src.push_back(B.x1a[c]);
      dst.push_back(B.x1b[c]);
      src.push_back(B.k1a[c]);
      dst.push_back(B.k1b[c]);
    }
  }
  std::vector<int32_t> esz(src.size(), 4);
  return permute_fields(st, b->nf, B.fperm, (int32_t)src.size(), src.data(), dst.data(),
                        esz.data());
}

int32_t sph_batch_pack(void* handle, void* stream, int32_t which) {
  Batch* b = as_batch(handle);
  if (!b || which < 0 || which > 2) {
    set_error("sph_batch_pack: bad arguments");
    return SPH_EINVAL;
  }
  hipStream_t st = as_stream(stream);
  const sph_batch_buffers& B = b->buf;
  const int64_t n = which == 2 ? b->nb : b->nf;
  if (n == 0) return SPH_OK;
  const sph_batch_set& S = which == 2 ? B.bnd : B.fb;
  const float *x = S.x, *y = S.y, *z = S.z, *u = S.u, *v = S.v, *w = S.w;
  if (which == 1) {  // X1 (in fb's order) with components >= dim from X
    const int dim = b->dim;
    x = B.x1a[0];
    y = B.x1a[1];
    u = B.x1a[3];
    v = B.x1a[4];
    if (dim > 2) {
      z = B.x1a[2];
      w = B.x1a[5];
    }
  }
  float* pos4 = which == 2 ? B.bpos4 : B.fpos4;
  float* vel4 = which == 2 ? B.bvel4 : B.fvel4;
  float* pm4 = which == 2 ? b->d_bpm4 : b->d_fpm4;
  hipLaunchKernelGGL(batch_pack_kernel, dim3(nblocks_for(n)), dim3(kBlock), 0, st, n, x, y, z,
                     S.h, u, v, w, S.m, (float4*)pos4, (float4*)vel4, (float4*)pm4);
  return launch_status("sph_batch_pack");
}

int32_t sph_batch_density(void* handle, void* stream, int32_t phase, int32_t set,
                          const sph_batch_sched* sched) {
  Batch* b = as_batch(handle);
  if (!b || phase < 0 || phase > 1 || set < 0 || set > 1) {
    set_error("sph_batch_density: bad arguments");
    return SPH_EINVAL;
  }
  SPH_TRY(check_sched(sched, "sph_batch_density"));
  const int nt = b->ntiles[set];
  if (nt == 0) return SPH_OK;
  hipStream_t st = as_stream(stream);
  int* ctr = b->d_counters + (phase * 2 + set) * kXcds;
  if (sched->mode == 2) SPH_TRY(fill_i32(st, ctr, kXcds, 0));
  if (b->dim == 3)
    hipLaunchKernelGGL(batch_density_kernel<3>, dim3(launch_blocks(b, sched, nt)), dim3(kBlock),
                       0, st, b->d_args, phase, set, b->d_tiles[set][sched->order], nt,
                       b->d_seg[set][sched->order], sched->mode, sched->grab, ctr);
  else
    hipLaunchKernelGGL(batch_density_kernel<2>, dim3(launch_blocks(b, sched, nt)), dim3(kBlock),
                       0, st, b->d_args, phase, set, b->d_tiles[set][sched->order], nt,
                       b->d_seg[set][sched->order], sched->mode, sched->grab, ctr);
  return launch_status("sph_batch_density");
}

int32_t sph_batch_stage(void* handle, void* stream, int32_t stage, const sph_batch_sched* sched) {
  Batch* b = as_batch(handle);
  if (!b || stage < 1 || stage > 2) {
    set_error("sph_batch_stage: bad arguments");
    return SPH_EINVAL;
  }
  SPH_TRY(check_sched(sched, "sph_batch_stage"));
  const int nt = b->ntiles[0];
  if (nt == 0) return SPH_OK;
  hipStream_t st = as_stream(stream);
  int* ctr = b->d_counters + (4 + stage - 1) * kXcds;
  if (sched->mode == 2) SPH_TRY(fill_i32(st, ctr, kXcds, 0));
  const int nbk = launch_blocks(b, sched, nt);
  const Tile* tl = b->d_tiles[0][sched->order];
  const int* sg = b->d_seg[0][sched->order];
  if (stage == 1) SPH_TRY(fill_i32(st, (int32_t*)b->d_disp, 1, 0));
#define SPH_STAGE(M, D)                                                                   \
  hipLaunchKernelGGL((batch_stage_kernel<M, D>), dim3(nbk), dim3(kBlock), 0, st, b->d_args, tl, \
                     nt, sg, sched->mode, sched->grab, ctr)
  if (stage == 1) {
    if (b->dim == 3) SPH_STAGE(1, 3);
    else SPH_STAGE(1, 2);
  } else {
    if (b->dim == 3) SPH_STAGE(2, 3);
    else SPH_STAGE(2, 2);
  }
#undef SPH_STAGE
  if (stage == 2) {
    // the step's final state in fa also needs the fields the stages do not
    // write (h, m, uid) in the same order
    const sph_batch_buffers& B = b->buf;
    hipLaunchKernelGGL(copy3_kernel, dim3(nblocks_for(b->nf)), dim3(kBlock), 0, st, b->nf,
                       (const uint32_t*)B.fb.h, (uint32_t*)B.fa.h, (const uint32_t*)B.fb.m,
                       (uint32_t*)B.fa.m, (const uint32_t*)B.fb.uid, (uint32_t*)B.fa.uid);
  }
  return launch_status("sph_batch_stage");
}

// ---- graphs ---------------------------------------------------------------
int32_t sph_graph_begin(void* stream) {
  hipError_t e = hipStreamBeginCapture(as_stream(stream), hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    set_error("sph_graph_begin: %s", hipGetErrorString(e));
    return (int32_t)e;
  }
  return SPH_OK;
}

int32_t sph_graph_end(void* stream, void** graph_exec) {
  if (!graph_exec) {
    set_error("sph_graph_end: NULL output");
    return SPH_EINVAL;
  }
  *graph_exec = nullptr;
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(as_stream(stream), &g);
  if (e != hipSuccess) {
    set_error("sph_graph_end: %s", hipGetErrorString(e));
    return (int32_t)e;
  }
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    set_error("sph_graph_end: instantiate: %s", hipGetErrorString(e));
    return (int32_t)e;
  }
  *graph_exec = x;
  return SPH_OK;
}

int32_t sph_graph_launch(void* graph_exec, void* stream) {
  if (!graph_exec) {
    set_error("sph_graph_launch: NULL graph");
    return SPH_EINVAL;
  }
  hipError_t e = hipGraphLaunch((hipGraphExec_t)graph_exec, as_stream(stream));
  if (e != hipSuccess) {
    set_error("sph_graph_launch: %s", hipGetErrorString(e));
    return (int32_t)e;
  }
  return SPH_OK;
}

int32_t sph_graph_destroy(void* graph_exec) {
  if (graph_exec) (void)hipGraphExecDestroy((hipGraphExec_t)graph_exec);
  return SPH_OK;
}

}  // extern "C"
