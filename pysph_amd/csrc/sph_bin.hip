// Cell binning (stable counting sort by cell key), the permutation gather,
// bounds, scans and the exact neighbour-list export.
//
// Reference behaviour being replaced: CellManager.update / Cell.update bin
// particles into a dict of cells keyed by floor(x/cell_size)
// (cell.pyx:42-82, :319-432, :927-1015), and FixedDestNbrParticleLocator
// caches, per destination, the source indices passing the strict FP64 test
// (nnps.pyx:271-356, :525-620).
#include <math.h>

#include <mutex>

#include "sph_internal.h"

namespace sph {

// ---------------------------------------------------------------------------
// block / wave scan helpers (64-wide wavefronts)
// ---------------------------------------------------------------------------
template <typename V>
__device__ __forceinline__ V wave_inclusive_scan(V x) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    V y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  return x;
}

// Exclusive scan over a block of kBlock threads; returns the thread's
// exclusive prefix and the block total.
template <typename V>
__device__ __forceinline__ V block_exclusive_scan(V x, V& total) {
  __shared__ V wsum[kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  V inc = wave_inclusive_scan(x);
  if (lane == 63) wsum[wid] = inc;
  __syncthreads();
  V off = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < kBlock / 64; ++k) {
    V s = wsum[k];
    if (k < wid) off += s;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return off + inc - x;
}

constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

template <typename In, typename Out>
__global__ __launch_bounds__(kBlock) void scan_reduce(const In* __restrict__ in, int64_t n,
                                                      Out* __restrict__ partial) {
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  Out s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = base + (int64_t)k * kBlock + threadIdx.x;
    if (i < n) s += (Out)in[i];
  }
  Out tot;
  block_exclusive_scan<Out>(s, tot);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// One block scans the partials in place (exclusive) and writes the grand
// total to *total.
template <typename Out>
__global__ __launch_bounds__(kBlock) void scan_partials(Out* __restrict__ partial, int nb,
                                                        Out* __restrict__ total) {
  Out carry = 0;
  for (int base = 0; base < nb; base += kBlock) {
    int i = base + threadIdx.x;
    Out x = i < nb ? partial[i] : 0;
    Out tot;
    Out ex = block_exclusive_scan<Out>(x, tot);
    if (i < nb) partial[i] = ex + carry;
    carry += tot;
  }
  if (threadIdx.x == 0) *total = carry;
}

template <typename In, typename Out>
__global__ __launch_bounds__(kBlock) void scan_final(const In* __restrict__ in, int64_t n,
                                                     const Out* __restrict__ partial,
                                                     Out* __restrict__ out) {
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  Out v[kScanItems];
  Out s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = base + k;
    v[k] = i < n ? (Out)in[i] : 0;
    s += v[k];
  }
  Out tot;
  Out ex = block_exclusive_scan<Out>(s, tot) + partial[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = base + k;
    if (i < n) out[i] = ex;
    ex += v[k];
  }
}

// Exclusive scan of n counts into out[0..n]; out[n] = total.  `partial`
// holds at least ceil(n / kScanTile) elements.
template <typename In, typename Out>
static int32_t exclusive_scan(hipStream_t st, const In* in, int64_t n, Out* out,
                              Out* partial) {
  const int nb = (int)((n + kScanTile - 1) / kScanTile);
  if (nb > 0) {
    hipLaunchKernelGGL((scan_reduce<In, Out>), dim3(nb), dim3(kBlock), 0, st, in, n,
                       partial);
  }
  hipLaunchKernelGGL((scan_partials<Out>), dim3(1), dim3(kBlock), 0, st, partial, nb,
                     out + n);
  if (nb > 0) {
    hipLaunchKernelGGL((scan_final<In, Out>), dim3(nb), dim3(kBlock), 0, st, in, n,
                       partial, out);
  }
  return launch_status("exclusive_scan");
}

__global__ void fill_i32_kernel(int32_t* __restrict__ p, int64_t n, int32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

int32_t fill_i32(hipStream_t st, int32_t* p, int64_t n, int32_t v) {
  if (n <= 0) return SPH_OK;
  int nb = (int)((n + kBlock - 1) / kBlock);
  if (nb > 65536) nb = 65536;
  hipLaunchKernelGGL(fill_i32_kernel, dim3(nb), dim3(kBlock), 0, st, p, n, v);
  return launch_status("fill_i32");
}

// ---------------------------------------------------------------------------
// bounds
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void bounds_partial(View<T> a, double* __restrict__ part) {
  double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[4] = {-INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < a.n;
       i += (int64_t)gridDim.x * kBlock) {
    double c[3] = {(double)a.x[i], (double)a.y[i], (double)a.z[i]};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      mn[d] = fmin(mn[d], c[d]);
      mx[d] = fmax(mx[d], c[d]);
    }
    mx[3] = fmax(mx[3], (double)a.h[i]);
  }
  __shared__ double red[7][kBlock];
  for (int d = 0; d < 3; ++d) {
    red[d][threadIdx.x] = mn[d];
    red[3 + d][threadIdx.x] = mx[d];
  }
  red[6][threadIdx.x] = mx[3];
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      for (int d = 0; d < 3; ++d)
        red[d][threadIdx.x] = fmin(red[d][threadIdx.x], red[d][threadIdx.x + s]);
      for (int d = 3; d < 7; ++d)
        red[d][threadIdx.x] = fmax(red[d][threadIdx.x], red[d][threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 7) part[blockIdx.x * 7 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void bounds_final(const double* __restrict__ part, int nb, double* __restrict__ out) {
  const int d = threadIdx.x;
  if (d >= 7) return;
  double v = d < 3 ? INFINITY : -INFINITY;
  for (int b = 0; b < nb; ++b) v = d < 3 ? fmin(v, part[b * 7 + d]) : fmax(v, part[b * 7 + d]);
  out[d] = v;
}

// ---------------------------------------------------------------------------
// binning
// ---------------------------------------------------------------------------
struct BinGrid {
  double cs;
  int lo[3];
  int dims[3];
  int k;  // bins per reference cell along each axis (the jump check's unit)
};

// floor(a / k) for k >= 1 and any sign of a.
__device__ __forceinline__ int floor_div(int a, int k) {
  return a >= 0 ? a / k : -((-a + k - 1) / k);
}

// floor(x/cs) as the reference computes it (cell.pyx:57), shifted into the
// grid and clamped.
__device__ __forceinline__ int grid_coord(double x, double cs, int lo, int dim) {
  const double f = floor(x / cs);
  const double c = f - (double)lo;
  if (!(c >= 0.0)) return 0;  // also catches NaN
  if (c > (double)(dim - 1)) return dim - 1;
  return (int)c;
}

// Runs of one cell among consecutive lanes of a wave.  Particles arrive in
// the slot order of the previous binning, i.e. mostly cell-sorted, so a
// run's first lane can do the run's atomic for all of it.  Valid lanes are
// a prefix of the wave (only the last wave of a launch has invalid lanes).
struct WaveRun {
  int head;      // lane of this lane's run head
  int len;       // run length (heads only)
  bool is_head;
};

__device__ __forceinline__ WaveRun wave_run(int key, bool valid) {
  const int lane = threadIdx.x & 63;
  const int prev = __shfl_up(key, 1, 64);
  WaveRun r;
  r.is_head = valid && (lane == 0 || prev != key);
  const unsigned long long heads = __ballot(r.is_head);
  const unsigned long long vmask = __ballot(valid);
  const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);
  const unsigned long long mine = heads & upto;
  r.head = mine ? 63 - __clzll((long long)mine) : lane;
  const unsigned long long after = heads & ~upto;
  const int end = after ? __ffsll((long long)after) - 1 : (vmask ? 64 - __clzll((long long)vmask) : 0);
  r.len = end - lane;
  return r;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void bin_count(View<T> a, BinGrid g,
                                                    int* __restrict__ cell_tmp,
                                                    int* __restrict__ counts,
                                                    const int* __restrict__ prev_cell,
                                                    int* __restrict__ jumps) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const bool valid = i < a.n;
  int c = -1, cx = 0, cy = 0, cz = 0;
  if (valid) {
    cx = grid_coord((double)a.x[i], g.cs, g.lo[0], g.dims[0]);
    cy = grid_coord((double)a.y[i], g.cs, g.lo[1], g.dims[1]);
    cz = grid_coord((double)a.z[i], g.cs, g.lo[2], g.dims[2]);
    c = (cz * g.dims[1] + cy) * g.dims[0] + cx;
    cell_tmp[i] = c;
  }
  const WaveRun r = wave_run(c, valid);
  if (r.is_head) atomicAdd(&counts[c], r.len);
  if (valid && prev_cell != nullptr) {
    const int pc = prev_cell[i];
    int px = pc % g.dims[0], t = pc / g.dims[0];
    int py = t % g.dims[1], pz = t / g.dims[1];
    int qx = cx, qy = cy, qz = cz;
    if (g.k > 1) {
      // compare reference cells, floor(x / cell_size) = floor(bin / k)
      px = floor_div(px + g.lo[0], g.k);
      py = floor_div(py + g.lo[1], g.k);
      pz = floor_div(pz + g.lo[2], g.k);
      qx = floor_div(qx + g.lo[0], g.k);
      qy = floor_div(qy + g.lo[1], g.k);
      qz = floor_div(qz + g.lo[2], g.k);
    }
    if (abs(px - qx) > 1 || abs(py - qy) > 1 || abs(pz - qz) > 1) atomicAdd(jumps, 1);
  }
}

// The jump check of CellManager.update (cell.pyx:1422-1437) on the
// reference's unbounded cells: the cell of a particle is floor(x/cs) per
// axis (cell.pyx:57, FP64), kept per particle between checks as one int64
// (21 bits per axis, biased by 2^20, coordinates clamped to +-2^20; a
// non-finite position has no cell).  Counting on these rather than on the
// grid's bins keeps the check independent of the binning grid: of its bin
// size, of grid changes, and of the clamping of particles outside it into
// its border bins.
constexpr int64_t kRefBias = (int64_t)1 << 20;
constexpr int64_t kRefNone = INT64_MIN;

__device__ __forceinline__ int64_t ref_coord(double x, double cs) {
  const double f = floor(x / cs);
  const double c = fmin(fmax(f, (double)-kRefBias), (double)(kRefBias - 1));
  return (int64_t)c + kRefBias;
}

__device__ __forceinline__ bool ref_moved(int64_t a, int64_t b) {
  const int64_t m = (int64_t)(1 << 21) - 1;
  bool jump = false;
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    const int64_t d = ((a >> (21 * s)) & m) - ((b >> (21 * s)) & m);
    jump = jump || d > 1 || d < -1;
  }
  return jump;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void jump_check_kernel(View<T> a, double cs,
                                                            int64_t* __restrict__ ref,
                                                            int* __restrict__ jumps) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  bool jump = false;
  if (i < a.n) {
    const double x = (double)a.x[i], y = (double)a.y[i], z = (double)a.z[i];
    int64_t code = kRefNone;
    if (isfinite(x) && isfinite(y) && isfinite(z))
      code = (ref_coord(x, cs) << 42) | (ref_coord(y, cs) << 21) | ref_coord(z, cs);
    const int64_t prev = ref[i];
    jump = prev != kRefNone && code != kRefNone && ref_moved(prev, code);
    ref[i] = code;
  }
  const unsigned long long b = __ballot(jump);
  if ((threadIdx.x & 63) == 0 && b != 0ull) atomicAdd(jumps, (int)__popcll(b));
}

// Multi-scene binning: particle i belongs to scene scene[i], whose grid
// occupies cells [cell_base, cell_base + nx*ny*nz) of one global cell space,
// so one counting sort bins every scene and keeps scenes contiguous.
__global__ __launch_bounds__(kBlock) void batch_bin_count(int n, const float* __restrict__ x,
                                                          const float* __restrict__ y,
                                                          const float* __restrict__ z,
                                                          const int* __restrict__ scene,
                                                          const BatchBinGrid* __restrict__ grids,
                                                          int* __restrict__ cell_tmp,
                                                          int* __restrict__ counts) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const bool valid = i < n;
  int c = -1;
  if (valid) {
    const BatchBinGrid g = grids[scene[i]];
    const int cx = grid_coord((double)x[i], g.cs, g.lo[0], g.dims[0]);
    const int cy = grid_coord((double)y[i], g.cs, g.lo[1], g.dims[1]);
    const int cz = grid_coord((double)z[i], g.cs, g.lo[2], g.dims[2]);
    c = g.cell_base + (cz * g.dims[1] + cy) * g.dims[0] + cx;
    cell_tmp[i] = c;
  }
  const WaveRun r = wave_run(c, valid);
  if (r.is_head) atomicAdd(&counts[c], r.len);
}

__global__ __launch_bounds__(kBlock) void bin_scatter(int n, const int* __restrict__ cell_tmp,
                                                      const int* __restrict__ cstart,
                                                      int* __restrict__ cursor,
                                                      int* __restrict__ bucket) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const bool valid = i < n;
  const int c = valid ? cell_tmp[i] : -1;
  // one atomic per run of a cell; the run's members take consecutive
  // positions in lane order
  const WaveRun r = wave_run(c, valid);
  int base = 0;
  if (r.is_head) base = atomicAdd(&cursor[c], r.len);
  base = __shfl(base, r.head, 64);
  if (valid) bucket[cstart[c] + base + ((threadIdx.x & 63) - r.head)] = i;
}

// The atomic scatter leaves each cell's members in arbitrary order; rank
// every member by its current slot to get a deterministic, stable layout.
__global__ __launch_bounds__(kBlock) void bin_rank(int n, const int* __restrict__ cell_tmp,
                                                   const int* __restrict__ cstart,
                                                   const int* __restrict__ bucket,
                                                   int* __restrict__ perm,
                                                   int* __restrict__ cell_out) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= n) return;
  const int me = bucket[p];
  const int c = cell_tmp[me];
  const int s = cstart[c], e = cstart[c + 1];
  int rank = 0;
  for (int q = s; q < e; ++q) rank += bucket[q] < me ? 1 : 0;
  perm[s + rank] = me;
  cell_out[s + rank] = c;
}

// ---------------------------------------------------------------------------
// permutation gather
// ---------------------------------------------------------------------------
constexpr int kPermFields = 24;
struct PermArgs {
  const void* src[kPermFields];
  void* dst[kPermFields];
  int esz[kPermFields];
  int nf;
};

// One thread moves four consecutive slots p..p+3 of every field.  Slots
// whose sources are consecutive too (a run: the common case, since few
// particles change cell between two binnings) are copied with one 16-byte
// load per 4-byte field (two per 8-byte field); other groups gather the
// four elements one by one.  Loads go through bounds-checked buffer
// resources (dword-aligned 16-byte reads are allowed there); the stores are
// 16-byte aligned because p is a multiple of 4.  No output aliases an input
// (the entry point checks).
constexpr int kPermPer = 4;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t perm_rsrc(const void* base, int64_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, (int)(uint32_t)bytes, 0x00020000);
}

__global__ __launch_bounds__(kBlock) void permute_kernel(int n, const int* __restrict__ perm,
                                                         PermArgs a) {
  const int p = (blockIdx.x * kBlock + threadIdx.x) * kPermPer;
  if (p >= n) return;
  const bool full = p + kPermPer <= n;
  int j[kPermPer];
  if (full) {
    const int4 jj = *(const int4*)(perm + p);
    j[0] = jj.x; j[1] = jj.y; j[2] = jj.z; j[3] = jj.w;
  } else {
#pragma unroll
    for (int k = 0; k < kPermPer; ++k) j[k] = p + k < n ? perm[p + k] : 0;
  }
  const bool run = full && j[1] == j[0] + 1 && j[2] == j[0] + 2 && j[3] == j[0] + 3;
#pragma unroll 4
  for (int f = 0; f < a.nf; ++f) {
    const int esz = a.esz[f];
    const __amdgpu_buffer_rsrc_t rs = perm_rsrc(a.src[f], (int64_t)n * esz);
    if (esz == 4) {
      uint32_t* __restrict__ d = (uint32_t*)a.dst[f] + p;
      if (run) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b128(rs, j[0] * 4, 0, 0);
        *(uint4*)d = make_uint4(v[0], v[1], v[2], v[3]);
      } else {
        uint32_t v[kPermPer];
#pragma unroll
        for (int k = 0; k < kPermPer; ++k) v[k] = __builtin_amdgcn_raw_buffer_load_b32(rs, j[k] * 4, 0, 0);
        if (full) {
          *(uint4*)d = make_uint4(v[0], v[1], v[2], v[3]);
        } else {
#pragma unroll
          for (int k = 0; k < kPermPer; ++k)
            if (p + k < n) d[k] = v[k];
        }
      }
    } else {
      uint64_t* __restrict__ d = (uint64_t*)a.dst[f] + p;
      if (run) {
        const auto v0 = __builtin_amdgcn_raw_buffer_load_b128(rs, j[0] * 8, 0, 0);
        const auto v1 = __builtin_amdgcn_raw_buffer_load_b128(rs, j[0] * 8 + 16, 0, 0);
        *(uint4*)d = make_uint4(v0[0], v0[1], v0[2], v0[3]);
        *((uint4*)d + 1) = make_uint4(v1[0], v1[1], v1[2], v1[3]);
      } else {
        uint64_t v[kPermPer];
#pragma unroll
        for (int k = 0; k < kPermPer; ++k) {
          const auto w = __builtin_amdgcn_raw_buffer_load_b64(rs, j[k] * 8, 0, 0);
          v[k] = (uint64_t)w[0] | ((uint64_t)w[1] << 32);
        }
#pragma unroll
        for (int k = 0; k < kPermPer; ++k)
          if (p + k < n) d[k] = v[k];
      }
    }
  }
}

// Arrays past kPermRunMax elements (byte offsets beyond 32 bits), and
// outputs or permutations that are not 16-byte aligned, take the
// element-wise gather.
constexpr int64_t kPermRunMax = (int64_t)1 << 28;

__global__ __launch_bounds__(kBlock) void permute_kernel_wide(int n, const int* __restrict__ perm,
                                                              PermArgs a) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= n) return;
  const int64_t j = perm[p];
  for (int f = 0; f < a.nf; ++f) {
    if (a.esz[f] == 4)
      ((uint32_t*)a.dst[f])[p] = ((const uint32_t*)a.src[f])[j];
    else
      ((uint64_t*)a.dst[f])[p] = ((const uint64_t*)a.src[f])[j];
  }
}

// ---------------------------------------------------------------------------
// exact neighbour lists
// ---------------------------------------------------------------------------
template <typename T>
struct NbrArgs {
  View<T> dst, src;
  double rs;
  const int64_t* ids;
  const int64_t* off;
  int64_t* out;
  int64_t* counts;
  double rmax_sym;  // > 0: symmetric variable-h sets, stencil for this radius
  int64_t cap;      // fill: entries at positions >= cap are dropped...
  int32_t* overflow;  // ... and flagged here (nullable: cap is not checked)
};

template <typename T>
__device__ __forceinline__ bool pair_within(T xs, T ys, T zs, T xd, T yd, T zd, float r2f,
                                            double r2d);

template <>
__device__ __forceinline__ bool pair_within<float>(float xs, float ys, float zs, float xd,
                                                   float yd, float zd, float r2f, double r2d) {
  const float dx = xs - xd, dy = ys - yd, dz = zs - zd;
  const float d2 = dx * dx + dy * dy + dz * dz;
  return Within32::test(xs, ys, zs, xd, yd, zd, d2, r2f, r2d);
}

template <>
__device__ __forceinline__ bool pair_within<double>(double xs, double ys, double zs, double xd,
                                                    double yd, double zd, float, double r2d) {
  return exact_within(xs, ys, zs, xd, yd, zd, r2d);
}

template <typename T, bool FILL>
__global__ __launch_bounds__(kBlock) void nbr_kernel(Grid g, NbrArgs<T> a) {
  const int bid = xcd_block(blockIdx.x, gridDim.x);
  const int i = bid * kBlock + threadIdx.x;
  if (i >= a.dst.n) return;
  const T xd = a.dst.x[i], yd = a.dst.y[i], zd = a.dst.z[i];
  const double r = (double)a.dst.h[i] * a.rs;  // eff_radius (nnps.pyx:552)
  const double r2d = r * r;
  int cx, cy, cz;
  cell_xyz(a.dst.cell[i], g, cx, cy, cz);
  // ring = ceil(r/cs) for this destination (cell.pyx:1127-1129); the
  // symmetric sets also reach sources whose own radius covers the pair
  const double rd = ceil((a.rmax_sym > r ? a.rmax_sym : r) / g.cs);
  const int full = max(g.dims[0], max(g.dims[1], g.dims[2]));
  int ring = !(rd < (double)full) ? full : (int)rd;
  if (ring < 1) ring = 1;
  int64_t k = FILL ? a.off[i] : 0;
  // FILL: whether the ids arrive ascending (then the row needs no sort)
  int64_t last_id = INT64_MIN;
  bool ascending = true;
  const View<T>& S = a.src;
  for_each_image(g, (double)xd, (double)yd, (double)zd, cx, cy, cz, ring * g.cs,
                 [&](double ox, double oy, double oz, int icx, int icy, int icz, bool image) {
    for_each_row(g, S.cstart, icx, icy, icz, image ? ring + 1 : ring, [&](int j0, int j1) {
      for (int j = j0; j < j1; ++j) {
        double r2 = r2d;
        if (a.rmax_sym > 0) {
          // the reverse locator's test: d2 < (h_s * rs)^2 (nnps.pyx:735-747)
          const double rsj = (double)S.h[j] * a.rs;
          if (rsj * rsj > r2) r2 = rsj * rsj;
        }
        bool hit;
        if (!image) {
          hit = pair_within<T>(S.x[j], S.y[j], S.z[j], xd, yd, zd, (float)r2, r2);
        } else {
          // a ghost copy's coordinate, x_j + s*L, in FP64 (cell.pyx:1277-1376)
          hit = exact_within((double)S.x[j] + ox, (double)S.y[j] + oy, (double)S.z[j] + oz,
                             (double)xd, (double)yd, (double)zd, r2);
        }
        if (hit) {
          if (FILL) {
            const int64_t id = a.ids[j];
            ascending = ascending && id > last_id;
            last_id = id;
            if (a.overflow == nullptr || k < a.cap) a.out[k] = id;
          }
          ++k;
        }
      }
    });
  });
  if (!FILL) {
    a.counts[i] = k;
  } else {
    if (a.overflow != nullptr && k > a.cap) {
      *a.overflow = 1;  // plain vector store of a flag
      k = a.cap;
    }
    // sort the row ascending (rows hold tens of entries); a row whose ids
    // arrived ascending (slot ids of cell-ordered sources) is already sorted
    const int64_t b = a.off[i];
    for (int64_t p = b + 1; !ascending && p < k; ++p) {
      const int64_t v = a.out[p];
      int64_t q = p - 1;
      while (q >= b && a.out[q] > v) {
        a.out[q + 1] = a.out[q];
        --q;
      }
      a.out[q + 1] = v;
    }
  }
}

}  // namespace sph

using namespace sph;

namespace sph {

// Binning of a multi-scene particle set (see batch_bin_count); outputs as
// sph_bin: perm (new slot -> old slot), cell (global cell id per new slot),
// cell_start (ncells + 1).  The workspace is sph_bin_workspace_bytes(n, ncells).
int32_t batch_bin(hipStream_t st, int64_t n, int64_t ncells, const float* x, const float* y,
                  const float* z, const int* scene, const BatchBinGrid* grids, int32_t* perm,
                  int32_t* cell, int32_t* cell_start, void* ws) {
  char* w = (char*)ws;
  int* counts = (int*)w;
  w += (size_t)(ncells + 64) * 4;
  int* cell_tmp = (int*)w;
  w += (size_t)(n + 64) * 4;
  int* bucket = (int*)w;
  w += (size_t)(n + 64) * 4;
  int* partial = (int*)w;
  SPH_TRY(fill_i32(st, counts, ncells, 0));
  const int nb = nblocks_for(n);
  if (n > 0)
    hipLaunchKernelGGL(batch_bin_count, dim3(nb), dim3(kBlock), 0, st, (int)n, x, y, z, scene,
                       grids, cell_tmp, counts);
  SPH_TRY(launch_status("batch_bin/count"));
  SPH_TRY((exclusive_scan<int, int>(st, counts, ncells, cell_start, partial)));
  if (n == 0) return SPH_OK;
  SPH_TRY(fill_i32(st, counts, ncells, 0));
  hipLaunchKernelGGL(bin_scatter, dim3(nb), dim3(kBlock), 0, st, (int)n, cell_tmp, cell_start,
                     counts, bucket);
  hipLaunchKernelGGL(bin_rank, dim3(nb), dim3(kBlock), 0, st, (int)n, cell_tmp, cell_start,
                     bucket, perm, cell);
  return launch_status("batch_bin");
}

int32_t permute_fields(hipStream_t st, int64_t n, const int32_t* perm, int32_t nfields,
                       const void* const* src, void* const* dst, const int32_t* esz) {
  return sph_permute(st, n, perm, nfields, src, dst, esz);
}

}  // namespace sph

extern "C" {

int32_t sph_fill_i32(void* stream, int32_t* dst_dev, int64_t n, int32_t value) {
  if (n > 0 && dst_dev == nullptr) {
    set_error("sph_fill_i32: NULL destination");
    return SPH_EINVAL;
  }
  return fill_i32(as_stream(stream), dst_dev, n, value);
}

int32_t sph_bounds(void* stream, const sph_view* a, double* out_dev) {
  SPH_TRY(check_view(a, "sph_bounds", false, false, false));
  if (out_dev == nullptr) {
    set_error("sph_bounds: NULL output");
    return SPH_EINVAL;
  }
  hipStream_t st = as_stream(stream);
  // partials live at the tail of the output buffer's companion: use a small
  // fixed number of blocks and reduce them in a second single-block pass.
  const int nb = 64;
  static_assert(nb * 7 <= 4096, "partials");
  double* part = nullptr;
  // The 7-double output is followed by caller-provided room for partials.
  part = out_dev + 8;
  if (a->dtype == SPH_F32)
    hipLaunchKernelGGL(bounds_partial<float>, dim3(nb), dim3(kBlock), 0, st, to_view<float>(*a),
                       part);
  else
    hipLaunchKernelGGL(bounds_partial<double>, dim3(nb), dim3(kBlock), 0, st,
                       to_view<double>(*a), part);
  hipLaunchKernelGGL(bounds_final, dim3(1), dim3(64), 0, st, part, nb, out_dev);
  return launch_status("sph_bounds");
}

size_t sph_bin_workspace_bytes(int64_t n, int64_t ncells) {
  // counts/cursor (ncells), cell_tmp (n), bucket (n), scan partials
  const int64_t nb = (ncells + 1 + kScanTile - 1) / kScanTile + 1;
  size_t b = 0;
  b += (size_t)(ncells + 64) * 4;
  b += (size_t)(n + 64) * 4;
  b += (size_t)(n + 64) * 4;
  b += (size_t)(nb + 64) * 4;
  return b + 1024;
}

int32_t sph_bin(void* stream, const sph_view* a, const sph_grid* g, int32_t* perm_dev,
                int32_t* cell_dev, int32_t* cell_start_dev, const int32_t* prev_cell_dev,
                int32_t* jump_count_dev, void* workspace_dev, size_t workspace_bytes) {
  SPH_TRY(check_view(a, "sph_bin", false, false, false));
  SPH_TRY(check_grid(g, "sph_bin"));
  const int64_t n = a->n;
  const int64_t ncells = (int64_t)g->dims[0] * g->dims[1] * g->dims[2];
  if (workspace_bytes < sph_bin_workspace_bytes(n, ncells) || workspace_dev == nullptr) {
    set_error("sph_bin: workspace of %zu bytes, %zu needed", workspace_bytes,
              sph_bin_workspace_bytes(n, ncells));
    return SPH_EINVAL;
  }
  if (!cell_start_dev || (n > 0 && (!perm_dev || !cell_dev))) {
    set_error("sph_bin: NULL output");
    return SPH_EINVAL;
  }
  if (prev_cell_dev != nullptr && jump_count_dev == nullptr) {
    set_error("sph_bin: prev_cell given without jump_count");
    return SPH_EINVAL;
  }
  hipStream_t st = as_stream(stream);
  char* w = (char*)workspace_dev;
  int* counts = (int*)w;
  w += (size_t)(ncells + 64) * 4;
  int* cell_tmp = (int*)w;
  w += (size_t)(n + 64) * 4;
  int* bucket = (int*)w;
  w += (size_t)(n + 64) * 4;
  int* partial = (int*)w;

  BinGrid bg;
  bg.cs = g->cell_size;
  for (int d = 0; d < 3; ++d) {
    bg.lo[d] = g->lo[d];
    bg.dims[d] = g->dims[d];
  }
  bg.k = g->bins_per_cell > 1 ? g->bins_per_cell : 1;
  SPH_TRY(fill_i32(st, counts, ncells, 0));
  const int nb = nblocks_for(n);
  if (n > 0) {
    if (a->dtype == SPH_F32)
      hipLaunchKernelGGL(bin_count<float>, dim3(nb), dim3(kBlock), 0, st, to_view<float>(*a),
                         bg, cell_tmp, counts, prev_cell_dev, jump_count_dev);
    else
      hipLaunchKernelGGL(bin_count<double>, dim3(nb), dim3(kBlock), 0, st,
                         to_view<double>(*a), bg, cell_tmp, counts, prev_cell_dev,
                         jump_count_dev);
    SPH_TRY(launch_status("sph_bin/count"));
  }
  SPH_TRY((exclusive_scan<int, int>(st, counts, ncells, cell_start_dev, partial)));
  if (n == 0) return SPH_OK;
  SPH_TRY(fill_i32(st, counts, ncells, 0));
  hipLaunchKernelGGL(bin_scatter, dim3(nb), dim3(kBlock), 0, st, (int)n, cell_tmp,
                     cell_start_dev, counts, bucket);
  hipLaunchKernelGGL(bin_rank, dim3(nb), dim3(kBlock), 0, st, (int)n, cell_tmp,
                     cell_start_dev, bucket, perm_dev, cell_dev);
  return launch_status("sph_bin");
}

int32_t sph_jump_check(void* stream, const sph_view* a, double cell_size, int64_t* refcell_dev,
                       int32_t* jump_count_dev) {
  SPH_TRY(check_view(a, "sph_jump_check", false, false, false));
  const int64_t n = a->n;
  if (!(cell_size > 0.0) || (n > 0 && (!refcell_dev || !jump_count_dev))) {
    set_error("sph_jump_check: bad arguments (cell size %g)", cell_size);
    return SPH_EINVAL;
  }
  if (n == 0) return SPH_OK;
  hipStream_t st = as_stream(stream);
  const int nb = nblocks_for(n);
  if (a->dtype == SPH_F32)
    hipLaunchKernelGGL(jump_check_kernel<float>, dim3(nb), dim3(kBlock), 0, st,
                       to_view<float>(*a), cell_size, refcell_dev, jump_count_dev);
  else
    hipLaunchKernelGGL(jump_check_kernel<double>, dim3(nb), dim3(kBlock), 0, st,
                       to_view<double>(*a), cell_size, refcell_dev, jump_count_dev);
  return launch_status("sph_jump_check");
}

int32_t sph_permute(void* stream, int64_t n, const int32_t* perm_dev, int32_t nfields,
                    const void* const* src_dev_host, void* const* dst_dev_host,
                    const int32_t* elem_bytes_host) {
  if (n < 0 || nfields < 0 || (n > 0 && nfields > 0 && (!perm_dev || !src_dev_host ||
                                                          !dst_dev_host || !elem_bytes_host))) {
    set_error("sph_permute: bad arguments");
    return SPH_EINVAL;
  }
  hipStream_t st = as_stream(stream);
  for (int f = 0; f < nfields; ++f)
    for (int g = 0; g < nfields; ++g)
      if (n > 0 && dst_dev_host[f] != nullptr && dst_dev_host[f] == src_dev_host[g]) {
        set_error("sph_permute: output %d aliases input %d", f, g);
        return SPH_EINVAL;
      }
  for (int f0 = 0; f0 < nfields; f0 += kPermFields) {
    PermArgs pa;
    pa.nf = nfields - f0 < kPermFields ? nfields - f0 : kPermFields;
    for (int k = 0; k < pa.nf; ++k) {
      const int f = f0 + k;
      if ((elem_bytes_host[f] != 4 && elem_bytes_host[f] != 8) || !src_dev_host[f] ||
          !dst_dev_host[f] || src_dev_host[f] == dst_dev_host[f]) {
        set_error("sph_permute: field %d invalid (elem bytes %d, in-place not allowed)", f,
                  elem_bytes_host[f]);
        return SPH_EINVAL;
      }
      pa.src[k] = src_dev_host[f];
      pa.dst[k] = dst_dev_host[f];
      pa.esz[k] = elem_bytes_host[f];
    }
    if (n == 0) continue;
    // the run kernel reads perm and writes every output 16 bytes at a time
    bool aligned = ((uintptr_t)perm_dev & 15) == 0;
    for (int k = 0; k < pa.nf; ++k) aligned = aligned && ((uintptr_t)pa.dst[k] & 15) == 0;
    if (n > kPermRunMax || !aligned)
      hipLaunchKernelGGL(permute_kernel_wide, dim3(nblocks_for(n)), dim3(kBlock), 0, st, (int)n,
                         perm_dev, pa);
    else
      hipLaunchKernelGGL(permute_kernel, dim3(nblocks_for(n, kBlock * kPermPer)), dim3(kBlock), 0,
                         st, (int)n, perm_dev, pa);
  }
  return launch_status("sph_permute");
}

static int32_t nbr_common(const sph_grid* g, const sph_view* dst, const sph_view* src,
                          double rs, const char* who) {
  SPH_TRY(check_grid(g, who));
  SPH_TRY(check_view(dst, who, false, false, true));
  SPH_TRY(check_view(src, who, false, false, true));
  if (dst->dtype != src->dtype) {
    set_error("%s: dst and src dtypes differ", who);
    return SPH_EUNSUPPORTED;
  }
  if (!(rs > 0)) {
    set_error("%s: radius_scale must be positive", who);
    return SPH_EINVAL;
  }
  return SPH_OK;
}

static int32_t nbr_counts(void* stream, const sph_grid* g, const sph_view* dst,
                          const sph_view* src, double radius_scale, double rmax_sym,
                          int64_t* counts_dev, const char* who) {
  SPH_TRY(nbr_common(g, dst, src, radius_scale, who));
  if (dst->n == 0) return SPH_OK;
  if (!counts_dev) {
    set_error("%s: NULL output", who);
    return SPH_EINVAL;
  }
  hipStream_t st = as_stream(stream);
  Grid G = make_grid(*g, 0.0);
  if (rmax_sym > 0) SPH_TRY(check_period(g, rmax_sym, who));
  const int nb = nblocks_for(dst->n);
  if (dst->dtype == SPH_F32) {
    NbrArgs<float> a{to_view<float>(*dst), to_view<float>(*src), radius_scale, nullptr,
                     nullptr, nullptr, counts_dev, rmax_sym};
    hipLaunchKernelGGL((nbr_kernel<float, false>), dim3(nb), dim3(kBlock), 0, st, G, a);
  } else {
    NbrArgs<double> a{to_view<double>(*dst), to_view<double>(*src), radius_scale, nullptr,
                      nullptr, nullptr, counts_dev, rmax_sym};
    hipLaunchKernelGGL((nbr_kernel<double, false>), dim3(nb), dim3(kBlock), 0, st, G, a);
  }
  return launch_status(who);
}

static int32_t nbr_fill(void* stream, const sph_grid* g, const sph_view* dst,
                        const sph_view* src, double radius_scale, double rmax_sym,
                        const int64_t* src_ids_dev, const int64_t* offsets_dev,
                        int64_t* nbrs_dev, int64_t cap, int64_t total, int64_t* needed_host,
                        const char* who, int32_t* overflow = nullptr) {
  SPH_TRY(nbr_common(g, dst, src, radius_scale, who));
  if (overflow == nullptr) {
    if (needed_host) *needed_host = total;
    if (cap < total) {
      set_error("%s: capacity %lld < %lld pairs", who, (long long)cap, (long long)total);
      return SPH_ECAPACITY;
    }
    if (total == 0) return SPH_OK;
  }
  if (dst->n == 0) return SPH_OK;
  if (!src_ids_dev || !offsets_dev || !nbrs_dev) {
    set_error("%s: NULL argument", who);
    return SPH_EINVAL;
  }
  hipStream_t st = as_stream(stream);
  Grid G = make_grid(*g, 0.0);
  const int nb = nblocks_for(dst->n);
  if (dst->dtype == SPH_F32) {
    NbrArgs<float> a{to_view<float>(*dst), to_view<float>(*src), radius_scale, src_ids_dev,
                     offsets_dev, nbrs_dev, nullptr, rmax_sym, cap, overflow};
    hipLaunchKernelGGL((nbr_kernel<float, true>), dim3(nb), dim3(kBlock), 0, st, G, a);
  } else {
    NbrArgs<double> a{to_view<double>(*dst), to_view<double>(*src), radius_scale, src_ids_dev,
                      offsets_dev, nbrs_dev, nullptr, rmax_sym, cap, overflow};
    hipLaunchKernelGGL((nbr_kernel<double, true>), dim3(nb), dim3(kBlock), 0, st, G, a);
  }
  return launch_status(who);
}

int32_t sph_neighbor_fill_capped(void* stream, const sph_grid* g, const sph_view* dst,
                                 const sph_view* src, double radius_scale,
                                 const int64_t* src_ids_dev, const int64_t* offsets_dev,
                                 int64_t* nbrs_dev, int64_t cap, int32_t* overflow_dev) {
  if (cap < 0 || overflow_dev == nullptr ||
      (dst != nullptr && dst->n > 0 && (!src_ids_dev || !offsets_dev || !nbrs_dev))) {
    set_error("sph_neighbor_fill_capped: bad arguments");
    return SPH_EINVAL;
  }
  return nbr_fill(stream, g, dst, src, radius_scale, 0.0, src_ids_dev, offsets_dev, nbrs_dev,
                  cap, -1, nullptr, "sph_neighbor_fill_capped", overflow_dev);
}

int32_t sph_neighbor_counts(void* stream, const sph_grid* g, const sph_view* dst,
                            const sph_view* src, double radius_scale, int64_t* counts_dev) {
  return nbr_counts(stream, g, dst, src, radius_scale, 0.0, counts_dev, "sph_neighbor_counts");
}

int32_t sph_varh_neighbor_counts(void* stream, const sph_grid* g, const sph_view* dst,
                                 const sph_view* src, double radius_scale, double rmax,
                                 int64_t* counts_dev) {
  if (!(rmax > 0)) {
    set_error("sph_varh_neighbor_counts: rmax must be positive");
    return SPH_EINVAL;
  }
  return nbr_counts(stream, g, dst, src, radius_scale, rmax, counts_dev,
                    "sph_varh_neighbor_counts");
}

}  // extern "C"

namespace sph {

// Scratch of sph_scan_counts_noalloc: the tile partials of the exclusive
// scan, in a file-scope device array instead of a per-call allocation.
// Each GPU holds its own copy until the process ends; host code uses the
// address hipGetSymbolAddress returns for the current device, never the
// array's name.  Calls that overlap on one GPU share it and race.
constexpr int kNoAllocPartials = 1 << 16;  // 512 KiB of int64 partials
constexpr int64_t kNoAllocMaxCounts = (int64_t)kNoAllocPartials * kScanTile;
static __device__ __align__(256) int64_t scratch[kNoAllocPartials];
static_assert(kNoAllocMaxCounts == SPH_SCAN_NOALLOC_MAX_COUNTS,
              "the header's limit is the scratch's capacity");

constexpr int kNoAllocMaxDevices = 64;
static std::mutex g_scratch_mu;
static int64_t* g_scratch_ptr[kNoAllocMaxDevices] = {};

// The current device's address of `scratch`, looked up once per device.
static int32_t noalloc_scratch(int64_t** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    set_error("sph_scan_counts_noalloc: %s", hipGetErrorString(e));
    return (int32_t)e;
  }
  if (dev < 0 || dev >= kNoAllocMaxDevices) {
    set_error("sph_scan_counts_noalloc: device %d beyond %d", dev, kNoAllocMaxDevices);
    return SPH_EUNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      set_error("sph_scan_counts_noalloc: hipGetSymbolAddress: %s",
                e != hipSuccess ? hipGetErrorString(e) : "null address");
      return e != hipSuccess ? (int32_t)e : SPH_EINVAL;
    }
    g_scratch_ptr[dev] = (int64_t*)p;
  }
  *out = g_scratch_ptr[dev];
  return SPH_OK;
}

}  // namespace sph

extern "C" {

int32_t sph_scan_counts(void* stream, int64_t n, const int64_t* counts_dev,
                        int64_t* offsets_dev, int64_t* total_host) {
  if (n < 0 || !offsets_dev || (n > 0 && !counts_dev)) {
    set_error("sph_scan_counts: bad arguments");
    return SPH_EINVAL;
  }
  hipStream_t st = as_stream(stream);
  const int64_t nb = (n + kScanTile - 1) / kScanTile + 1;
  int64_t* partial = nullptr;
  hipError_t e = hipMallocAsync((void**)&partial, (size_t)nb * 8, st);
  if (e != hipSuccess) {
    set_error("sph_scan_counts: %s", hipGetErrorString(e));
    return (int32_t)e;
  }
  int32_t rc = exclusive_scan<int64_t, int64_t>(st, counts_dev, n, offsets_dev, partial);
  (void)hipFreeAsync(partial, st);
  if (rc != SPH_OK) return rc;
  if (total_host) {
    e = hipMemcpyAsync(total_host, offsets_dev + n, 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      set_error("sph_scan_counts: %s", hipGetErrorString(e));
      return (int32_t)e;
    }
  }
  return SPH_OK;
}

int32_t sph_scan_counts_noalloc(void* stream, int64_t n, const int64_t* counts_dev,
                                int64_t* offsets_dev, int64_t* total_host) {
  if (n < 0 || !offsets_dev || (n > 0 && !counts_dev)) {
    set_error("sph_scan_counts_noalloc: bad arguments");
    return SPH_EINVAL;
  }
  if (n > kNoAllocMaxCounts) {
    set_error("sph_scan_counts_noalloc: %lld counts exceed the maximum of %lld (scratch of "
              "%d tile partials); use sph_scan_counts",
              (long long)n, (long long)kNoAllocMaxCounts, kNoAllocPartials);
    return SPH_EINVAL;
  }
  int64_t* partial = nullptr;
  SPH_TRY(noalloc_scratch(&partial));
  hipStream_t st = as_stream(stream);
  // every partial the scan reads is written by the same call first, so the
  // previous call's contents need no clearing
  SPH_TRY((exclusive_scan<int64_t, int64_t>(st, counts_dev, n, offsets_dev, partial)));
  if (total_host) {
    hipError_t e = hipMemcpyAsync(total_host, offsets_dev + n, 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      set_error("sph_scan_counts_noalloc: %s", hipGetErrorString(e));
      return (int32_t)e;
    }
  }
  return SPH_OK;
}

int32_t sph_neighbor_fill(void* stream, const sph_grid* g, const sph_view* dst,
                          const sph_view* src, double radius_scale, const int64_t* src_ids_dev,
                          const int64_t* offsets_dev, int64_t* nbrs_dev, int64_t cap,
                          int64_t total, int64_t* needed_host) {
  return nbr_fill(stream, g, dst, src, radius_scale, 0.0, src_ids_dev, offsets_dev, nbrs_dev,
                  cap, total, needed_host, "sph_neighbor_fill");
}

int32_t sph_varh_neighbor_fill(void* stream, const sph_grid* g, const sph_view* dst,
                               const sph_view* src, double radius_scale, double rmax,
                               const int64_t* src_ids_dev, const int64_t* offsets_dev,
                               int64_t* nbrs_dev, int64_t cap, int64_t total,
                               int64_t* needed_host) {
  if (!(rmax > 0)) {
    set_error("sph_varh_neighbor_fill: rmax must be positive");
    return SPH_EINVAL;
  }
  return nbr_fill(stream, g, dst, src, radius_scale, rmax, src_ids_dev, offsets_dev, nbrs_dev,
                  cap, total, needed_host, "sph_varh_neighbor_fill");
}

}  // extern "C"
