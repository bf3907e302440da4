// Device-side building blocks shared by the gfx950 kernels: the cubic-spline
// kernel, the exact neighbour test, the cell-row stencil walk and the
// XCD-aware block mapping.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sph_mi355x.h"

namespace sph {

constexpr double kPi = 3.14159265358979323846;
constexpr int kBlock = 256;  // threads per block of the particle kernels
constexpr int kXcds = 8;     // MI355X: 8 XCDs, blocks dispatched round-robin

// Dispatch sends block b to XCD b % 8.  Remap so that each XCD walks one
// contiguous chunk of the (cell-sorted) particle range: neighbouring
// destinations share source cells, so this keeps each XCD's L2 hot.
__device__ __forceinline__ int xcd_block(int bid, int nblocks) {
  if (nblocks < 2 * kXcds) return bid;
  const int per = nblocks / kXcds;
  const int full = per * kXcds;
  if (bid >= full) return bid;  // tail blocks keep their id
  return (bid % kXcds) * per + bid / kXcds;
}

// Grid and search parameters passed to every neighbour kernel by value.
struct Grid {
  double cs;     // cell size
  int lo[3];     // integer cell coordinate of grid cell (0,0,0)
  int dims[3];
  float cs_f;
  float inv_cs_f;  // 1 / cs_f (row culling only; conservative by its margin)
  double rmax;   // largest search radius of the call (radius_scale * max h)
  int ring0;     // ceil(rmax / cs)
  int cell_base; // cell ids stored per particle are cell_base + local id
  int periodic;  // any periodic axis
  double per[3];  // period per axis, 0 = not periodic
  double pmin[3]; // lower face of the periodic domain
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

// Upper bound, in FP32, of a particle's straight-line displacement from its
// per-axis moves summed in squares (mv2, FP64).  A kept binning is reused
// with every search radius widened by the largest such bound: a source can
// lie that far (in any direction, not only along one axis) from the cell it
// was binned in, so the straight-line bound, not the largest per-axis move,
// keeps the culled stencil exact.  It also bounds every per-axis move, which
// is what the ring below needs.
__device__ __forceinline__ float move_bound(double mv2) {
  return (float)(sqrt(mv2) * (1.0 + 1e-6));
}

// Stencil ring: ceil(r/cs) cells (cell.pyx:1127-1129) widened by the largest
// displacement since binning, so that a binning stays exact when reused.
__device__ __forceinline__ int ring_of(const Grid& g, const float* disp) {
  if (disp == nullptr) return g.ring0;
  const double d = (double)(*disp);
  const double rd = ceil((g.rmax + 2.0 * d) / g.cs);
  // a ring as wide as the grid already covers every cell (also for inf/NaN)
  const int full = max(g.dims[0], max(g.dims[1], g.dims[2]));
  if (!(rd < (double)full)) return full;
  const int r = (int)rd;
  return r < g.ring0 ? g.ring0 : r;
}

// Cubic spline normalisation fac(1) (kernels.pyx:437-443).
__host__ __device__ __forceinline__ double fac_unit(int dim) {
  return dim == 1 ? 2.0 / 3.0 : (dim == 2 ? 10.0 / (7.0 * kPi) : 1.0 / kPi);
}

template <typename CT>
__device__ __forceinline__ CT h_dim_inv(CT inv_h, int dim) {
  return dim == 1 ? inv_h : (dim == 2 ? inv_h * inv_h : inv_h * inv_h * inv_h);
}

// W(q) without the normalisation (kernels.pyx:367-374).
template <typename CT>
__device__ __forceinline__ CT w_shape(CT q) {
  if (q > CT(2)) return CT(0);
  if (q > CT(1)) {
    const CT t = CT(2) - q;
    return CT(0.25) * t * t * t;
  }
  return CT(1) - CT(1.5) * (q * q) * (CT(1) - CT(0.5) * q);
}

// Scalar gradient factor: grad W = r_ab * gfac (kernels.pyx:403-412), without
// the normalisation.  r is |r_ab|, q = r/h, inv_h = 1/h.
template <typename CT>
__device__ __forceinline__ CT grad_shape(CT q, CT r, CT h, CT inv_h) {
  if (q > CT(2)) return CT(0);
  if (q >= CT(1)) {
    const CT t = CT(2) - q;
    return CT(-0.75) * t * t / (h * r);
  }
  if (q > CT(1e-14)) return CT(3) * (CT(0.75) * q - CT(1)) * (inv_h * inv_h);
  return CT(0);
}

// Exact neighbour test of the reference, computed in FP64 without
// contraction: ((xs-xd)^2 + (ys-yd)^2) + (zs-zd)^2 < r2 (point.pxd:35-37,
// nnps.pyx:337).
__device__ __forceinline__ bool exact_within(double xs, double ys, double zs, double xd,
                                             double yd, double zd, double r2) {
#pragma clang fp contract(off)
  const double dx = xs - xd, dy = ys - yd, dz = zs - zd;
  const double d2 = (dx * dx + dy * dy) + dz * dz;
  return d2 < r2;
}

// FP32 positions: screen in FP32, decide exactly in FP64 only inside a band
// of +-2e-6 relative around the cut-off.  The FP32 distance has a relative
// error below 5*2^-24 (3e-7) and r2f below 2^-24, so outside the band the
// FP32 verdict equals the FP64 verdict.
struct Within32 {
  __device__ __forceinline__ static bool test(float xs, float ys, float zs, float xd,
                                              float yd, float zd, float d2f,
                                              float r2f, double r2d) {
    if (d2f < r2f * 0.999998f) return true;
    if (d2f > r2f * 1.000002f) return false;
    return exact_within((double)xs, (double)ys, (double)zs, (double)xd, (double)yd,
                        (double)zd, r2d);
  }
};

// Cell coordinate decode of a linear cell id (x fastest).
__device__ __forceinline__ void cell_xyz(int c, const Grid& g, int& cx, int& cy, int& cz) {
  c -= g.cell_base;
  cx = c % g.dims[0];
  const int t = c / g.dims[0];
  cy = t % g.dims[1];
  cz = t / g.dims[1];
}

// Walk the stencil of cell (cx,cy,cz) with the given ring: cells of one
// (dy,dz) row are contiguous in the cell-sorted layout, so each row is one
// particle range [cell_start[first], cell_start[last+1]).
template <class F>
__device__ __forceinline__ void for_each_row(const Grid& g, const int* __restrict__ cstart,
                                             int cx, int cy, int cz, int ring, F&& f) {
  const int x0 = max(cx - ring, 0), x1 = min(cx + ring, g.dims[0] - 1);
  const int y0 = max(cy - ring, 0), y1 = min(cy + ring, g.dims[1] - 1);
  const int z0 = max(cz - ring, 0), z1 = min(cz + ring, g.dims[2] - 1);
  for (int z = z0; z <= z1; ++z) {
    for (int y = y0; y <= y1; ++y) {
      const int base = (z * g.dims[1] + y) * g.dims[0];
      const int j0 = cstart[base + x0];
      const int j1 = cstart[base + x1 + 1];
      f(j0, j1);
    }
  }
}

// Periodic images of a destination at p in cell (cx,cy,cz): f(ox,oy,oz,
// icx,icy,icz,image) runs for the destination itself (offset 0) and, along
// periodic axes, for every shift s in {-1,0,1}^3 whose image point p - s*L
// comes within `reach` of the domain: sources are then seen at x_j + s*L,
// the ghost copy's coordinate (cell.pyx:1277-1376).  With period > 2*reach
// a source matches at most one image.  Image cells come from the image
// point in FP64; unshifted axes keep the destination's own cell.
template <class F>
__device__ __forceinline__ void for_each_image(const Grid& g, double px, double py, double pz,
                                               int cx, int cy, int cz, double reach, F&& f) {
  f(0.0, 0.0, 0.0, cx, cy, cz, false);
  if (!g.periodic) return;
  const double p[3] = {px, py, pz};
  const int c0[3] = {cx, cy, cz};
  int lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = 0;
    hi[a] = 0;
    if (g.per[a] > 0.0) {
      if (p[a] + reach >= g.pmin[a] + g.per[a]) hi[a] = 1;
      if (p[a] - reach < g.pmin[a]) lo[a] = -1;
    }
  }
  for (int sx = lo[0]; sx <= hi[0]; ++sx)
    for (int sy = lo[1]; sy <= hi[1]; ++sy)
      for (int sz = lo[2]; sz <= hi[2]; ++sz) {
        if (sx == 0 && sy == 0 && sz == 0) continue;
        const int sh[3] = {sx, sy, sz};
        double o[3];
        int c[3];
        for (int a = 0; a < 3; ++a) {
          o[a] = sh[a] * g.per[a];
          if (sh[a] == 0) {
            c[a] = c0[a];
          } else {
            const double k = floor((p[a] - o[a]) / g.cs) - (double)g.lo[a];
            c[a] = k < 0.0 ? 0 : (k > (double)(g.dims[a] - 1) ? g.dims[a] - 1 : (int)k);
          }
        }
        f(o[0], o[1], o[2], c[0], c[1], c[2], true);
      }
}

template <typename T>
struct View {
  int n;
  const T *x, *y, *z, *u, *v, *w, *h, *m, *rho, *p, *cs;
  const int64_t* tag;
  const int* cell;
  const int* cstart;
  const float4* pos4;  // packed (x,y,z,h), FP32 storage only
  const float4* vel4;  // packed (u,v,w,m)
  float4* aux4;        // packed (rho,cs,p/rho^2,p)
  const float* pos3;   // packed (x,y,z), 3 floats per particle
  const float4* pm4;   // packed (x,y,z,m)
  const int* hrange;   // h-uniformity pair (FP32 bits; equal: uniform h)
  const int* mrange;   // mass-uniformity pair (FP32 bits; equal: uniform m)
};

template <typename T>
inline View<T> to_view(const sph_view& a) {
  View<T> v;
  v.n = (int)a.n;
  v.x = (const T*)a.x; v.y = (const T*)a.y; v.z = (const T*)a.z;
  v.u = (const T*)a.u; v.v = (const T*)a.v; v.w = (const T*)a.w;
  v.h = (const T*)a.h; v.m = (const T*)a.m; v.rho = (const T*)a.rho;
  v.p = (const T*)a.p; v.cs = (const T*)a.cs;
  v.tag = a.tag; v.cell = a.cell; v.cstart = a.cell_start;
  v.pos4 = (const float4*)a.pos4;
  v.vel4 = (const float4*)a.vel4;
  v.aux4 = (float4*)a.aux4;
  v.pos3 = (const float*)a.pos3;
  v.pm4 = (const float4*)a.pm4;
  v.hrange = (const int*)a.hrange;
  v.mrange = (const int*)a.mrange;
  return v;
}

}  // namespace sph
