// Edge-MLP message-passing surrogate: forward, backward and Adam on the GPU.
//
// Network (per node i of a symmetric radius graph, hidden width 128):
//   H0   = ReLU(X_i W_in + b_in)                       X_i: 16 node features
//   for r in 1..2:
//     [P | Q] = H W_ab                                  (MFMA GEMM, N x 256)
//     A_i  = sum_j ReLU(P_i + Q_j + e_ij W_c + b_e)     e_ij: 16 edge features
//     U    = ReLU([H | A] W_u + c_u)                    (MFMA GEMM, N x 128)
//     H    = H + U
//   Y_i  = H_i W_out + b_out + a_i / g                  (2 outputs: acceleration / g)
// where a_i is the node's SPH pressure + artificial-viscosity acceleration
// (a parameter-free skip path kept in FP32): the network learns the
// correction that turns it into the acceleration over a 4x time step.
// Weights are kept in FP32 (Adam) with BF16 copies that feed the MFMA GEMMs
// (v_mfma_f32_32x32x16_bf16, FP32 accumulation).  Edge work is a fused
// gather / ReLU / sum per destination (one wave per node, two channels per
// lane); the backward pass recomputes the edge pre-activations and uses the
// graph's symmetry (j in N(i) <=> i in N(j)) to turn the scatter into Q's
// gradient into a gather.  Weight gradients are reduced through per-block
// partials, so every result is deterministic.
#include <algorithm>
#include <vector>

#include "sph_internal.h"

namespace sph {
namespace mlp {

constexpr int kH = SPH_MLP_HIDDEN;   // 128
constexpr int kIn = SPH_MLP_NIN;     // 16 node features
constexpr int kE = SPH_MLP_NE;       // 16 edge features
constexpr int kR = SPH_MLP_ROUNDS;   // 2
constexpr int kOut = 2;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// ---- parameter layout (FP32 master; the BF16 copy uses the same offsets) --
struct Layout {
  int64_t w_in, b_in, w_ab[kR], w_c[kR], b_e[kR], w_u[kR], c_u[kR], w_out, b_out, total;
};
static Layout layout() {
  Layout L;
  int64_t o = 0;
  L.w_in = o; o += kIn * kH;
  L.b_in = o; o += kH;
  for (int r = 0; r < kR; ++r) {
    L.w_ab[r] = o; o += kH * 2 * kH;
    L.w_c[r] = o; o += kE * kH;
    L.b_e[r] = o; o += kH;
    L.w_u[r] = o; o += 2 * kH * kH;
    L.c_u[r] = o; o += kH;
  }
  L.w_out = o; o += kH * kOut;
  L.b_out = o; o += 4;
  L.total = o;
  return L;
}

// ---------------------------------------------------------------------------
// GEMM: C[M x N] = op(A)[M x K] op(B)[K x N], BF16 operands, FP32 accumulate.
// A(m,k) = TA ? A[k*lda + m] : A[m*lda + k];  B(k,n) = TB ? B[n*ldb + k] : B[k*ldb + n].
// Block tile 64 x 64 (four 32 x 32 wave tiles), K in steps of 16.  Split-K:
// block z reduces k in [z*kc, (z+1)*kc) and writes its own partial slab.
// ---------------------------------------------------------------------------
struct GemmArgs {
  int M, N, K, kc;
  const __bf16* A;
  int64_t lda;
  const __bf16* B;
  int64_t ldb;
  float* C;            // FP32 output (may be null)
  int64_t ldc;
  int64_t c_zstride;   // split-K: slab z at C + z * c_zstride
  int accumulate;      // C += result
  const float* bias;   // [N] or null
  int relu;
  float* U;            // activation before the residual (null: not stored)
  int64_t ldu;
  const float* res;    // residual added after the activation (null: none)
  int64_t ldr;
  __bf16* Cb;          // BF16 copy of the output (null: none)
  int64_t ldcb;
};

template <bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_kernel(GemmArgs g) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wm = w & 1, wn = w >> 1;
  const int row = blockIdx.y * 64 + wm * 32 + (lane & 31);   // A row of this lane
  const int col = blockIdx.x * 64 + wn * 32 + (lane & 31);   // B column of this lane
  const int kh = 8 * (lane >> 5);
  const int k_begin = blockIdx.z * g.kc;
  const int k_end = min(g.K, k_begin + g.kc);
  f32x16 acc = {};
  const bool rok = row < g.M, cok = col < g.N;
  for (int k0 = k_begin; k0 < k_end; k0 += 16) {
    bf16x8 a, b;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int k = k0 + kh + i;
      const bool kok = k < k_end;
      a[i] = (rok && kok) ? (TA ? g.A[(int64_t)k * g.lda + row] : g.A[(int64_t)row * g.lda + k])
                          : (__bf16)0.f;
      b[i] = (cok && kok) ? (TB ? g.B[(int64_t)col * g.ldb + k] : g.B[(int64_t)k * g.ldb + col])
                          : (__bf16)0.f;
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
  }
  float* C = g.C ? g.C + (int64_t)blockIdx.z * g.c_zstride : nullptr;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int orow = blockIdx.y * 64 + wm * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    const int ocol = col;
    if (orow >= g.M || ocol >= g.N) continue;
    float v = acc[r];
    if (g.bias) v += g.bias[ocol];
    if (g.relu) v = fmaxf(v, 0.f);
    if (g.U) g.U[(int64_t)orow * g.ldu + ocol] = v;
    if (g.res) v += g.res[(int64_t)orow * g.ldr + ocol];
    if (C) {
      float* p = C + (int64_t)orow * g.ldc + ocol;
      if (g.accumulate) v += *p;
      *p = v;
    }
    if (g.Cb) g.Cb[(int64_t)orow * g.ldcb + ocol] = (__bf16)v;
  }
}

// Epilogue of the A B^T product with its loads issued ahead: the bias value
// of the lane's column once, and the residual and accumulated values of all
// 16 output elements before the first store, so that their latencies overlap
// (per element, each load would be waited for in turn).  Rows and columns
// past the matrix read row / column 0 and are not stored.  The arithmetic is
// that of `gemm_kernel`'s epilogue, in the same order.
struct EpiIn {
  int rows[16];  // output row of element r, -1 when outside the matrix
  float bv, rv[16], cv[16];
};

__device__ __forceinline__ void epi_load(const GemmArgs& g, int orow0, int col, int lane,
                                         EpiIn& in) {
  const bool cok = col < g.N;
  const int ccol = cok ? col : 0;
  in.bv = 0.f;
  if (g.bias) in.bv = g.bias[ccol];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int orow = orow0 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    in.rows[r] = (cok && orow < g.M) ? orow : -1;
  }
  if (g.res) {
#pragma unroll
    for (int r = 0; r < 16; ++r)
      in.rv[r] = g.res[(int64_t)(in.rows[r] < 0 ? 0 : in.rows[r]) * g.ldr + ccol];
  }
  if (g.C && g.accumulate) {
#pragma unroll
    for (int r = 0; r < 16; ++r)
      in.cv[r] = g.C[(int64_t)(in.rows[r] < 0 ? 0 : in.rows[r]) * g.ldc + ccol];
  }
}

__device__ __forceinline__ void epi_store(const GemmArgs& g, const f32x16& acc, int col,
                                          const EpiIn& in) {
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    float v = acc[r];
    if (g.bias) v += in.bv;
    if (g.relu) v = fmaxf(v, 0.f);
    const bool ok = in.rows[r] >= 0;
    if (g.U && ok) g.U[(int64_t)in.rows[r] * g.ldu + col] = v;
    if (g.res) v += in.rv[r];
    if (g.C) {
      if (g.accumulate) v += in.cv[r];
      if (ok) g.C[(int64_t)in.rows[r] * g.ldc + col] = v;
    }
    if (g.Cb && ok) g.Cb[(int64_t)in.rows[r] * g.ldcb + col] = (__bf16)v;
  }
}

// k-steps whose fragments are loaded before their MFMAs
constexpr int kGemmChunk = 8;

// C = A B^T-stored: A [M x K] row-major, B given as B^T [N x K] row-major
// (k contiguous for both), so every fragment is one 16-byte load from
// global memory / L2 and no LDS staging is needed.  The forward layers use
// transposed BF16 copies of the weights; the input gradients use the
// weights themselves.  Requires lda, ldb % 8 == 0 and K % 16 == 0.
// KS = K / 16 when the depth is one of the network's (1, 8 or 16 steps):
// the fragments of kGemmChunk k-steps are loaded before their MFMAs, so
// their latencies overlap; KS = 0 walks a run-time K one step at a time.
// Rows (columns) past M (N) load row (column) 0 and their outputs are not
// stored, so no load is predicated.  The MFMAs run in k order either way.
template <int KS>
__global__ __launch_bounds__(256) void gemm_nt_kernel(GemmArgs g) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wm = w & 1, wn = w >> 1;
  const int row = blockIdx.y * 64 + wm * 32 + (lane & 31);
  const int col = blockIdx.x * 64 + wn * 32 + (lane & 31);
  const int kh = 8 * (lane >> 5);
  const __bf16* ap = g.A + (int64_t)(row < g.M ? row : 0) * g.lda + kh;
  const __bf16* bp = g.B + (int64_t)(col < g.N ? col : 0) * g.ldb + kh;
  f32x16 acc = {};
  if constexpr (KS > 0) {
    constexpr int kChunk = KS < kGemmChunk ? KS : kGemmChunk;
#pragma unroll
    for (int c = 0; c < KS; c += kChunk) {
      bf16x8 a[kChunk], b[kChunk];
#pragma unroll
      for (int s = 0; s < kChunk; ++s) {
        a[s] = *(const bf16x8*)(ap + 16 * (c + s));
        b[s] = *(const bf16x8*)(bp + 16 * (c + s));
      }
#pragma unroll
      for (int s = 0; s < kChunk; ++s)
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[s], b[s], acc, 0, 0, 0);
    }
  } else {
    for (int k0 = 0; k0 < g.K; k0 += 16) {
      const bf16x8 a = *(const bf16x8*)(ap + k0);
      const bf16x8 b = *(const bf16x8*)(bp + k0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
  }
  EpiIn in;
  epi_load(g, blockIdx.y * 64 + wm * 32, col, lane, in);
  epi_store(g, acc, col, in);
}

#ifndef SPH_GEMM_LDS
#define SPH_GEMM_LDS 1
#endif

// The same product with both 64 x K tiles staged in LDS: the block loads A's
// 64 rows and B^T's 64 rows with 16-byte loads in which consecutive threads
// read consecutive bytes of a row (whole cache lines), then each wave reads
// its MFMA fragments from LDS.  Per-lane fragment loads from global memory
// touch 32 cache lines per instruction and kept the texture units 73 %
// busy.  The row pitch is K + 8 elements, so the eight 16-byte reads of a
// pass start 16 bytes apart modulo 128 (no bank conflicts).  Rows (columns)
// past M (N) load row (column) 0 and are not stored.  The MFMAs run in k
// order and the epilogue is `epi_load` / `epi_store`: the results are those
// of gemm_nt_kernel bit for bit.  K = 16 KS (128 or 256: 35 or 68 KB of LDS).
template <int KS>
__global__ __launch_bounds__(256) void gemm_nt_lds_kernel(GemmArgs g) {
  constexpr int K = 16 * KS;
  constexpr int kPitch = K + 8;
  constexpr int kPerRow = K / 8;  // 16-byte chunks per row
  __shared__ __align__(16) __bf16 sA[64 * kPitch];
  __shared__ __align__(16) __bf16 sB[64 * kPitch];
  const int tid = threadIdx.x;
  const int row0 = blockIdx.y * 64, col0 = blockIdx.x * 64;
#pragma unroll
  for (int i = 0; i < 64 * kPerRow / 256; ++i) {
    const int c = tid + 256 * i;
    const int r = c / kPerRow, kc = c % kPerRow;
    const int ar = row0 + r < g.M ? row0 + r : 0;
    const int bc = col0 + r < g.N ? col0 + r : 0;
    const bf16x8 va = *(const bf16x8*)(g.A + (int64_t)ar * g.lda + kc * 8);
    const bf16x8 vb = *(const bf16x8*)(g.B + (int64_t)bc * g.ldb + kc * 8);
    *(bf16x8*)(sA + r * kPitch + kc * 8) = va;
    *(bf16x8*)(sB + r * kPitch + kc * 8) = vb;
  }
  __syncthreads();
  const int lane = tid & 63, w = tid >> 6;
  const int wm = w & 1, wn = w >> 1;
  const int kh = 8 * (lane >> 5);
  const __bf16* ap = sA + (wm * 32 + (lane & 31)) * kPitch + kh;
  const __bf16* bp = sB + (wn * 32 + (lane & 31)) * kPitch + kh;
  f32x16 acc = {};
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    const bf16x8 a = *(const bf16x8*)(ap + 16 * s);
    const bf16x8 b = *(const bf16x8*)(bp + 16 * s);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
  }
  const int col = col0 + wn * 32 + (lane & 31);
  EpiIn in;
  epi_load(g, row0 + wm * 32, col, lane, in);
  epi_store(g, acc, col, in);
}

// wt[n*K + k] = bf16(w[k*N + n]) for one K x N weight matrix.
__global__ __launch_bounds__(256) void transpose_bf16_kernel(int K, int N,
                                                             const float* __restrict__ w,
                                                             __bf16* __restrict__ wt) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= K * N) return;
  const int k = t / N, n = t % N;
  wt[(int64_t)n * K + k] = (__bf16)w[t];
}

// Sum of nslab slabs [M x N] (ld) into out (accumulate when acc != 0).
__global__ __launch_bounds__(256) void slab_sum_kernel(const float* __restrict__ part, int nslab,
                                                       int64_t zstride, int M, int N,
                                                       int64_t ld, float* __restrict__ out,
                                                       int acc) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (int64_t)M * N) return;
  const int m = (int)(i / N), n = (int)(i % N);
  float s = 0.f;
  for (int z = 0; z < nslab; ++z) s += part[z * zstride + (int64_t)m * ld + n];
  if (acc) s += out[i];
  out[i] = s;
}

// ---------------------------------------------------------------------------
// features
// ---------------------------------------------------------------------------
struct FeatArgs {
  int n;
  const float4* pos4;
  const float4* vel4;
  const float* rho;
  const int* btype;
  const int64_t* off;
  const int64_t* nbr;
  float rho0, co, gamma, B, u_scale, g, rho_scale, alpha, eta;
  int dim;
};

__device__ __forceinline__ float tait_p(const FeatArgs& f, float rho) {
  return f.B * (powf(rho / f.rho0, f.gamma) - 1.f);
}

// Cubic spline W and grad factor of the 2D/3D kernel (kernels.pyx:367-412),
// h = the pair's mean smoothing length.
__device__ __forceinline__ void spline(float r, float h, int dim, float& w, float& gf) {
  const float q = r / h;
  const float fac = dim == 2 ? 10.f / (7.f * 3.14159265358979f * h * h)
                             : 1.f / (3.14159265358979f * h * h * h);
  const float t = 2.f - q;
  if (q > 2.f) {
    w = 0.f;
    gf = 0.f;
  } else if (q > 1.f) {
    w = fac * 0.25f * t * t * t;
    gf = fac * (-0.75f * t * t / (h * r));
  } else {
    w = fac * (1.f - 1.5f * q * q * (1.f - 0.5f * q));
    gf = q > 1e-14f ? fac * 3.f * (0.75f * q - 1.f) / (h * h) : 0.f;
  }
}

// Node features: u/U, v/U, (rho/rho0 - 1) S, p/B, boundary flag, 1, and the
// node's SPH pressure and artificial-viscosity accelerations / g (the sums
// over its edges, from edge_feat_kernel).
__global__ __launch_bounds__(256) void node_feat_kernel(FeatArgs f, const float* __restrict__ acc4,
                                                        __bf16* __restrict__ X) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= f.n) return;
  const float4 v = f.vel4[i];
  const float rho = f.rho[i];
  float x[kIn] = {};
  x[0] = v.x / f.u_scale;
  x[1] = v.y / f.u_scale;
  x[2] = (rho / f.rho0 - 1.f) * f.rho_scale;
  x[3] = tait_p(f, rho) / f.B;
  x[4] = f.btype[i] ? 1.f : 0.f;
  x[5] = 1.f;
  // the SPH accelerations enter the network scaled down (they reach 100s of
  // g at a summation-density free surface); the decoder's skip path adds
  // them unscaled
#pragma unroll
  for (int c = 0; c < 4; ++c) x[6 + c] = 0.01f * acc4[(int64_t)i * 4 + c];
#pragma unroll
  for (int k = 0; k < kIn; ++k) X[(int64_t)i * kIn + k] = (__bf16)x[k];
}

// Edge features of edge (i <- j), one lane per edge.  Antisymmetric parts
// change sign and the rho_i / rho_j slots swap for the reverse edge
// (see reverse_feat), which the backward pass relies on.
// kFeatLanes lanes per node: a node has about 21 edges in 2D, so a whole
// wave per node would leave two thirds of its lanes idle (the features
// launch 0.077 -> 0.048 ms on the 64k-particle dam break)
constexpr int kFeatLanes = 16;
constexpr int kFeatNodes = 256 / kFeatLanes;  // nodes per 256-thread block

__global__ __launch_bounds__(256) void edge_feat_kernel(FeatArgs f, float* __restrict__ E,
                                                        float* __restrict__ acc4) {
  const int i = blockIdx.x * kFeatNodes + (int)(threadIdx.x / kFeatLanes);
  const int lane = threadIdx.x % kFeatLanes;
  // every lane stays for the group reductions below (no early return)
  const bool node = i < f.n;
  const int64_t e0 = node ? f.off[i] : 0, e1 = node ? f.off[i + 1] : 0;
  const float4 pi = f.pos4[node ? i : 0], vi = f.vel4[node ? i : 0];
  const float ri = f.rho[node ? i : 0];
  const float pri = tait_p(f, ri);
  float sum[4] = {0.f, 0.f, 0.f, 0.f};
  for (int64_t e = e0 + lane; e < e1; e += kFeatLanes) {
    const int j = (int)f.nbr[e];
    const float4 pj = f.pos4[j], vj = f.vel4[j];
    const float rj = f.rho[j];
    const float prj = tait_p(f, rj);
    const float h = 0.5f * (pi.w + pj.w);
    const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
    const float r = sqrtf(dx * dx + dy * dy + dz * dz);
    float w, gf;
    spline(r, h, f.dim, w, gf);
    const float du = vi.x - vj.x, dv = vi.y - vj.y, dw = vi.z - vj.z;
    const float mbar = 0.5f * (vi.w + vj.w);
    const float pt = -mbar * (pri / (ri * ri) + prj / (rj * rj)) * gf / f.g;
    const float dot = du * dx + dv * dy + dw * dz;
    float visc = 0.f;
    if (dot < 0.f) {
      const float cbar = 0.5f * f.co * (powf(ri / f.rho0, 0.5f * (f.gamma - 1.f)) +
                                        powf(rj / f.rho0, 0.5f * (f.gamma - 1.f)));
      const float mu = h * dot / (r * r + f.eta * f.eta * h * h);
      visc = -mbar * (-f.alpha * cbar * mu) / (0.5f * (ri + rj)) * gf / f.g;
    }
    sum[0] += pt * dx;
    sum[1] += pt * dy;
    sum[2] += visc * dx;
    sum[3] += visc * dy;
    // a pair's pressure term is of order (p/rho)/h ~ g H / h; scale to O(1)
    const float ps = 2.f * f.g * h / (f.u_scale * f.u_scale);
    float x[kE] = {};
    x[0] = dx / h;
    x[1] = dy / h;
    x[2] = r / h;
    x[3] = du / f.u_scale;
    x[4] = dv / f.u_scale;
    x[5] = (ri / f.rho0 - 1.f) * f.rho_scale;
    x[6] = (rj / f.rho0 - 1.f) * f.rho_scale;
    x[7] = w * h * h;
    x[8] = pt * dx * ps;
    x[9] = pt * dy * ps;
    x[10] = visc * dx * ps;
    x[11] = visc * dy * ps;
    x[12] = 1.f;
#pragma unroll
    for (int k = 0; k < kE; ++k) E[e * kE + k] = x[k];
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    float v = sum[c];
    for (int off = kFeatLanes / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kFeatLanes);
    if (lane == 0 && node) acc4[(int64_t)i * 4 + c] = v;
  }
}

// Features of edge (j <- i) from those of (i <- j).
__device__ __forceinline__ float reverse_feat(const float* x, int k) {
  switch (k) {
    case 0: case 1: case 3: case 4: case 8: case 9: case 10: case 11: return -x[k];
    case 5: return x[6];
    case 6: return x[5];
    default: return x[k];
  }
}

// ---------------------------------------------------------------------------
// message passing (one wave per destination node, channels lane, lane + 64)
// ---------------------------------------------------------------------------
struct EdgeArgs {
  int n;
  const int64_t* off;
  const int64_t* nbr;
  const float* E;       // [edges x kE]
  const float* PQ;      // [n x 2H]: P | Q
  const float* Wc;      // [kE x H]
  const float* be;      // [H]
  __bf16* A;            // forward: aggregate written at A + i*lda (BF16)
  int64_t lda;
  // backward
  const float* dA;      // [n x ldda] gradient of the aggregate
  int64_t ldda;
  float* dPQ;           // [n x 2H]
  __bf16* dPQb;         // BF16 copy
  float* part;          // [nwaves x (kE + 1) x H] partial dWc, dbe
};

__global__ __launch_bounds__(256) void edge_fwd_kernel(EdgeArgs a) {
  const int i = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  if (i >= a.n) return;
  float wc0[kE], wc1[kE];
#pragma unroll
  for (int k = 0; k < kE; ++k) {
    wc0[k] = a.Wc[k * kH + lane];
    wc1[k] = a.Wc[k * kH + lane + 64];
  }
  const float p0 = a.PQ[(int64_t)i * 2 * kH + lane] + a.be[lane];
  const float p1 = a.PQ[(int64_t)i * 2 * kH + lane + 64] + a.be[lane + 64];
  float s0 = 0.f, s1 = 0.f;
  const int64_t e0 = a.off[i], e1 = a.off[i + 1];
  // four edges per iteration: their neighbour rows are in flight together
  constexpr int U = 4;
  for (int64_t e = e0; e < e1; e += U) {
    float q0[U], q1[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t eu = e + u < e1 ? e + u : e0;
      const int j = (int)a.nbr[eu];
      q0[u] = a.PQ[(int64_t)j * 2 * kH + kH + lane];
      q1[u] = a.PQ[(int64_t)j * 2 * kH + kH + lane + 64];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (e + u >= e1) break;
      const float* x = a.E + (e + u) * kE;
      float z0 = p0 + q0[u], z1 = p1 + q1[u];
#pragma unroll
      for (int k = 0; k < kE; ++k) {
        const float xk = x[k];
        z0 += xk * wc0[k];
        z1 += xk * wc1[k];
      }
      s0 += fmaxf(z0, 0.f);
      s1 += fmaxf(z1, 0.f);
    }
  }
  a.A[(int64_t)i * a.lda + lane] = (__bf16)s0;
  a.A[(int64_t)i * a.lda + lane + 64] = (__bf16)s1;
}

#ifndef SPH_EDGE_MFMA
#define SPH_EDGE_MFMA 1
#endif

// x = hi + lo with hi = bf16(x) and lo = bf16(x - hi): hi*w_hi + hi*w_lo +
// lo*w_hi reproduces an FP32 product to ~2^-16 relative.
__device__ __forceinline__ void split_bf16(float x, __bf16& hi, __bf16& lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
}

// The aggregation A_i = sum_j ReLU(P_i + b + Q_j + e_ij W_c) on the matrix
// cores.  A wave takes the edges of node i 32 at a time: the 32 x 16 block
// of edge features times W_c (16 x 128) is four v_mfma_f32_32x32x16_bf16
// tiles of 32 hidden units, each as three products of the features' and
// W_c's BF16 high and low parts (the FP32 products to ~2^-16).  In the
// product's layout, lane l holds hidden unit 32t + (l & 31) of the 16 edge
// rows 8(r / 4) + 4(l / 32) + (r % 4); it adds P_i + b and Q_j of each row,
// applies the ReLU and sums its rows, and the two half-waves' row sums are
// added at the end.  Rows past the node's last edge read that edge and are
// masked out of the sums.  The block splits W_c into its BF16 fragments in
// LDS once (8 KB, in the lanes' fragment order, so each read is one
// conflict-free ds_read_b128); then each wave takes nodes w, w + nwaves, ...
#ifndef SPH_EDGE_WAVES
#define SPH_EDGE_WAVES 3  // 154 VGPRs: three waves per SIMD without spills
#endif
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SPH_EDGE_WAVES)))
void edge_fwd_mfma_kernel(EdgeArgs a) {
  __shared__ bf16x8 wfrag[2 * 4 * 64];  // [hi, lo][tile][lane]
  for (int f = threadIdx.x; f < 2 * 4 * 64; f += 256) {
    const int l = f & 63, t = (f >> 6) & 3, lo = f >> 8;
    const int c = 32 * t + (l & 31), k0 = 8 * (l >> 5);
    bf16x8 v;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      __bf16 h, l2;
      split_bf16(a.Wc[(k0 + s) * kH + c], h, l2);
      v[s] = lo ? l2 : h;
    }
    wfrag[f] = v;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
  const int nwaves = gridDim.x * 4;
  const int m = lane & 31, kh = 8 * (lane >> 5);
  const int mr0 = 4 * (lane >> 5);  // edge row of r is mr0 + 8(r / 4) + r % 4
  float be[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) be[t] = a.be[32 * t + m];
  for (int i = wave; i < a.n; i += nwaves) {
    const int64_t e0 = a.off[i], e1 = a.off[i + 1];
    float pb[4], sum[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      pb[t] = a.PQ[(int64_t)i * 2 * kH + 32 * t + m] + be[t];
      sum[t] = 0.f;
    }
    for (int64_t eb = e0; eb < e1; eb += 32) {
      const int64_t em = eb + m < e1 ? eb + m : e1 - 1;  // feature row of this lane
      const float4 x0 = *(const float4*)(a.E + em * kE + kh);
      const float4 x1 = *(const float4*)(a.E + em * kE + kh + 4);
      const int jm = (int)a.nbr[em];
      const int lim = (int)(e1 - eb);  // rows below lim are edges of node i
      const float xs[8] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
      bf16x8 ah, al;
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        __bf16 h, l;
        split_bf16(xs[s], h, l);
        ah[s] = h;
        al[s] = l;
      }
      int qoff[16];  // Q_j row of edge row r, in floats
      float okf[16];  // 1 for the edges of node i, 0 for the rows past them
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int mr = mr0 + 8 * (r >> 2) + (r & 3);
        const int j = __shfl(jm, mr);
        qoff[r] = j * 2 * kH + kH + m;
        okf[r] = mr < lim ? 1.f : 0.f;
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        float q[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) q[r] = a.PQ[(int64_t)qoff[r] + 32 * t];
        const bf16x8 bh = wfrag[t * 64 + lane], bl = wfrag[256 + t * 64 + lane];
        f32x16 acc = {};
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          // a masked row's z is finite (it re-reads a valid edge): x 0 adds 0
          const float z = acc[r] + pb[t] + q[r];
          sum[t] = fmaf(fmaxf(z, 0.f), okf[r], sum[t]);
        }
        // one tile's 16 neighbour-row loads in flight at a time: hoisting all
        // 64 spills at the occupancy set above
        __builtin_amdgcn_sched_barrier(0);
      }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const float tot = sum[t] + __shfl_xor(sum[t], 32);
      if (lane < 32) a.A[(int64_t)i * a.lda + 32 * t + m] = (__bf16)tot;
    }
  }
}

// Persistent: each wave walks nodes i = wave, wave + nwaves, ...  and keeps
// its dWc / dbe partial sums in registers.
__global__ __launch_bounds__(256) void edge_bwd_kernel(EdgeArgs a) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
  const int nwaves = gridDim.x * 4;
  float wc0[kE], wc1[kE], g0[kE], g1[kE];
#pragma unroll
  for (int k = 0; k < kE; ++k) {
    wc0[k] = a.Wc[k * kH + lane];
    wc1[k] = a.Wc[k * kH + lane + 64];
    g0[k] = 0.f;
    g1[k] = 0.f;
  }
  const float be0 = a.be[lane], be1 = a.be[lane + 64];
  float gb0 = 0.f, gb1 = 0.f;
  for (int i = wave; i < a.n; i += nwaves) {
    const float* PQi = a.PQ + (int64_t)i * 2 * kH;
    const float pi0 = PQi[lane] + be0, pi1 = PQi[lane + 64] + be1;
    const float qi0 = PQi[kH + lane], qi1 = PQi[kH + lane + 64];
    const float dai0 = a.dA[(int64_t)i * a.ldda + lane];
    const float dai1 = a.dA[(int64_t)i * a.ldda + lane + 64];
    float dp0 = 0.f, dp1 = 0.f, dq0 = 0.f, dq1 = 0.f;
    const int64_t e0 = a.off[i], e1 = a.off[i + 1];
    for (int64_t e = e0; e < e1; ++e) {
      const int j = (int)a.nbr[e];
      const float* x = a.E + e * kE;
      const float* PQj = a.PQ + (int64_t)j * 2 * kH;
      // edge (i <- j): s = P_i + Q_j + e_ij Wc + b
      float s0 = pi0 + PQj[kH + lane], s1 = pi1 + PQj[kH + lane + 64];
      // edge (j <- i): s' = P_j + Q_i + e_ji Wc + b
      float t0 = PQj[lane] + be0 + qi0, t1 = PQj[lane + 64] + be1 + qi1;
      float xk[kE];
#pragma unroll
      for (int k = 0; k < kE; ++k) xk[k] = x[k];
#pragma unroll
      for (int k = 0; k < kE; ++k) {
        s0 += xk[k] * wc0[k];
        s1 += xk[k] * wc1[k];
        const float rk = reverse_feat(xk, k);
        t0 += rk * wc0[k];
        t1 += rk * wc1[k];
      }
      const float ds0 = s0 > 0.f ? dai0 : 0.f, ds1 = s1 > 0.f ? dai1 : 0.f;
      dp0 += ds0;
      dp1 += ds1;
#pragma unroll
      for (int k = 0; k < kE; ++k) {
        g0[k] += xk[k] * ds0;
        g1[k] += xk[k] * ds1;
      }
      gb0 += ds0;
      gb1 += ds1;
      const float daj0 = a.dA[(int64_t)j * a.ldda + lane];
      const float daj1 = a.dA[(int64_t)j * a.ldda + lane + 64];
      dq0 += t0 > 0.f ? daj0 : 0.f;
      dq1 += t1 > 0.f ? daj1 : 0.f;
    }
    float* d = a.dPQ + (int64_t)i * 2 * kH;
    d[lane] = dp0;
    d[lane + 64] = dp1;
    d[kH + lane] = dq0;
    d[kH + lane + 64] = dq1;
    __bf16* db = a.dPQb + (int64_t)i * 2 * kH;
    db[lane] = (__bf16)dp0;
    db[lane + 64] = (__bf16)dp1;
    db[kH + lane] = (__bf16)dq0;
    db[kH + lane + 64] = (__bf16)dq1;
  }
  float* p = a.part + (int64_t)wave * (kE + 1) * kH;
#pragma unroll
  for (int k = 0; k < kE; ++k) {
    p[k * kH + lane] = g0[k];
    p[k * kH + lane + 64] = g1[k];
  }
  p[kE * kH + lane] = gb0;
  p[kE * kH + lane + 64] = gb1;
}

// ---------------------------------------------------------------------------
// decoder, loss, elementwise pieces
// ---------------------------------------------------------------------------
// Y = H W_out + b_out + the node's SPH pressure and viscosity
// accelerations / g (FP32 skip path; one thread per node).
__global__ __launch_bounds__(256) void decode_kernel(int n, const float* __restrict__ H,
                                                     const float* __restrict__ W,
                                                     const float* __restrict__ b,
                                                     const float* __restrict__ acc4,
                                                     float* __restrict__ Y) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float* a4 = acc4 + (int64_t)i * 4;
  float y0 = b[0] + a4[0] + a4[2], y1 = b[1] + a4[1] + a4[3];
  const float* h = H + (int64_t)i * kH;
  for (int c = 0; c < kH; ++c) {
    const float v = h[c];
    y0 += v * W[c * kOut];
    y1 += v * W[c * kOut + 1];
  }
  Y[2 * (int64_t)i] = y0;
  Y[2 * (int64_t)i + 1] = y1;
}

// dH = dY W_out^T; per-block (64 nodes) partials of dW_out = H^T dY and db_out.
__global__ __launch_bounds__(256) void decode_bwd_kernel(int n, const float* __restrict__ H,
                                                         const float* __restrict__ W,
                                                         const float* __restrict__ dY,
                                                         float* __restrict__ dH,
                                                         float* __restrict__ part) {
  __shared__ float sw[2][256];
  __shared__ float sb[2][64];
  const int node0 = blockIdx.x * 64;
  const int t = threadIdx.x;
  const int c = t & (kH - 1);
  const int half = t >> 7;
  const float w0 = W[c * kOut], w1 = W[c * kOut + 1];
  float gw0 = 0.f, gw1 = 0.f;
  for (int k = half; k < 64; k += 2) {
    const int i = node0 + k;
    if (i >= n) break;
    const float d0 = dY[2 * (int64_t)i], d1 = dY[2 * (int64_t)i + 1];
    dH[(int64_t)i * kH + c] = d0 * w0 + d1 * w1;
    const float h = H[(int64_t)i * kH + c];
    gw0 += h * d0;
    gw1 += h * d1;
  }
  sw[0][t] = gw0;
  sw[1][t] = gw1;
  if (t < 64) {
    const int i = node0 + t;
    sb[0][t] = i < n ? dY[2 * (int64_t)i] : 0.f;
    sb[1][t] = i < n ? dY[2 * (int64_t)i + 1] : 0.f;
  }
  __syncthreads();
  float* p = part + (int64_t)blockIdx.x * (kH * kOut + kOut);
  if (t < kH) {
    p[t * kOut] = sw[0][t] + sw[0][t + kH];
    p[t * kOut + 1] = sw[1][t] + sw[1][t + kH];
  }
  if (t == 0) {
    float s0 = 0.f, s1 = 0.f;
    for (int k = 0; k < 64; ++k) {
      s0 += sb[0][k];
      s1 += sb[1][k];
    }
    p[kH * kOut] = s0;
    p[kH * kOut + 1] = s1;
  }
}

// MSE over fluid nodes of (Y - T): per-block partial losses and dY.
__global__ __launch_bounds__(256) void loss_kernel(int n, const int* __restrict__ btype,
                                                   const float* __restrict__ Y,
                                                   const float* __restrict__ T, float inv_nf,
                                                   float* __restrict__ dY,
                                                   float* __restrict__ part) {
  __shared__ float red[256];
  const int i = blockIdx.x * 256 + threadIdx.x;
  float l = 0.f;
  if (i < n) {
    const bool fl = btype[i] == 0;
    const float e0 = fl ? Y[2 * (int64_t)i] - T[2 * (int64_t)i] : 0.f;
    const float e1 = fl ? Y[2 * (int64_t)i + 1] - T[2 * (int64_t)i + 1] : 0.f;
    l = (e0 * e0 + e1 * e1) * inv_nf;
    dY[2 * (int64_t)i] = 2.f * e0 * inv_nf;
    dY[2 * (int64_t)i + 1] = 2.f * e1 * inv_nf;
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// out = mask(act > 0) * d  (FP32 and BF16 copies); ld strides per array.
__global__ __launch_bounds__(256) void relu_bwd_kernel(int64_t count, int cols,
                                                       const float* __restrict__ act,
                                                       int64_t lda_, const float* __restrict__ d,
                                                       int64_t ldd, float* __restrict__ out,
                                                       __bf16* __restrict__ outb) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= count) return;
  const int64_t r = t / cols;
  const int c = (int)(t % cols);
  const float v = act[r * lda_ + c] > 0.f ? d[r * ldd + c] : 0.f;
  out[t] = v;
  outb[t] = (__bf16)v;
}

// dst[r][c] = a[r][c] + b[r * ldb + c]
__global__ __launch_bounds__(256) void add_kernel(int64_t count, int cols,
                                                  const float* __restrict__ a,
                                                  const float* __restrict__ b, int64_t ldb,
                                                  float* __restrict__ dst) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= count) return;
  const int64_t r = t / cols;
  const int c = (int)(t % cols);
  dst[t] = a[t] + b[r * ldb + c];
}

// Column sums of x [rows x cols] (ld) per block of 256 rows -> part.
__global__ __launch_bounds__(256) void colsum_part_kernel(int rows, int cols,
                                                          const float* __restrict__ x,
                                                          int64_t ld, float* __restrict__ part) {
  const int r0 = blockIdx.x * 256;
  for (int c = threadIdx.x; c < cols; c += 256) {
    float s = 0.f;
    for (int r = r0; r < min(rows, r0 + 256); ++r) s += x[(int64_t)r * ld + c];
    part[(int64_t)blockIdx.x * cols + c] = s;
  }
}

__global__ __launch_bounds__(256) void adam_kernel(int64_t n, float* __restrict__ w,
                                                   __bf16* __restrict__ wb,
                                                   const float* __restrict__ g,
                                                   float* __restrict__ m, float* __restrict__ v,
                                                   float lr, float b1, float b2, float eps,
                                                   float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  wb[i] = (__bf16)wi;
}

__global__ __launch_bounds__(256) void to_bf16_kernel(int64_t n, const float* __restrict__ w,
                                                      __bf16* __restrict__ wb) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) wb[i] = (__bf16)w[i];
}

__global__ __launch_bounds__(256) void fill_f32_kernel(int64_t n, float* __restrict__ p,
                                                       float v) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

// Surrogate update of fluid nodes (symplectic Euler): u += dt a g, x += dt u.
__global__ __launch_bounds__(256) void integrate_kernel(int64_t n, const int* __restrict__ btype,
                                                        const float* __restrict__ Y, float ag,
                                                        float dt, float* __restrict__ x,
                                                        float* __restrict__ y,
                                                        float* __restrict__ u,
                                                        float* __restrict__ v) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n || btype[i] != 0) return;
  const float u1 = u[i] + dt * ag * Y[2 * i];
  const float v1 = v[i] + dt * ag * Y[2 * i + 1];
  u[i] = u1;
  v[i] = v1;
  x[i] = x[i] + dt * u1;
  y[i] = y[i] + dt * v1;
}

// ---------------------------------------------------------------------------
// model object
// ---------------------------------------------------------------------------
struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
};

struct Model {
  Layout L;
  float* w = nullptr;     // FP32 master
  __bf16* wb = nullptr;   // BF16 copy
  __bf16* wbt = nullptr;  // BF16 copies of the GEMM weights, transposed
  float* g = nullptr;     // gradients
  float* m = nullptr;     // Adam moments
  float* v = nullptr;
  // activations and scratch, grown on demand
  // optional per-category launch timing of the forward pass: features,
  // GEMMs, edge aggregation, decoder
  int timing = 0;
  double t_ms[4] = {0, 0, 0, 0};
  int64_t t_calls = 0;
  std::vector<hipEvent_t> ev;
  Buf X, E, N4, H[kR + 1], HA[kR], PQ[kR], U[kR], dH, dZ, dZb, dHA, dPQ, dPQb, part, lossp;
  int64_t n_act = -1, e_act = -1;
  bool acts = false;  // the last forward pass stored what the backward pass reads
};

// Events around each launch of a timed forward pass.
struct Timer {
  Model* M;
  hipStream_t st;
  std::vector<std::pair<int, int>> marks;  // (category, event index of the start)
  int used = 0;
  Timer(Model* m, hipStream_t s) : M(m), st(s) {}
  void begin(int cat) {
    if (!M->timing) return;
    while ((int)M->ev.size() < used + 2) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) {
        M->timing = 0;
        return;
      }
      M->ev.push_back(e);
    }
    (void)hipEventRecord(M->ev[used], st);
    marks.push_back({cat, used});
    used += 2;
  }
  void end() {
    if (!M->timing || marks.empty()) return;
    (void)hipEventRecord(M->ev[marks.back().second + 1], st);
  }
  void collect() {
    if (!M->timing || marks.empty()) return;
    (void)hipEventSynchronize(M->ev[marks.back().second + 1]);
    for (auto& m : marks) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, M->ev[m.second], M->ev[m.second + 1]) == hipSuccess)
        M->t_ms[m.first] += ms;
    }
    M->t_calls += 1;
  }
};

static int32_t ensure(Buf& b, size_t bytes) {
  if (b.bytes >= bytes && b.p) return SPH_OK;
  if (b.p) {
    // earlier launches may still read the old buffer
    (void)hipDeviceSynchronize();
    (void)hipFree(b.p);
  }
  b.p = nullptr;
  b.bytes = 0;
  hipError_t e = hipMalloc(&b.p, bytes < 256 ? 256 : bytes);
  if (e != hipSuccess) {
    set_error("sph_mlp: allocation of %zu bytes: %s", bytes, hipGetErrorString(e));
    return (int32_t)e;
  }
  b.bytes = bytes;
  return SPH_OK;
}

static int nblk(int64_t n) { return (int)((n + 255) / 256); }

// Persistent edge aggregation: up to SPH_EDGE_WAVES waves per SIMD of the
// 256 CUs, each wave then walking nodes with stride gridDim.x * 4.
static int edge_fwd_blocks(int64_t n) {
  const int64_t waves = std::min<int64_t>(n, 256 * 4 * SPH_EDGE_WAVES);
  return (int)std::max<int64_t>(1, (waves + 3) / 4);
}

static int32_t gemm(hipStream_t st, bool ta, bool tb, GemmArgs a, int splits = 1) {
  if (!ta && tb && splits == 1 && a.K % 16 == 0 && a.lda % 8 == 0 && a.ldb % 8 == 0 &&
      ((uintptr_t)a.A & 15) == 0 && ((uintptr_t)a.B & 15) == 0) {
    a.kc = a.K;
    dim3 grid((a.N + 63) / 64, (a.M + 63) / 64, 1);
    if (a.K == 16) hipLaunchKernelGGL(gemm_nt_kernel<1>, grid, dim3(256), 0, st, a);
    else if (SPH_GEMM_LDS && a.K == 128)
      hipLaunchKernelGGL(gemm_nt_lds_kernel<8>, grid, dim3(256), 0, st, a);
    else if (SPH_GEMM_LDS && a.K == 256)
      hipLaunchKernelGGL(gemm_nt_lds_kernel<16>, grid, dim3(256), 0, st, a);
    else if (a.K == 128) hipLaunchKernelGGL(gemm_nt_kernel<8>, grid, dim3(256), 0, st, a);
    else if (a.K == 256) hipLaunchKernelGGL(gemm_nt_kernel<16>, grid, dim3(256), 0, st, a);
    else hipLaunchKernelGGL(gemm_nt_kernel<0>, grid, dim3(256), 0, st, a);
    return launch_status("sph_mlp/gemm_nt");
  }
  a.kc = (a.K + splits - 1) / splits;
  a.kc = (a.kc + 15) / 16 * 16;
  const int z = (a.K + a.kc - 1) / a.kc;
  dim3 grid((a.N + 63) / 64, (a.M + 63) / 64, z < 1 ? 1 : z);
  if (!ta && !tb) hipLaunchKernelGGL((gemm_kernel<false, false>), grid, dim3(256), 0, st, a);
  else if (!ta && tb) hipLaunchKernelGGL((gemm_kernel<false, true>), grid, dim3(256), 0, st, a);
  else if (ta && !tb) hipLaunchKernelGGL((gemm_kernel<true, false>), grid, dim3(256), 0, st, a);
  else hipLaunchKernelGGL((gemm_kernel<true, true>), grid, dim3(256), 0, st, a);
  return launch_status("sph_mlp/gemm");
}

static GemmArgs gargs(int M, int N, int K, const __bf16* A, int64_t lda, const __bf16* B,
                      int64_t ldb) {
  GemmArgs a{};
  a.M = M;
  a.N = N;
  a.K = K;
  a.A = A;
  a.lda = lda;
  a.B = B;
  a.ldb = ldb;
  return a;
}

// Weight gradient G[Kdim x N] += X^T D with X [rows x Kdim] (ldx), D [rows x N]
// (ldd), reduced over rows in split-K slabs then summed deterministically.
static int32_t weight_grad(Model* M, hipStream_t st, int rows, int kdim, int ncols,
                           const __bf16* X, int64_t ldx, const __bf16* D, int64_t ldd,
                           float* G) {
  const int splits = rows > 4096 ? std::min(64, rows / 2048) : 1;
  const int64_t slab = (int64_t)kdim * ncols;
  SPH_TRY(ensure(M->part, (size_t)splits * slab * 4 + 1024));
  GemmArgs a = gargs(kdim, ncols, rows, X, ldx, D, ldd);
  a.C = (float*)M->part.p;
  a.ldc = ncols;
  a.c_zstride = slab;
  SPH_TRY(gemm(st, true, false, a, splits));
  const int kc = ((rows + splits - 1) / splits + 15) / 16 * 16;
  const int nz = (rows + kc - 1) / kc;
  hipLaunchKernelGGL(slab_sum_kernel, dim3(nblk(slab)), dim3(256), 0, st,
                     (const float*)M->part.p, nz < 1 ? 1 : nz, slab, kdim, ncols,
                     (int64_t)ncols, G, 1);
  return launch_status("sph_mlp/weight_grad");
}

// bias gradient: G[cols] += column sums of x[rows x cols]
static int32_t bias_grad(Model* M, hipStream_t st, int rows, int cols, const float* x,
                         int64_t ld, float* G) {
  const int nb = (rows + 255) / 256;
  SPH_TRY(ensure(M->part, (size_t)nb * cols * 4 + 1024));
  hipLaunchKernelGGL(colsum_part_kernel, dim3(nb), dim3(256), 0, st, rows, cols, x, ld,
                     (float*)M->part.p);
  hipLaunchKernelGGL(slab_sum_kernel, dim3(nblk(cols)), dim3(256), 0, st,
                     (const float*)M->part.p, nb, (int64_t)cols, 1, cols, (int64_t)cols, G, 1);
  return launch_status("sph_mlp/bias_grad");
}

static FeatArgs feat_args(const sph_mlp_graph* g) {
  FeatArgs f;
  f.n = (int)g->n;
  f.pos4 = (const float4*)g->pos4;
  f.vel4 = (const float4*)g->vel4;
  f.rho = g->rho;
  f.btype = g->btype;
  f.off = g->off;
  f.nbr = g->nbr;
  f.rho0 = (float)g->rho0;
  f.co = (float)g->co;
  f.gamma = (float)g->gamma;
  f.B = (float)(g->co * g->co * g->rho0 / g->gamma);
  f.u_scale = (float)g->u_scale;
  f.g = (float)g->g;
  f.rho_scale = (float)g->rho_scale;
  f.alpha = (float)g->alpha;
  f.eta = (float)g->eta;
  f.dim = g->dim;
  return f;
}

static int32_t check_graph(const sph_mlp_graph* g, const char* who) {
  if (!g || g->n < 0 || g->n > INT32_MAX / 4 || g->nedges < 0 ||
      (g->n > 0 && (!g->pos4 || !g->vel4 || !g->rho || !g->btype || !g->off || !g->nbr)) ||
      !(g->rho0 > 0) || !(g->u_scale > 0) || !(g->g > 0) || (g->dim != 2 && g->dim != 3)) {
    set_error("%s: invalid graph", who);
    return SPH_EINVAL;
  }
  return SPH_OK;
}

// Transposed BF16 copies of the GEMM weights (W_in, W_ab, W_u).
static int32_t refresh_transposed(Model* M, hipStream_t st) {
  const Layout& L = M->L;
  struct Mat {
    int64_t off;
    int K, N;
  };
  std::vector<Mat> mats = {{L.w_in, kIn, kH}};
  for (int r = 0; r < kR; ++r) {
    mats.push_back({L.w_ab[r], kH, 2 * kH});
    mats.push_back({L.w_u[r], 2 * kH, kH});
  }
  for (const Mat& m : mats)
    hipLaunchKernelGGL(transpose_bf16_kernel, dim3(nblk((int64_t)m.K * m.N)), dim3(256), 0, st,
                       m.K, m.N, M->w + m.off, M->wbt + m.off);
  return launch_status("sph_mlp/transpose");
}

}  // namespace mlp
}  // namespace sph

using namespace sph;
using namespace sph::mlp;

extern "C" {

int64_t sph_mlp_param_count(void) { return layout().total; }

int32_t sph_mlp_create(void** model) {
  if (!model) {
    set_error("sph_mlp_create: NULL output");
    return SPH_EINVAL;
  }
  Model* M = new Model();
  M->L = layout();
  const size_t n = (size_t)M->L.total;
  hipError_t e = hipMalloc((void**)&M->w, n * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&M->wb, n * 2);
  if (e == hipSuccess) e = hipMalloc((void**)&M->wbt, n * 2);
  if (e == hipSuccess) e = hipMalloc((void**)&M->g, n * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&M->m, n * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&M->v, n * 4);
  if (e == hipSuccess) e = hipMemset(M->w, 0, n * 4);
  if (e == hipSuccess) e = hipMemset(M->wb, 0, n * 2);
  if (e == hipSuccess) e = hipMemset(M->g, 0, n * 4);
  if (e == hipSuccess) e = hipMemset(M->m, 0, n * 4);
  if (e == hipSuccess) e = hipMemset(M->v, 0, n * 4);
  if (e != hipSuccess) {
    set_error("sph_mlp_create: %s", hipGetErrorString(e));
    return (int32_t)e;
  }
  *model = M;
  return SPH_OK;
}

int32_t sph_mlp_destroy(void* model) {
  Model* M = (Model*)model;
  if (!M) return SPH_OK;
  for (void* p : {(void*)M->w, (void*)M->wb, (void*)M->wbt, (void*)M->g, (void*)M->m,
                  (void*)M->v})
    (void)hipFree(p);
  for (hipEvent_t e : M->ev) (void)hipEventDestroy(e);
  for (Buf* b : {&M->X, &M->E, &M->N4, &M->dH, &M->dZ, &M->dZb, &M->dHA, &M->dPQ, &M->dPQb, &M->part,
                 &M->lossp})
    (void)hipFree(b->p);
  for (int r = 0; r <= kR; ++r) (void)hipFree(M->H[r].p);
  for (int r = 0; r < kR; ++r) {
    (void)hipFree(M->HA[r].p);
    (void)hipFree(M->PQ[r].p);
    (void)hipFree(M->U[r].p);
  }
  delete M;
  return SPH_OK;
}

// Synchronous copies of all parameters / gradients (FP32, layout order).
int32_t sph_mlp_set_params(void* model, const float* host) {
  Model* M = (Model*)model;
  if (!M || !host) {
    set_error("sph_mlp_set_params: bad arguments");
    return SPH_EINVAL;
  }
  hipError_t e = hipMemcpy(M->w, host, M->L.total * 4, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    set_error("sph_mlp_set_params: %s", hipGetErrorString(e));
    return (int32_t)e;
  }
  hipLaunchKernelGGL(to_bf16_kernel, dim3(nblk(M->L.total)), dim3(256), 0, 0, M->L.total, M->w,
                     M->wb);
  SPH_TRY(refresh_transposed(M, 0));
  e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    set_error("sph_mlp_set_params: %s", hipGetErrorString(e));
    return (int32_t)e;
  }
  return SPH_OK;
}

int32_t sph_mlp_get(void* model, int32_t which, float* host) {
  Model* M = (Model*)model;
  if (!M || !host || which < 0 || which > 1) {
    set_error("sph_mlp_get: bad arguments");
    return SPH_EINVAL;
  }
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess)
    e = hipMemcpy(host, which == 0 ? M->w : M->g, M->L.total * 4, hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    set_error("sph_mlp_get: %s", hipGetErrorString(e));
    return (int32_t)e;
  }
  return SPH_OK;
}

int32_t sph_mlp_timing(void* model, int32_t enable, double* ms_host, int64_t* calls_host) {
  Model* M = (Model*)model;
  if (!M) {
    set_error("sph_mlp_timing: NULL model");
    return SPH_EINVAL;
  }
  if (ms_host)
    for (int k = 0; k < 4; ++k) ms_host[k] = M->t_ms[k];
  if (calls_host) *calls_host = M->t_calls;
  if (enable >= 0) {
    M->timing = enable ? 1 : 0;
    for (int k = 0; k < 4; ++k) M->t_ms[k] = 0.0;
    M->t_calls = 0;
  }
  return SPH_OK;
}

int32_t sph_mlp_zero_grad(void* model, void* stream) {
  Model* M = (Model*)model;
  if (!M) {
    set_error("sph_mlp_zero_grad: NULL model");
    return SPH_EINVAL;
  }
  hipLaunchKernelGGL(fill_f32_kernel, dim3(nblk(M->L.total)), dim3(256), 0, as_stream(stream),
                     M->L.total, M->g, 0.f);
  return launch_status("sph_mlp_zero_grad");
}

int32_t sph_mlp_forward(void* model, void* stream, const sph_mlp_graph* gr, float* out_dev) {
  Model* M = (Model*)model;
  if (!M || !out_dev) {
    set_error("sph_mlp_forward: bad arguments");
    return SPH_EINVAL;
  }
  SPH_TRY(check_graph(gr, "sph_mlp_forward"));
  const int64_t n = gr->n, ne = gr->nedges;
  if (n == 0) return SPH_OK;
  hipStream_t st = as_stream(stream);
  const Layout& L = M->L;
  SPH_TRY(ensure(M->X, (size_t)n * kIn * 2));
  SPH_TRY(ensure(M->E, (size_t)std::max<int64_t>(ne, 1) * kE * 4));
  for (int r = 0; r <= kR; ++r) SPH_TRY(ensure(M->H[r], (size_t)n * kH * 4));
  for (int r = 0; r < kR; ++r) {
    SPH_TRY(ensure(M->HA[r], (size_t)n * 2 * kH * 2));
    SPH_TRY(ensure(M->PQ[r], (size_t)n * 2 * kH * 4));
    SPH_TRY(ensure(M->U[r], (size_t)n * kH * 4));
  }
  M->n_act = n;
  M->e_act = ne;
  // inference: the pre-residual activations (the backward pass's ReLU
  // masks) are not stored
  const bool infer = (gr->flags & SPH_MLP_INFER) != 0;
  M->acts = !infer;
  SPH_TRY(ensure(M->N4, (size_t)n * 4 * 4));
  Timer T(M, st);
  const FeatArgs f = feat_args(gr);
  T.begin(0);
  hipLaunchKernelGGL(edge_feat_kernel, dim3((int)((n + kFeatNodes - 1) / kFeatNodes)), dim3(256),
                     0, st, f,
                     (float*)M->E.p, (float*)M->N4.p);
  hipLaunchKernelGGL(node_feat_kernel, dim3(nblk(n)), dim3(256), 0, st, f,
                     (const float*)M->N4.p, (__bf16*)M->X.p);
  T.end();
  SPH_TRY(launch_status("sph_mlp_forward/features"));
  // encoder: H0 = ReLU(X W_in + b_in) -> H[0] (FP32), HA[0] left half (BF16)
  {
    GemmArgs a = gargs((int)n, kH, kIn, (const __bf16*)M->X.p, kIn, M->wbt + L.w_in, kIn);
    a.C = (float*)M->H[0].p;
    a.ldc = kH;
    a.bias = M->w + L.b_in;
    a.relu = 1;
    a.Cb = (__bf16*)M->HA[0].p;
    a.ldcb = 2 * kH;
    T.begin(1);
    SPH_TRY(gemm(st, false, true, a));
    T.end();
  }
  for (int r = 0; r < kR; ++r) {
    __bf16* HA = (__bf16*)M->HA[r].p;
    float* PQ = (float*)M->PQ[r].p;
    {  // [P | Q] = H W_ab
      GemmArgs a = gargs((int)n, 2 * kH, kH, HA, 2 * kH, M->wbt + L.w_ab[r], kH);
      a.C = PQ;
      a.ldc = 2 * kH;
      T.begin(1);
      SPH_TRY(gemm(st, false, true, a));
      T.end();
    }
    EdgeArgs e{};
    e.n = (int)n;
    e.off = gr->off;
    e.nbr = gr->nbr;
    e.E = (const float*)M->E.p;
    e.PQ = PQ;
    e.Wc = M->w + L.w_c[r];
    e.be = M->w + L.b_e[r];
    e.A = HA + kH;
    e.lda = 2 * kH;
    T.begin(2);
    // the MFMA kernel indexes Q rows with 32-bit offsets (n * 2 * kH < 2^31)
    if (SPH_EDGE_MFMA && !(gr->flags & SPH_MLP_EDGE_FP32) && n < (int64_t(1) << 22))
      hipLaunchKernelGGL(edge_fwd_mfma_kernel, dim3(edge_fwd_blocks(n)), dim3(256), 0, st, e);
    else
      hipLaunchKernelGGL(edge_fwd_kernel, dim3((int)((n + 3) / 4)), dim3(256), 0, st, e);
    T.end();
    SPH_TRY(launch_status("sph_mlp_forward/edges"));
    {  // U = ReLU([H | A] W_u + c_u); H' = H + U
      GemmArgs a = gargs((int)n, kH, 2 * kH, HA, 2 * kH, M->wbt + L.w_u[r], 2 * kH);
      a.bias = M->w + L.c_u[r];
      a.relu = 1;
      a.U = infer ? nullptr : (float*)M->U[r].p;
      a.ldu = kH;
      a.res = (const float*)M->H[r].p;
      a.ldr = kH;
      a.C = (float*)M->H[r + 1].p;
      a.ldc = kH;
      if (r + 1 < kR) {
        a.Cb = (__bf16*)M->HA[r + 1].p;
        a.ldcb = 2 * kH;
      }
      T.begin(1);
      SPH_TRY(gemm(st, false, true, a));
      T.end();
    }
  }
  T.begin(3);
  hipLaunchKernelGGL(decode_kernel, dim3(nblk(n)), dim3(256), 0, st, (int)n,
                     (const float*)M->H[kR].p, M->w + L.w_out, M->w + L.b_out,
                     (const float*)M->N4.p, out_dev);
  T.end();
  SPH_TRY(launch_status("sph_mlp_forward"));
  T.collect();
  return SPH_OK;
}

int32_t sph_mlp_loss(void* model, void* stream, const sph_mlp_graph* gr, const float* out_dev,
                     const float* target_dev, float* loss_dev, float* dout_dev) {
  Model* M = (Model*)model;
  if (!M || !out_dev || !target_dev || !loss_dev || !dout_dev) {
    set_error("sph_mlp_loss: bad arguments");
    return SPH_EINVAL;
  }
  SPH_TRY(check_graph(gr, "sph_mlp_loss"));
  const int64_t n = gr->n;
  if (n == 0) return SPH_OK;
  hipStream_t st = as_stream(stream);
  const int nb = nblk(n);
  SPH_TRY(ensure(M->lossp, (size_t)nb * 4 + 256));
  const float inv_nf = gr->n_fluid > 0 ? 1.f / (float)gr->n_fluid : 0.f;
  hipLaunchKernelGGL(loss_kernel, dim3(nb), dim3(256), 0, st, (int)n, gr->btype, out_dev,
                     target_dev, inv_nf, dout_dev, (float*)M->lossp.p);
  hipLaunchKernelGGL(slab_sum_kernel, dim3(1), dim3(256), 0, st, (const float*)M->lossp.p, nb,
                     (int64_t)1, 1, 1, (int64_t)1, loss_dev, 0);
  return launch_status("sph_mlp_loss");
}

int32_t sph_mlp_backward(void* model, void* stream, const sph_mlp_graph* gr,
                         const float* dout_dev) {
  Model* M = (Model*)model;
  if (!M || !dout_dev) {
    set_error("sph_mlp_backward: bad arguments");
    return SPH_EINVAL;
  }
  SPH_TRY(check_graph(gr, "sph_mlp_backward"));
  const int64_t n = gr->n;
  if (n == 0) return SPH_OK;
  if (M->n_act != n || M->e_act != gr->nedges || !M->acts) {
    set_error("sph_mlp_backward: needs the forward pass of the same graph first (without "
              "SPH_MLP_INFER)");
    return SPH_EINVAL;
  }
  hipStream_t st = as_stream(stream);
  const Layout& L = M->L;
  SPH_TRY(ensure(M->dH, (size_t)n * kH * 4));
  SPH_TRY(ensure(M->dZ, (size_t)n * kH * 4));
  SPH_TRY(ensure(M->dZb, (size_t)n * 2 * kH * 2));
  SPH_TRY(ensure(M->dHA, (size_t)n * 2 * kH * 4));
  SPH_TRY(ensure(M->dPQ, (size_t)n * 2 * kH * 4));
  SPH_TRY(ensure(M->dPQb, (size_t)n * 2 * kH * 2));
  {
    // one scratch buffer serves every partial reduction of this pass: size
    // it once, before any launch reads it
    const int64_t wg = 64LL * 2 * kH * kH;                       // split-K slabs
    const int64_t eg = 1024LL * 4 * (kE + 1) * kH;                // edge partials
    const int64_t dg = ((n + 63) / 64) * (kH * kOut + kOut);      // decoder partials
    const int64_t bg = ((n + 255) / 256) * kH;                    // bias partials
    SPH_TRY(ensure(M->part, (size_t)std::max(std::max(wg, eg), std::max(dg, bg)) * 4 + 1024));
  }
  float* dH = (float*)M->dH.p;
  // decoder
  {
    const int nb = (int)((n + 63) / 64);
    const int64_t per = kH * kOut + kOut;
    SPH_TRY(ensure(M->part, (size_t)nb * per * 4 + 1024));
    hipLaunchKernelGGL(decode_bwd_kernel, dim3(nb), dim3(256), 0, st, (int)n,
                       (const float*)M->H[kR].p, M->w + L.w_out, dout_dev, dH,
                       (float*)M->part.p);
    // W_out and b_out are contiguous: one slab sum over blocks
    hipLaunchKernelGGL(slab_sum_kernel, dim3(nblk(per)), dim3(256), 0, st,
                       (const float*)M->part.p, nb, per, 1, (int)per, per, M->g + L.w_out, 1);
    SPH_TRY(launch_status("sph_mlp_backward/decoder"));
  }
  for (int r = kR - 1; r >= 0; --r) {
    const int64_t cnt = n * kH;
    float* dZ = (float*)M->dZ.p;
    __bf16* dZb = (__bf16*)M->dZb.p;
    // dZ = dH' * (U > 0)
    hipLaunchKernelGGL(relu_bwd_kernel, dim3(nblk(cnt)), dim3(256), 0, st, cnt, kH,
                       (const float*)M->U[r].p, (int64_t)kH, (const float*)dH, (int64_t)kH, dZ,
                       dZb);
    SPH_TRY(bias_grad(M, st, (int)n, kH, dZ, kH, M->g + L.c_u[r]));
    // dW_u += [H | A]^T dZ
    SPH_TRY(weight_grad(M, st, (int)n, 2 * kH, kH, (const __bf16*)M->HA[r].p, 2 * kH, dZb, kH,
                        M->g + L.w_u[r]));
    {  // d[H | A] = dZ W_u^T
      GemmArgs a = gargs((int)n, 2 * kH, kH, dZb, kH, M->wb + L.w_u[r], kH);
      a.C = (float*)M->dHA.p;
      a.ldc = 2 * kH;
      SPH_TRY(gemm(st, false, true, a));
    }
    // dH_r = dH' + d[H]
    hipLaunchKernelGGL(add_kernel, dim3(nblk(cnt)), dim3(256), 0, st, cnt, kH, (const float*)dH,
                       (const float*)M->dHA.p, (int64_t)2 * kH, dH);
    // edges: dP, dQ, dWc, dbe
    EdgeArgs e{};
    e.n = (int)n;
    e.off = gr->off;
    e.nbr = gr->nbr;
    e.E = (const float*)M->E.p;
    e.PQ = (const float*)M->PQ[r].p;
    e.Wc = M->w + L.w_c[r];
    e.be = M->w + L.b_e[r];
    e.dA = (const float*)M->dHA.p + kH;
    e.ldda = 2 * kH;
    e.dPQ = (float*)M->dPQ.p;
    e.dPQb = (__bf16*)M->dPQb.p;
    const int nbk = (int)std::min<int64_t>(1024, (n + 3) / 4);
    const int64_t per = (int64_t)(kE + 1) * kH;
    SPH_TRY(ensure(M->part, (size_t)nbk * 4 * per * 4 + 1024));
    e.part = (float*)M->part.p;
    hipLaunchKernelGGL(edge_bwd_kernel, dim3(nbk), dim3(256), 0, st, e);
    // W_c and b_e are contiguous
    hipLaunchKernelGGL(slab_sum_kernel, dim3(nblk(per)), dim3(256), 0, st,
                       (const float*)M->part.p, nbk * 4, per, 1, (int)per, per,
                       M->g + L.w_c[r], 1);
    SPH_TRY(launch_status("sph_mlp_backward/edges"));
    // dW_ab += H^T dPQ
    SPH_TRY(weight_grad(M, st, (int)n, kH, 2 * kH, (const __bf16*)M->HA[r].p, 2 * kH,
                        (const __bf16*)M->dPQb.p, 2 * kH, M->g + L.w_ab[r]));
    {  // dH += dPQ W_ab^T
      GemmArgs a = gargs((int)n, kH, 2 * kH, (const __bf16*)M->dPQb.p, 2 * kH,
                         M->wb + L.w_ab[r], 2 * kH);
      a.C = dH;
      a.ldc = kH;
      a.accumulate = 1;
      SPH_TRY(gemm(st, false, true, a));
    }
  }
  // encoder: dZ0 = dH * (H0 > 0); dW_in += X^T dZ0; db_in
  {
    const int64_t cnt = n * kH;
    float* dZ = (float*)M->dZ.p;
    __bf16* dZb = (__bf16*)M->dZb.p;
    hipLaunchKernelGGL(relu_bwd_kernel, dim3(nblk(cnt)), dim3(256), 0, st, cnt, kH,
                       (const float*)M->H[0].p, (int64_t)kH, (const float*)dH, (int64_t)kH, dZ,
                       dZb);
    SPH_TRY(bias_grad(M, st, (int)n, kH, dZ, kH, M->g + L.b_in));
    SPH_TRY(weight_grad(M, st, (int)n, kIn, kH, (const __bf16*)M->X.p, kIn, dZb, kH,
                        M->g + L.w_in));
  }
  return launch_status("sph_mlp_backward");
}

int32_t sph_mlp_adam(void* model, void* stream, float lr, float b1, float b2, float eps,
                     int32_t t) {
  Model* M = (Model*)model;
  if (!M || t < 1) {
    set_error("sph_mlp_adam: bad arguments");
    return SPH_EINVAL;
  }
  const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
  hipLaunchKernelGGL(adam_kernel, dim3(nblk(M->L.total)), dim3(256), 0, as_stream(stream),
                     M->L.total, M->w, M->wb, M->g, M->m, M->v, lr, b1, b2, eps, c1, c2);
  SPH_TRY(launch_status("sph_mlp_adam"));
  return refresh_transposed(M, as_stream(stream));
}

int32_t sph_mlp_integrate(void* stream, int64_t n, const int32_t* btype, const float* out_dev,
                          double a_scale, double dt, float* x, float* y, float* u, float* v) {
  if (n < 0 || (n > 0 && (!btype || !out_dev || !x || !y || !u || !v))) {
    set_error("sph_mlp_integrate: bad arguments");
    return SPH_EINVAL;
  }
  if (n == 0) return SPH_OK;
  hipLaunchKernelGGL(integrate_kernel, dim3(nblk(n)), dim3(256), 0, as_stream(stream), n, btype,
                     out_dev, (float)a_scale, (float)dt, x, y, u, v);
  return launch_status("sph_mlp_integrate");
}

}  // extern "C"
