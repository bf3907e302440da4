/*
 * sph_mi355x.h -- C ABI of libsph_mi355x.so, the MI355X (gfx950) particle
 * step of PySPH's WCSPH hot path.
 *
 * Conventions
 *   - Every pointer inside a struct or argument named *_dev, and every field
 *     pointer of sph_view, is a DEVICE pointer (e.g. torch.Tensor.data_ptr()).
 *     Arguments named *_host are host pointers.
 *   - Every call is asynchronous on `stream` (a hipStream_t passed as void*;
 *     NULL = the default stream) unless its comment says it synchronises.
 *   - Return codes: 0 ok; >0 a hipError_t; SPH_EINVAL (-1) bad argument;
 *     SPH_ECAPACITY (-2) output too small (`needed` is set);
 *     SPH_EJUMP (-3) a particle moved more than one cell since the previous
 *     binning; SPH_EUNSUPPORTED (-4) unsupported dimension or dtype.
 *     sph_last_error() returns the message of the last failure (per thread).
 *   - The library never frees or keeps caller buffers past a call.
 *   - Outputs are written, not accumulated: one call over all sources equals
 *     the reference's zero-then-accumulate (sph_calc.pyx:223-276).  Particles
 *     whose tag is not LocalReal (0) get 0 (sph_func.pyx:211-234).
 *
 * Reference paths are relative to source/pysph/ of pankajp/pysph.
 */
#ifndef SPH_MI355X_H
#define SPH_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPH_ABI_VERSION 1

#define SPH_OK 0
#define SPH_EINVAL (-1)
#define SPH_ECAPACITY (-2)
#define SPH_EJUMP (-3)
#define SPH_EUNSUPPORTED (-4)

#define SPH_F32 0
#define SPH_F64 1

#define SPH_MAX_SRC 4

/* Source roles of the fused momentum pass. */
#define SPH_SRC_MOMENTUM 1u
#define SPH_SRC_XSPH 2u

/* Pair arithmetic of the neighbour passes when the storage is FP32:
 * SPH_PAIR_NATIVE computes in the storage type, SPH_PAIR_F64 evaluates each
 * pair and accumulates in FP64 (storage stays FP32). */
#define SPH_PAIR_NATIVE 0
#define SPH_PAIR_F64 1

/* One ParticleArray on the device (particle_array.pyx:96-1279): SoA, one
 * contiguous array per property, all real fields of type `dtype`.  Fields a
 * call does not read may be NULL. */
typedef struct sph_view {
    int64_t n;
    int32_t dtype; /* SPH_F32 or SPH_F64 */
    int32_t reserved;
    const void *x, *y, *z, *u, *v, *w, *h, *m, *rho, *p, *cs;
    const int64_t *tag;        /* NULL: every particle LocalReal */
    const int32_t *cell;       /* per particle: linear cell from sph_bin */
    const int32_t *cell_start; /* ncells + 1 offsets from sph_bin */
    /* Optional packed FP32 copies (float4 per particle) read by the fast
     * neighbour passes: pos4 = (x,y,z,h), vel4 = (u,v,w,m) from sph_pack,
     * aux4 = (rho, cs, p/rho^2, p) written by sph_density.  When pos4 and
     * vel4 are set on the destination and every source (FP32 storage, at
     * most 2^27 particles each), the passes read them instead of the SoA
     * fields; larger arrays run the SoA kernels. */
    const void *pos4, *vel4, *aux4;
    /* Optional packed (x,y,z) FP32 copy, 12 bytes per particle, from
     * sph_pack: when a source carries it, the candidate screen of the fast
     * passes reads four candidates in three 16-byte loads. */
    const void *pos3;
    /* Optional packed (x,y,z,m) FP32 copy (float4 per particle) and the
     * array's h-uniformity pair (2 int32 on the device: FP32 bit patterns,
     * equal exactly when every particle has the same h, which is then
     * hrange[0]), both from sph_pack: when a density source carries both
     * and its h is uniform, the density pass reads one 16-byte record per
     * pair (same arithmetic, same bits). */
    const void *pm4;
    const int32_t *hrange;
    /* Optional mass-uniformity pair from sph_pack (like hrange, for m):
     * when a density source of uniform h also carries pos3 and its mass is
     * uniform, the density pass reads the 12-byte (x,y,z) records and takes
     * m = mrange[0] (same products, same bits). */
    const int32_t *mrange;
} sph_view;

/* The cell grid shared by all arrays of one Particles set.  Cell (i,j,k) of
 * the reference (floor(x/cell_size), cell.pyx:42-82) is grid cell
 * (i-lo[0], j-lo[1], k-lo[2]), clamped into [0, dims).  Clamping keeps the
 * stencil search complete because it never increases a cell distance. */
typedef struct sph_grid {
    double cell_size;
    int32_t lo[3];
    int32_t dims[3]; /* 1 for unused axes */
    /* Periodic axes (PeriodicDomain, cell.pyx:1568-1596): period[a] > 0
     * makes axis a periodic over [pmin[a], pmin[a] + period[a]); 0 = not
     * periodic.  The neighbour searches then also visit the destination's
     * periodic images and see a source at x_j + s*period (s = -1, 0, 1), the
     * coordinate of the reference's ghost copy (cell.pyx:1277-1376), added in
     * FP64.  A periodic call needs period > 2 * the search radius on that
     * axis.  The packed fast path and neighbour lists are not used then. */
    double period[3];
    double pmin[3];
    /* Bins of the grid per reference cell along each axis: cell_size above
     * is the bin size, the reference's cell (cell.pyx:893-925) is
     * bins_per_cell bins wide, and the jump check of sph_bin compares
     * reference cells, floor((lo + bin) / bins_per_cell).  0 or 1: the bins
     * are the reference's cells.  Searches only read cell_size. */
    int32_t bins_per_cell;
} sph_grid;

/* Parameters of the WCSPH operations.
 *   TaitEquation(co, ro, gamma)             eos_funcs.pyx:77-92
 *   MomentumEquation(alpha, beta, eta)      pressure_funcs.pyx:117-134
 *   XSPHCorrection(eps)                     xsph_funcs.pyx:13-18
 *   GravityForce(gx, gy, gz)                external_force.pyx:19-34 */
typedef struct sph_params {
    int32_t dim;        /* 1, 2 or 3 (CubicSplineKernel(dim)) */
    int32_t pair_prec;  /* SPH_PAIR_NATIVE or SPH_PAIR_F64 */
    double radius_scale; /* kernel.radius() == 2 */
    double co, ro, gamma;
    double alpha, beta, eta;
    double eps;
    double g[3];
    int32_t has_gravity;
    int32_t has_xsph;
    /* 1: variable-h symmetric pairs (VarHNbrParticleLocator,
     * nnps.pyx:654-749): a source interacts with a destination when
     * d2 < (h_dst*radius_scale)^2 OR d2 < (h_src*radius_scale)^2, same
     * distance arithmetic as the FixedDest test.  rmax must then be
     * radius_scale * the largest h of the destination and the sources.
     * The neighbour passes run their SoA kernels in this mode (packed
     * inputs and neighbour lists are ignored); the batched engine rejects it. */
    int32_t symmetric;
    /* Bit s set: the summation density drops the self pair of source s, which
     * must be the destination array: the intended semantics of
     * SPHFunctionParticle exclude_self (sph_func.pyx:381-394), whose own code
     * reads an attribute its class does not declare and so never takes
     * effect; no reference output pins this flag.
     * Read by sph_density / sph_density_cached only: in the other passes
     * the self pair contributes 0.  The density then runs its SoA kernels;
     * the batched engine rejects it. */
    int32_t exclude_self;
    /* 1: the queue-free FP32 density scan (sources of uniform h) stages each
     * stencil row in LDS: the wave loads the union of its lanes' candidate
     * ranges with coalesced 16-byte loads and every lane screens its own
     * range from LDS, the same sums as the default per-lane loads (rows
     * whose union exceeds 256 records take those).  Opt-in: slower on
     * MI355X at the sizes measured.  The batched engine ignores it. */
    int32_t lds_rows;
} sph_params;

int32_t sph_abi_version(void);
const char *sph_last_error(void);
/* Number of visible devices (synchronous; 0 when none). */
int32_t sph_device_count(void);

/* ---- binning: CellManager.update / Cell.update (cell.pyx:319-432,
 *      :927-1015) and the cell size rule (cell.pyx:893-925) ---------------- */

/* min/max of x,y,z and max of h over the array -> out_dev[0..6] (double):
 * xmin,ymin,zmin,xmax,ymax,zmax,hmax.  out_dev must hold SPH_BOUNDS_WORDS
 * doubles; the words after the first 8 are the reduction's scratch.  Reads
 * only x,y,z,h. */
#define SPH_BOUNDS_WORDS 456
int32_t sph_bounds(void *stream, const sph_view *a, double *out_dev);

/* Workspace bytes sph_bin needs for n particles and ncells cells. */
size_t sph_bin_workspace_bytes(int64_t n, int64_t ncells);

/* Counting-sort binning by cell key.  Writes, for the NEW (cell-sorted)
 * layout: perm_dev[p] = the current slot of the particle that goes to slot
 * p (stable: current-slot order inside a cell), cell_dev[p] its linear
 * cell, cell_start_dev[ncells+1].  When prev_cell_dev is not NULL it holds
 * each particle's previous cell in the CURRENT layout; a particle whose
 * reference cell (see bins_per_cell) moved by more than one in any axis is
 * counted into *jump_count_dev (the reference's jump tolerance,
 * cell.pyx:1422-1437).  Reads x,y,z. */
int32_t sph_bin(void *stream, const sph_view *a, const sph_grid *g,
                int32_t *perm_dev, int32_t *cell_dev, int32_t *cell_start_dev,
                const int32_t *prev_cell_dev, int32_t *jump_count_dev,
                void *workspace_dev, size_t workspace_bytes);

/* The jump tolerance check of CellManager.update (cell.pyx:1422-1437) on
 * the reference's unbounded cells floor(x / cell_size) per axis
 * (cell.pyx:57, FP64).  refcell_dev[i] holds particle i's cell from the
 * previous check (INT64_MIN: none) and receives its current one (21 bits per
 * axis, biased by 2^20, coordinates clamped to +-2^20; INT64_MIN for a
 * non-finite position); a particle whose cell moved by more than one in any
 * axis is counted into *jump_count_dev.  The caller keeps refcell_dev in the
 * particles' slot order (permuting it with them).  Unlike the check of
 * sph_bin, this one does not depend on the binning grid.  Reads x, y, z. */
int32_t sph_jump_check(void *stream, const sph_view *a, double cell_size,
                       int64_t *refcell_dev, int32_t *jump_count_dev);

/* Gather nfields arrays through perm: dst[f][p] = src[f][perm[p]].
 * elem_bytes[f] is 4 or 8.  Host arrays of device pointers; no output may
 * be one of the inputs.  Runs of consecutive sources are copied 16 bytes
 * at a time when perm and every output are 16-byte aligned (as device
 * allocations are); otherwise element by element. */
int32_t sph_permute(void *stream, int64_t n, const int32_t *perm_dev,
                    int32_t nfields, const void *const *src_dev_host,
                    void *const *dst_dev_host, const int32_t *elem_bytes_host);

/* Wrap coordinates into the periodic domain of `g`: per periodic axis a,
 * a particle with pos[a] >= pmin + period moves by -period, one with
 * pos[a] < pmin by +period (the binning point rule of cell.pyx:392-398,
 * applied to the stored coordinate).  Each of the nextra arrays of
 * extra_dev_host[3*k + a] (e.g. the integrator's copies of x, y, z) moves
 * by the same amount as pos[a].  Arrays of dtype, n elements each; NULL
 * entries for non-periodic axes are allowed. */
int32_t sph_periodic_wrap(void *stream, const sph_grid *g, int32_t dtype, int64_t n,
                          void *const *pos_dev_host, int32_t nextra,
                          void *const *extra_dev_host);

/* Write the packed copies pos4[i] = (x,y,z,h), vel4[i] = (u,v,w,m),
 * pm4[i] = (x,y,z,m) (float4 each) and pos3[i] = (x,y,z) (3 floats) of an
 * FP32 array, its h-uniformity pair hrange[0..1] (the bit pattern of h[0]
 * twice when every h equals it; otherwise hrange[1] holds another h) and
 * its mass-uniformity pair mrange[0..1] (the same for m); any output may be
 * NULL. */
int32_t sph_pack(void *stream, const sph_view *a, void *pos4_dev, void *vel4_dev,
                 void *pos3_dev, void *pm4_dev, int32_t *hrange_dev, int32_t *mrange_dev);

/* ---- FixedDestNbrParticleLocator.get_nearest_particles (nnps.pyx:458-620)
 *      as a CSR export: the exact FP64 test
 *      ((xs-xd)^2 + (ys-yd)^2) + (zs-zd)^2 < (h_d*radius_scale)^2
 *      (point.pxd:35-37, nnps.pyx:311-354, :552), rows in the destination's
 *      current layout, entries = src_ids[j] sorted ascending. ------------- */

/* counts_dev[i] = number of neighbours of destination i. */
int32_t sph_neighbor_counts(void *stream, const sph_grid *g, const sph_view *dst,
                            const sph_view *src, double radius_scale,
                            int64_t *counts_dev);
/* Exclusive scan of counts into offsets_dev[n+1]; synchronises and returns
 * the total through *total_host. */
int32_t sph_scan_counts(void *stream, int64_t n, const int64_t *counts_dev,
                        int64_t *offsets_dev, int64_t *total_host);
/* sph_scan_counts without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The scan's
 * tile partials live in a fixed device array inside the library (512 KiB per
 * GPU, held until the process ends), so n is limited to
 * SPH_SCAN_NOALLOC_MAX_COUNTS (2^27) and larger n returns SPH_EINVAL.  Calls
 * that overlap on one GPU share that array and race; serialise them (one
 * stream, or events).  total_host may be NULL (no synchronisation). */
#define SPH_SCAN_NOALLOC_MAX_COUNTS (1LL << 27)
int32_t sph_scan_counts_noalloc(void *stream, int64_t n, const int64_t *counts_dev,
                                int64_t *offsets_dev, int64_t *total_host);
/* Fill the CSR rows.  cap < offsets[n] -> SPH_ECAPACITY with *needed_host. */
int32_t sph_neighbor_fill(void *stream, const sph_grid *g, const sph_view *dst,
                          const sph_view *src, double radius_scale,
                          const int64_t *src_ids_dev, const int64_t *offsets_dev,
                          int64_t *nbrs_dev, int64_t cap, int64_t total,
                          int64_t *needed_host);

/* sph_neighbor_fill without a host read-back of the total (no
 * synchronisation): offsets_dev come from sph_scan_counts with a NULL
 * total; entries at positions >= cap are not written and set
 * *overflow_dev to 1 (the caller zeroes it and checks it later).  For
 * callers that size the CSR from a known bound, such as a graph rebuilt
 * every step. */
int32_t sph_neighbor_fill_capped(void *stream, const sph_grid *g, const sph_view *dst,
                                 const sph_view *src, double radius_scale,
                                 const int64_t *src_ids_dev, const int64_t *offsets_dev,
                                 int64_t *nbrs_dev, int64_t cap, int32_t *overflow_dev);

/* ---- VarHNbrParticleLocator (nnps.pyx:654-749) as a CSR export: the
 *      symmetric set d2 < (h_d*rs)^2 OR d2 < (h_s*rs)^2, rows sorted
 *      ascending like the FixedDest export.  rmax = radius_scale * the
 *      largest h of dst and src (it sizes the stencil). ---------------- */
int32_t sph_varh_neighbor_counts(void *stream, const sph_grid *g, const sph_view *dst,
                                 const sph_view *src, double radius_scale, double rmax,
                                 int64_t *counts_dev);
int32_t sph_varh_neighbor_fill(void *stream, const sph_grid *g, const sph_view *dst,
                               const sph_view *src, double radius_scale, double rmax,
                               const int64_t *src_ids_dev, const int64_t *offsets_dev,
                               int64_t *nbrs_dev, int64_t cap, int64_t total,
                               int64_t *needed_host);

/* ---- neighbour passes -------------------------------------------------- */

/* Optional neighbour list of one destination array, built by a density
 * pass and read by the later passes of the same step.
 *   mode 0 (build): sph_density searches the cells and records every source
 *     within radius*(1+skin) of each destination (a superset of the
 *     neighbours) while it sums the density over the exact neighbours.
 *   mode 1 (use):   sph_density, and always sph_momentum_xsph /
 *     sph_wcsph_stage, visit the recorded sources and apply the exact test.
 * The list stays exact while no particle has moved by more than
 * skin*radius/2 since it was built; a pass given disp_dev checks that on
 * the device and searches the cells instead where it does not hold.
 * Entry k of destination i is idx[k*n + i] (n = the destination count),
 * grouped by source in call order; count[s*n + i] holds the number of
 * entries of source s (idx: max_per*n words, count: SPH_MAX_SRC*n).  A
 * destination with more than max_per entries keeps max_per of them and is
 * searched in the cells by the passes that read the list.  The sources and
 * their order, and the slot order of all arrays, must not change between
 * the build and the uses.  Only the packed FP32 fast path uses lists. */
typedef struct sph_nlist {
    int32_t *idx;
    int32_t *count;
    int32_t max_per;
    int32_t mode;
    float skin;
    int32_t reserved;
} sph_nlist;

/* SPHRho over all sources (density_funcs.pyx:38-84), optionally followed by
 * TaitEquation on the destination (eos_funcs.pyx:105-116) when p_dev and
 * cs_dev are both non-NULL.  With packed inputs the destination's aux4
 * (when set) receives (rho, cs, p/rho^2, p).  `disp_dev` (nullable, one float) bounds the
 * straight-line displacement |x_now - x_binned| of every particle since the
 * arrays were binned; the stencil ring and the culled search radius grow by
 * it so that a binning can be reused across an RK2 stage.  rmax = the
 * largest search radius (radius_scale * max h).  `nlist_out` (nullable)
 * records the accepted sources; lists are written by the packed FP32 path
 * only, so a call that takes the SoA kernels (FP64 storage, no packed
 * inputs, symmetric pairs, periodic axes or exclude_self) with a non-NULL
 * `nlist_out` returns SPH_EINVAL instead of leaving the list unwritten. */
int32_t sph_density(void *stream, const sph_grid *g, const sph_params *prm,
                    const sph_view *dst, int32_t nsrc, const sph_view *srcs,
                    double rmax, const float *disp_dev, void *rho_dev,
                    void *p_dev, void *cs_dev, const sph_nlist *nlist_out);

/* sph_density for ndst (1 or 2) destination arrays over the same sources:
 * the SPHRho (+ TaitEquation) calcs of every array of a step, e.g. the
 * fluid's and the boundary's (density_funcs.pyx:38-84, eos_funcs.pyx:
 * 105-116).  When both destinations take the packed FP32-pair scan with
 * the same epilogue (both or neither with p, cs), one launch serves them:
 * its blocks cover both destination sets, so the shorter pass fills the GPU
 * while the longer one drains; otherwise each destination is one
 * sph_density call.  rho_dev[k], p_dev[k], cs_dev[k] belong to dsts[k]
 * (p_dev / cs_dev may be NULL); no lists, source caches or exclude_self. */
int32_t sph_density_multi(void *stream, const sph_grid *g, const sph_params *prm,
                          int32_t ndst, const sph_view *dsts, int32_t nsrc,
                          const sph_view *srcs, double rmax, const float *disp_dev,
                          void *const *rho_dev, void *const *p_dev, void *const *cs_dev);

/* SPHDensityRate over all sources (density_funcs.pyx:89-158, hks off):
 * rate_dev[a] = sum_b m_b (v_a - v_b) . grad_a W(r_a - r_b, (h_a+h_b)/2),
 * the continuity-equation density derivative.  Reads x,y,z,h,u,v,w of the
 * destination and x,y,z,h,u,v,w,m of the sources. */
int32_t sph_density_rate(void *stream, const sph_grid *g, const sph_params *prm,
                         const sph_view *dst, int32_t nsrc, const sph_view *srcs,
                         double rmax, const float *disp_dev, void *rate_dev);

/* sph_density with the partial sums of one source kept between calls.
 * When a source does not move relative to the destination (boundary
 * particles from boundary particles), its contribution to each
 * destination's density is the same at every call: mode SPH_CACHE_FILL
 * computes it as usual and also stores it in sums_dev (one double per
 * destination); mode SPH_CACHE_USE adds the stored value instead of
 * scanning that source.  The partial sum enters the total in the same
 * position, so the result has the same bits.  The caller returns to
 * SPH_CACHE_FILL whenever the destination or that source moved, changed
 * mass or smoothing length, or was re-binned.  Not combinable with a
 * neighbour list (nlist_out must be NULL). */
#define SPH_CACHE_FILL 1
#define SPH_CACHE_USE 2
typedef struct sph_src_cache {
    double *sums;
    int32_t src;  /* index of the cached source in srcs */
    int32_t mode; /* SPH_CACHE_FILL or SPH_CACHE_USE */
} sph_src_cache;
int32_t sph_density_cached(void *stream, const sph_grid *g, const sph_params *prm,
                           const sph_view *dst, int32_t nsrc, const sph_view *srcs,
                           double rmax, const float *disp_dev, void *rho_dev, void *p_dev,
                           void *cs_dev, const sph_src_cache *cache);

/* TaitEquation alone: p, cs from rho (eos_funcs.pyx:105-116). */
int32_t sph_tait(void *stream, const sph_params *prm, const sph_view *a,
                 void *p_dev, void *cs_dev);

/* MomentumEquation (pressure_funcs.pyx:154-245) over sources flagged
 * SPH_SRC_MOMENTUM and XSPHCorrection (xsph_funcs.pyx:31-88) over sources
 * flagged SPH_SRC_XSPH, one pass.  acc_dev[3] / dv_dev[3] are the outputs
 * (components >= dim may be NULL; a NULL acc_dev[0] or dv_dev[0] skips that
 * operation).  GravityForce is NOT added here (it has its own k array). */
int32_t sph_momentum_xsph(void *stream, const sph_grid *g, const sph_params *prm,
                          const sph_view *dst, int32_t nsrc, const sph_view *srcs,
                          const uint32_t *src_flags_host, double rmax,
                          const float *disp_dev, void *const *acc_dev,
                          void *const *dv_dev, const sph_nlist *nlist_in);

/* Fused WCSPH integrator stage for one fluid destination array: the same
 * pair sums as sph_momentum_xsph followed by the reference integrator's
 * update of that particle (integrator.py:472-514, :534-551, :597-619), with
 * PositionStepping (k_x = u, position_funcs.pyx:40-60) and GravityForce
 * (k_u = g, external_force.pyx:39-45) as separate calcs in this order:
 * mom, gravity, step, xsph.
 *   stage 1 (RK2 predictor):  k1 = (a, dv) -> k1_dev[6];
 *            out.u = (u0 + a*dt) + g*dt;  out.x = (x0 + u0*dt) + dv*dt
 *   stage 2 (RK2 final):      out.u = (u0 + hdt*(k1a+a)) + hdt*(g+g)
 *            out.x = (x0 + hdt*(u0+u)) + hdt*(k1dv+dv),  hdt = 0.5*dt
 *   stage 0 (Euler):          out.u = (u0 + a*dt) + g*dt;
 *            out.x = (x0 + u0*dt) + dv*dt
 * `dst` is the state the right-hand side is evaluated at; state0_dev[6]
 * holds x0,y0,z0,u0,v0,w0; out_dev[6] receives the new state (it may alias
 * state0_dev only for stage 2, whose pair sums read `dst` and `srcs`, never
 * state0).  disp_out_dev (nullable) receives, via atomic max, an upper bound
 * of the largest straight-line move |out.x - x0| (Euclidean over the dim
 * components; the stage-2 passes widen their ring and culled search radius
 * by it); it must be zeroed by the caller.  `nlist_in` (nullable) is read
 * by the packed FP32 path only: a call that takes the SoA kernels with a
 * non-NULL `nlist_in` returns SPH_EINVAL. */
int32_t sph_wcsph_stage(void *stream, int32_t stage, const sph_grid *g,
                        const sph_params *prm, const sph_view *dst, int32_t nsrc,
                        const sph_view *srcs, const uint32_t *src_flags_host,
                        double rmax, const float *disp_dev, double dt,
                        const void *const *state0_dev, void *const *k1_dev,
                        void *const *out_dev, float *disp_out_dev,
                        const sph_nlist *nlist_in);

/* ---- integrator updates of the generic path (integrator.py:472-514,
 *      :534-551, :597-619) --------------------------------------------- */

/* mode 0: out = a + b*s                 (Integrator.step, integrator.py:505)
 * mode 1: out = a + s*(b + c); a = out  (RK2 final_step, s = 0.5*dt, :616)
 * mode 2: out = a + b*s; a = out        (Euler final_step, :550)
 * mode 3: out = a                        (set_initial_arrays / reset copy,
 *                                         PositionStepping k = u, :40-60)
 * mode 4: out = s                        (GravityForce k = g, :39-45)
 * mode 5: out = (a + 2b) + 2c            (RK4 partial sum k1 + 2k2 + 2k3, :669-689)
 * mode 7: out = 2a - b                    (PredictorCorrector final_step, :756-772)
 * mode 8: out = a + s*(b - c)            (LeapFrog final_step, s = 0.5*dt, :866-883)
 * With tag_dev non-NULL, modes 0, 3 and 4 write 0 where tag != LocalReal
 * (sph_func.pyx:211-234: SPHFunction.eval accumulates into the outputs of
 * LocalReal destinations and sets the others to 0).  nops arrays of n elements each, one dtype. */
typedef struct sph_update_op {
    void *out;
    void *a;
    const void *b;
    const void *c;
} sph_update_op;
int32_t sph_update(void *stream, int32_t dtype, int64_t n, int32_t mode, double s,
                   int32_t nops, const sph_update_op *ops_host, const int64_t *tag_dev);

/* MoveCircleX (axis 0) / MoveCircleY (axis 1) of external_force.pyx:99-162:
 * the components of unit-speed circular motion about the origin into
 * out0_dev / out1_dev (either may be NULL).  Reads x, y, z (and tag). */
int32_t sph_move_circle(void *stream, const sph_view *a, int32_t axis, void *out0_dev,
                        void *out1_dev);

/* FP64 conservation sums over LocalReal particles of one array, added to
 * out_dev[8]: mass, momentum x,y,z, kinetic energy, -sum m g.x, Tait elastic
 * energy sum m e(rho), total (no reference counterpart; the reference
 * reports none).  The caller zeroes out_dev once per set of arrays. */
int32_t sph_conservation_f64(void *stream, const sph_params *prm, const sph_view *a,
                             double *out_dev);

/* Fill n int32 words with a value (graph-capturable, no hipMemset). */
int32_t sph_fill_i32(void *stream, int32_t *dst_dev, int64_t n, int32_t value);

/* ---- batched multi-scene WCSPH ------------------------------------------
 * No counterpart in the reference, which steps one Particles set at a time
 * (solver.py:457-508).  Many independent scenes share one set of device
 * buffers: every scene's fluid particles are one slice of a concatenated
 * fluid set, its boundary particles one slice of a boundary set.  Each scene
 * has its own grid in one global cell space (cells [cell_base, cell_base +
 * nx*ny*nz)), so one counting sort bins all scenes and keeps them
 * contiguous.  The per-particle arithmetic is that of sph_density and
 * sph_wcsph_stage (packed FP32 path, RK2), so a scene stepped in a batch
 * gives the bits it gives when stepped alone with the same grid. */
typedef struct sph_batch_scene {
    int64_t f0, nf;     /* fluid slice [f0, f0 + nf) */
    int64_t b0, nb;     /* boundary slice (nb = 0: no boundary) */
    sph_grid grid;      /* cell grid of the scene */
    sph_params prm;     /* dim, radius_scale, Tait, viscosity, XSPH, gravity */
    double rmax;        /* radius_scale * max h of the scene */
    double dt;          /* RK2 time step */
    double cost;        /* relative cost per particle (schedule order) */
} sph_batch_scene;

/* One concatenated particle set (FP32 SoA; uid = original index, carried
 * along by every binning). */
typedef struct sph_batch_set {
    float *x, *y, *z, *u, *v, *w, *h, *m, *rho, *p, *cs;
    int32_t *uid;
} sph_batch_set;

/* Device buffers of a batch, owned by the caller and fixed for the life of
 * the handle (so that a captured step stays valid).  The fluid state is in
 * fa between steps and in fb (the step's binning order) during a step. */
typedef struct sph_batch_buffers {
    int64_t nf_total, nb_total;
    sph_batch_set fa, fb;        /* fluid sets */
    sph_batch_set bnd, bnd_in;   /* boundary set (binned) and its input order */
    float *x1a[6], *x1b[6];      /* RK2 stage state X1 (x1b: only for which=1 binning) */
    float *k1a[6], *k1b[6];      /* RK2 k1 (momentum 3, XSPH 3), likewise */
    float *fpos4, *fvel4, *faux4;  /* packed fluid copies (float4 per particle) */
    float *bpos4, *bvel4, *baux4;  /* packed boundary copies */
    int32_t *fscene, *bscene;    /* scene of each slot */
    int32_t *fperm, *fcell, *fcell_start;  /* fluid binning */
    int32_t *bperm, *bcell, *bcell_start;  /* boundary binning */
    void *ws;                    /* binning workspace */
    size_t ws_bytes;             /* >= sph_bin_workspace_bytes(max(nf, nb), ncells) */
} sph_batch_buffers;

/* Work schedule of a neighbour phase (never changes results).  Tiles are 64
 * consecutive particles of one scene, one wave each.
 *   mode 0: one wave per tile (a plain grid);
 *   mode 1: persistent: blocks_per_cu blocks per CU, each XCD walks its own
 *           segment of the tile list in a fixed stride;
 *   mode 2: persistent, waves take `grab` tiles at a time from their XCD's
 *           segment with an atomic counter.
 *   order 0: tiles in scene order, XCD segments of equal tile count;
 *   order 1: XCD segments of equal estimated cost (scene cost x particles),
 *           most expensive tiles first within a segment. */
typedef struct sph_batch_sched {
    int32_t mode;
    int32_t blocks_per_cu;
    int32_t grab;
    int32_t order;
} sph_batch_sched;

/* Number of cells of the global cell space for these scenes. */
int64_t sph_batch_ncells(int32_t nscenes, const sph_batch_scene *scenes_host);
int32_t sph_batch_create(int32_t nscenes, const sph_batch_scene *scenes_host,
                         const sph_batch_buffers *buffers_host, void **handle);
int32_t sph_batch_destroy(void *handle);
/* One RK2 step of every scene is: bin(0), pack(0), density(0, 0),
 * density(0, 1), stage(1), pack(1), density(1, 0), density(1, 1), stage(2).
 * The step bins once: stage 2 evaluates X1 in the order of the step's
 * binning with the stencil widened by the predictor's largest displacement
 * (recorded on the device by stage 1), the neighbour sets being the same.
 * which 0: fluid by X (fa -> fb); 1: fluid by X1 (fb -> fa, x1a -> x1b,
 * k1a -> k1b; not part of the step); 2: boundary (bnd_in -> bnd). */
int32_t sph_batch_bin(void *handle, void *stream, int32_t which);
/* which 0: fluid from fb; 1: fluid from X1 (x1a, components >= dim and h, m
 * from fb); 2: boundary. */
int32_t sph_batch_pack(void *handle, void *stream, int32_t which);
/* SPHRho + TaitEquation of every scene on fb's order: phase 0 at X (stage
 * 1, results in fb), phase 1 at X1 (stage 2, results in fa); set 0 fluid
 * destinations, 1 boundary. */
int32_t sph_batch_density(void *handle, void *stream, int32_t phase, int32_t set,
                          const sph_batch_sched *sched);
/* RK2 stage of every scene's fluid: 1 = predictor (fb -> x1a, k1a, and the
 * largest displacement), 2 = final update (fb, x1a, k1a -> fa, with h, m, uid
 * copied along, so that fa holds the whole state for the next step). */
int32_t sph_batch_stage(void *handle, void *stream, int32_t stage,
                        const sph_batch_sched *sched);

/* ---- edge-MLP surrogate ---------------------------------------------------
 * No counterpart in the reference.  A message-passing network over the
 * symmetric radius graph of a particle set (every node a destination and a
 * source, self edges included) predicts each fluid particle's acceleration
 * (in units of g) over a time step 4x the plain one.  Node features:
 * velocity / u_scale, (rho/rho0 - 1) rho_scale, Tait pressure, boundary
 * flag.  Edge features: relative position / h, distance / h, relative
 * velocity / u_scale, both densities, W, and the pair's pressure and
 * artificial-viscosity accelerations (from the same cubic spline and Tait
 * law).  Hidden width 128, 2 rounds; FP32 master weights with BF16 copies
 * for the MFMA GEMMs, FP32 accumulation everywhere. */
#define SPH_MLP_HIDDEN 128
#define SPH_MLP_NIN 16
#define SPH_MLP_NE 16
#define SPH_MLP_ROUNDS 2
typedef struct sph_mlp_graph {
    int64_t n;            /* nodes */
    int64_t nedges;
    int64_t n_fluid;      /* nodes with btype 0 (loss normalisation) */
    int32_t dim;          /* 2 or 3 */
    /* SPH_MLP_INFER: sph_mlp_forward for inference only; it does not store
     * the activations sph_mlp_backward reads, which then returns SPH_EINVAL
     * until a forward pass without the flag.  SPH_MLP_EDGE_FP32: aggregate
     * the edges with FP32 FMAs on the vector units instead of split-BF16
     * products on the matrix cores (the path graphs of 2^22 nodes or more
     * always take; the products agree to ~2^-16, the outputs to ~1e-3, as
     * the aggregate's BF16 copy can round either way) */
    int32_t flags;
    const float *pos4;    /* (x, y, z, h) per node */
    const float *vel4;    /* (u, v, w, m) per node */
    const float *rho;     /* density per node */
    const int32_t *btype; /* 0 fluid, 1 boundary */
    const int64_t *off;   /* CSR offsets, n + 1 */
    const int64_t *nbr;   /* neighbour node of each edge */
    double rho0, co, gamma, u_scale, g, rho_scale, alpha, eta;
} sph_mlp_graph;
#define SPH_MLP_INFER 1
#define SPH_MLP_EDGE_FP32 2
int64_t sph_mlp_param_count(void);
int32_t sph_mlp_create(void **model);
int32_t sph_mlp_destroy(void *model);
/* Parameters as FP32 in layout order (synchronous copies). */
int32_t sph_mlp_set_params(void *model, const float *params_host);
/* which 0: parameters, 1: accumulated gradients (synchronous). */
int32_t sph_mlp_get(void *model, int32_t which, float *host);
int32_t sph_mlp_zero_grad(void *model, void *stream);
/* Launch timing of sph_mlp_forward (synchronises after each timed call):
 * ms_host[4] receives the accumulated milliseconds of the feature, GEMM,
 * edge-aggregation and decoder launches and *calls_host the number of timed
 * calls (either may be NULL); enable 1 / 0 switches timing on / off and
 * resets the totals, -1 only reads them. */
int32_t sph_mlp_timing(void *model, int32_t enable, double *ms_host, int64_t *calls_host);
/* out_dev: n x 2 predicted accelerations / g.  Keeps the activations for a
 * following sph_mlp_backward of the same graph. */
int32_t sph_mlp_forward(void *model, void *stream, const sph_mlp_graph *g, float *out_dev);
/* Mean over fluid nodes of |out - target|^2 -> *loss_dev; its gradient
 * with respect to out -> dout_dev (n x 2). */
int32_t sph_mlp_loss(void *model, void *stream, const sph_mlp_graph *g, const float *out_dev,
                     const float *target_dev, float *loss_dev, float *dout_dev);
/* Gradients of the parameters, added to the model's gradient buffer. */
int32_t sph_mlp_backward(void *model, void *stream, const sph_mlp_graph *g,
                         const float *dout_dev);
int32_t sph_mlp_adam(void *model, void *stream, float lr, float beta1, float beta2, float eps,
                     int32_t step);
/* Surrogate step of the fluid nodes (btype 0), symplectic Euler:
 * u += dt a_scale out, then x += dt u (2D components). */
int32_t sph_mlp_integrate(void *stream, int64_t n, const int32_t *btype, const float *out_dev,
                          double a_scale, double dt, float *x, float *y, float *u, float *v);

/* ---- stream capture into HIP graphs (any sequence of the calls above) -- */
int32_t sph_graph_begin(void *stream);
int32_t sph_graph_end(void *stream, void **graph_exec);
int32_t sph_graph_launch(void *graph_exec, void *stream);
int32_t sph_graph_destroy(void *graph_exec);

#ifdef __cplusplus
}
#endif
#endif /* SPH_MI355X_H */
