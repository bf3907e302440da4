/*
 * sph_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, FP64, CPU restatement of the reference's Cython particle step.
 * It is the checker for the HIP path: only tests/, the smoke entry point and
 * the CPU-baseline leg of bench.py may load it.  No product code links it.
 *
 * What it restates (reference paths relative to source/pysph/):
 *   cell binning .......... base/cell.pyx:42-82 (real_to_int / find_cell_id),
 *                           :93-116 (immediate neighbour list, x-major loops),
 *                           :893-925 (cell size = max_radius_scale * max_h),
 *                           :1106-1142 (potential cells, ring ceil(r/cs))
 *   FixedDest locator ..... base/nnps.pyx:271-356 (strict d2 < r2 filter),
 *                           :525-559 (eff. radius = dest h * radius_scale),
 *                           :596-620 (self index swapped to the last slot)
 *   distance order ........ base/point.pxd:35-37
 *   cubic spline .......... base/kernels.pyx:47-53 (h_dim), :345-449
 *   SPHFunction.eval ...... sph/sph_func.pyx:199-234 (non-LocalReal -> 0),
 *                           :381-394, :466-484 (per-neighbour accumulation)
 *   SPHCalc.sph ........... sph/sph_calc.pyx:223-276 (zero, then += per src)
 *   SPHRho ................ sph/funcs/density_funcs.pyx:38-84
 *   TaitEquation .......... sph/funcs/eos_funcs.pyx:84-116
 *   MomentumEquation ...... sph/funcs/pressure_funcs.pyx:154-245
 *   XSPHCorrection ........ sph/funcs/xsph_funcs.pyx:31-88
 *   SPHDensityRate ........ sph/funcs/density_funcs.pyx:89-158
 *   VarH locator .......... base/nnps.pyx:654-749 (forward query, then the
 *                           reverse locator's hits appended if absent)
 *   PositionStepping ...... sph/funcs/position_funcs.pyx:32-60
 *   GravityForce .......... sph/funcs/external_force.pyx:39-45
 *   Euler / RK2 ........... solver/integrator.py:339-514 (set_initial_arrays,
 *                           reset_current_arrays, eval, step), :525-573,
 *                           :581-645 (cumulative updates through a shared _p_0)
 *
 * Two execution modes share the same arithmetic:
 *   nthreads == 1 : "reference structure" -- hash-map cells, a per-(src,dst)
 *                   neighbour cache rebuilt once per stage, one kernel call
 *                   per pair, single thread (the reference is single-threaded).
 *   nthreads  > 1 : the same loops with an OpenMP parallel-for over
 *                   destination particles.
 * Built with -O3 -fno-fast-math -ffp-contract=off so that no FMA contraction
 * changes the rounding of the reference's expressions.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846

/* One ParticleArray as SoA double arrays (the reference's DoubleArrays). */
typedef struct {
    int64_t n;
    double *x, *y, *z, *u, *v, *w, *h, *m, *rho, *p, *cs;
    const int64_t *tag; /* NULL: every particle is LocalReal (tag 0) */
    int32_t type;       /* ParticleType: 0 Fluid, 1 Solid */
} orc_arr;

/* Neighbour cache of one (src, dst) pair in CSR form, neighbour order as the
 * reference builds it (cell stencil order, index order inside a cell, self
 * swapped last when src is dst). */
typedef struct {
    int64_t n_dst;
    int64_t *off; /* n_dst + 1 */
    int64_t *nbr;
} orc_nbrs;

static int g_threads = 1;

int orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n < 1) n = omp_get_max_threads();
    g_threads = n;
    return n;
#else
    (void)n;
    g_threads = 1;
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* cubic spline kernel: base/kernels.pyx:47-53, :345-449                    */
/* ------------------------------------------------------------------------ */
static double h_dim(double h, int dim)
{
    if (dim == 1) return 1 / h;
    if (dim == 2) return 1 / (h * h);
    return 1 / (h * h * h);
}

/* KernelBase.fac = _fac(1.0) (kernels.pyx:73, :437-443) */
static double fac_unit(int dim)
{
    if (dim == 1) return (2. / 3.) / 1.0;
    if (dim == 2) return 10 / (7 * ORC_PI) / (1.0 * 1.0);
    return 1. / ORC_PI / (1.0 * 1.0 * 1.0);
}

/* W(pa - pb, h): kernels.pyx:351-374 */
double orc_cubic_w(double pax, double pay, double paz, double pbx, double pby,
                   double pbz, double h, int dim)
{
    double rab = sqrt((pax - pbx) * (pax - pbx) + (pay - pby) * (pay - pby) +
                      (paz - pbz) * (paz - pbz));
    double q = rab / h;
    double val;
    double fac = h_dim(h, dim) * fac_unit(dim);
    if (q > 2.0)
        val = 0.0;
    else if (q > 1.0)
        val = 0.25 * (2 - q) * (2 - q) * (2 - q);
    else
        val = 1 - 1.5 * (q * q) * (1 - 0.5 * q);
    return val * fac;
}

/* grad_a W(pa - pb, h): kernels.pyx:376-413 */
void orc_cubic_grad(double pax, double pay, double paz, double pbx, double pby,
                    double pbz, double h, int dim, double *g)
{
    double rab = sqrt((pax - pbx) * (pax - pbx) + (pay - pby) * (pay - pby) +
                      (paz - pbz) * (paz - pbz));
    double q = rab / h;
    double val = 0.0;
    double fac = h_dim(h, dim) * fac_unit(dim);
    double rx = pax - pbx, ry = pay - pby, rz = paz - pbz;
    if (q > 2.0) {
    } else if (q >= 1.0) {
        val = -0.75 * (2 - q) * (2 - q) / (h * rab);
    } else if (q > 1e-14) {
        val = 3.0 * (0.75 * q - 1) / (h * h);
    }
    g[0] = rx * (val * fac);
    g[1] = ry * (val * fac);
    g[2] = rz * (val * fac);
}

/* ------------------------------------------------------------------------ */
/* cell binning: base/cell.pyx:42-82 and a hash map standing in for the     */
/* reference's cells_dict keyed by IntPoint.                                */
/* ------------------------------------------------------------------------ */
int32_t orc_real_to_int(double real_val, double step)
{
    return (int32_t)floor(real_val / step); /* cell.pyx:57 */
}

typedef struct {
    int64_t cap;      /* power of two */
    int32_t *kx, *ky, *kz;
    int64_t *slot;    /* -1 empty, else cell number */
    int64_t ncells;
    int64_t *start;   /* per cell: first position in `items` */
    int64_t *count;
    int64_t *items;   /* particle indices grouped by cell, index order */
} orc_cells;

static uint64_t cell_hash(int32_t x, int32_t y, int32_t z)
{
    uint64_t hsh = (uint64_t)(uint32_t)x * 0x9E3779B97F4A7C15ull;
    hsh ^= (uint64_t)(uint32_t)y * 0xC2B2AE3D27D4EB4Full + (hsh << 6) + (hsh >> 2);
    hsh ^= (uint64_t)(uint32_t)z * 0x165667B19E3779F9ull + (hsh << 6) + (hsh >> 2);
    return hsh;
}

static int64_t cells_find(const orc_cells *c, int32_t x, int32_t y, int32_t z)
{
    uint64_t mask = (uint64_t)c->cap - 1, s = cell_hash(x, y, z) & mask;
    for (;;) {
        int64_t v = c->slot[s];
        if (v < 0) return -1;
        if (c->kx[s] == x && c->ky[s] == y && c->kz[s] == z) return v;
        s = (s + 1) & mask;
    }
}

static int64_t cells_insert(orc_cells *c, int32_t x, int32_t y, int32_t z)
{
    uint64_t mask = (uint64_t)c->cap - 1, s = cell_hash(x, y, z) & mask;
    for (;;) {
        int64_t v = c->slot[s];
        if (v < 0) {
            c->slot[s] = c->ncells;
            c->kx[s] = x;
            c->ky[s] = y;
            c->kz[s] = z;
            return c->ncells++;
        }
        if (c->kx[s] == x && c->ky[s] == y && c->kz[s] == z) return v;
        s = (s + 1) & mask;
    }
}

static void cells_free(orc_cells *c)
{
    free(c->kx); free(c->ky); free(c->kz); free(c->slot);
    free(c->start); free(c->count); free(c->items);
    memset(c, 0, sizeof(*c));
}

/* Bin one array: particle i goes to cell find_cell_id(x_i) (cell.pyx:61-82). */
static int cells_build(orc_cells *c, const orc_arr *a, double cs)
{
    int64_t n = a->n, cap = 64, i;
    int64_t *cell_of;
    memset(c, 0, sizeof(*c));
    while (cap < 2 * (n + 1)) cap <<= 1;
    c->cap = cap;
    c->kx = (int32_t *)malloc(sizeof(int32_t) * cap);
    c->ky = (int32_t *)malloc(sizeof(int32_t) * cap);
    c->kz = (int32_t *)malloc(sizeof(int32_t) * cap);
    c->slot = (int64_t *)malloc(sizeof(int64_t) * cap);
    cell_of = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
    if (!c->kx || !c->ky || !c->kz || !c->slot || !cell_of) return -1;
    for (i = 0; i < cap; ++i) c->slot[i] = -1;
    for (i = 0; i < n; ++i)
        cell_of[i] = cells_insert(c, orc_real_to_int(a->x[i], cs),
                                  orc_real_to_int(a->y[i], cs),
                                  orc_real_to_int(a->z[i], cs));
    c->start = (int64_t *)calloc((size_t)c->ncells + 1, sizeof(int64_t));
    c->count = (int64_t *)calloc((size_t)c->ncells + 1, sizeof(int64_t));
    c->items = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
    if (!c->start || !c->count || !c->items) return -1;
    for (i = 0; i < n; ++i) c->count[cell_of[i]]++;
    for (i = 1; i < c->ncells; ++i) c->start[i] = c->start[i - 1] + c->count[i - 1];
    memset(c->count, 0, sizeof(int64_t) * ((size_t)c->ncells + 1));
    for (i = 0; i < n; ++i) {
        int64_t k = cell_of[i];
        c->items[c->start[k] + c->count[k]++] = i;
    }
    free(cell_of);
    return 0;
}

/* FixedDest neighbour query for destination i (nnps.pyx:209-242, :271-356,
 * :525-559).  Appends to buf; returns the new length. */
static int64_t query_one(const orc_cells *c, const orc_arr *src, double cs,
                         double px, double py, double pz, double radius,
                         int64_t *buf, int64_t len, int64_t cap)
{
    double radius2 = radius * radius;
    int32_t d = (int32_t)ceil(radius / cs); /* cell.pyx:1127-1129 */
    int32_t cx = orc_real_to_int(px, cs), cy = orc_real_to_int(py, cs),
            cz = orc_real_to_int(pz, cs);
    int32_t i, j, k;
    /* construct_immediate_neighbor_list: x outer, y, z inner (cell.pyx:105-114) */
    for (i = -d; i <= d; ++i)
        for (j = -d; j <= d; ++j)
            for (k = -d; k <= d; ++k) {
                int64_t cell = cells_find(c, cx + i, cy + j, cz + k), t;
                if (cell < 0) continue;
                for (t = 0; t < c->count[cell]; ++t) {
                    int64_t idx = c->items[c->start[cell] + t];
                    /* cPoint_distance2(src, dst) (point.pxd:35-37) */
                    double d2 = ((src->x[idx] - px) * (src->x[idx] - px) +
                                 (src->y[idx] - py) * (src->y[idx] - py)) +
                                (src->z[idx] - pz) * (src->z[idx] - pz);
                    if (d2 < radius2) {
                        if (len < cap) buf[len] = idx;
                        ++len;
                    }
                }
            }
    return len;
}

/* Build the neighbour cache of (src, dst): returns 0 or -1 on allocation
 * failure.  cell_size <= 0 is an error (the caller resolves it). */
int orc_nbrs_build(orc_nbrs *out, const orc_arr *dst, const orc_arr *src,
                   double cell_size, double radius_scale)
{
    orc_cells c;
    int64_t i, total = 0;
    int same = (dst == src) || (dst->x == src->x);
    if (cell_size <= 0) return -1;
    if (cells_build(&c, src, cell_size)) return -1;
    out->n_dst = dst->n;
    out->off = (int64_t *)malloc(sizeof(int64_t) * (dst->n + 1));
    if (!out->off) return -1;
    /* pass 1: counts */
#pragma omp parallel for schedule(dynamic, 256) num_threads(g_threads)
    for (i = 0; i < dst->n; ++i)
        out->off[i + 1] = query_one(&c, src, cell_size, dst->x[i], dst->y[i],
                                    dst->z[i], dst->h[i] * radius_scale,
                                    NULL, 0, 0);
    out->off[0] = 0;
    for (i = 0; i < dst->n; ++i) {
        total += out->off[i + 1];
        out->off[i + 1] = total;
    }
    out->nbr = (int64_t *)malloc(sizeof(int64_t) * (total + 1));
    if (!out->nbr) return -1;
#pragma omp parallel for schedule(dynamic, 256) num_threads(g_threads)
    for (i = 0; i < dst->n; ++i) {
        int64_t b = out->off[i], len = out->off[i + 1] - b, j;
        int64_t *buf = out->nbr + b;
        query_one(&c, src, cell_size, dst->x[i], dst->y[i], dst->z[i],
                  dst->h[i] * radius_scale, buf, 0, len);
        /* keep the self index last (nnps.pyx:612-618) */
        if (same)
            for (j = 0; j < len - 1; ++j)
                if (buf[j] == i) {
                    buf[j] = buf[len - 1];
                    buf[len - 1] = i;
                    break;
                }
    }
    cells_free(&c);
    return 0;
}

/* VarHNbrParticleLocator._update_cache (nnps.pyx:654-749): the row of
 * destination k is its FixedDest query without the self swap, followed by
 * every source j (ascending) whose own query over the destination array,
 * with radius h_j * radius_scale, finds k and that the row does not hold
 * yet.  The union is the symmetric set: d2 < (h_k rs)^2 or d2 < (h_j rs)^2.
 * Returns 0 or -1 on allocation failure. */
int orc_nbrs_build_varh(orc_nbrs *out, const orc_arr *dst, const orc_arr *src,
                        double cell_size, double radius_scale)
{
    orc_cells cs_src, cd;
    int64_t i, j, total = 0, rtotal = 0;
    int64_t *foff, *fnbr, *roff, *rnbr, *eoff, *ecnt, *enbr;
    if (cell_size <= 0) return -1;
    if (cells_build(&cs_src, src, cell_size)) return -1;
    if (cells_build(&cd, dst, cell_size)) return -1;
    foff = (int64_t *)calloc((size_t)dst->n + 1, sizeof(int64_t));
    roff = (int64_t *)calloc((size_t)src->n + 1, sizeof(int64_t));
    if (!foff || !roff) return -1;
    /* forward rows: FixedDest queries of dst from src */
    for (i = 0; i < dst->n; ++i)
        foff[i + 1] = foff[i] + query_one(&cs_src, src, cell_size, dst->x[i], dst->y[i],
                                          dst->z[i], dst->h[i] * radius_scale, NULL, 0, 0);
    fnbr = (int64_t *)malloc(sizeof(int64_t) * (foff[dst->n] + 1));
    if (!fnbr) return -1;
    for (i = 0; i < dst->n; ++i)
        query_one(&cs_src, src, cell_size, dst->x[i], dst->y[i], dst->z[i],
                  dst->h[i] * radius_scale, fnbr + foff[i], 0, foff[i + 1] - foff[i]);
    /* reverse rows: queries of src from dst with the source's radius */
    for (j = 0; j < src->n; ++j)
        roff[j + 1] = roff[j] + query_one(&cd, dst, cell_size, src->x[j], src->y[j],
                                          src->z[j], src->h[j] * radius_scale, NULL, 0, 0);
    rtotal = roff[src->n];
    rnbr = (int64_t *)malloc(sizeof(int64_t) * (rtotal + 1));
    if (!rnbr) return -1;
    for (j = 0; j < src->n; ++j)
        query_one(&cd, dst, cell_size, src->x[j], src->y[j], src->z[j],
                  src->h[j] * radius_scale, rnbr + roff[j], 0, roff[j + 1] - roff[j]);
    /* candidates appended to row k: the j (ascending) whose reverse row has k */
    eoff = (int64_t *)calloc((size_t)dst->n + 1, sizeof(int64_t));
    ecnt = (int64_t *)calloc((size_t)dst->n + 1, sizeof(int64_t));
    enbr = (int64_t *)malloc(sizeof(int64_t) * (rtotal + 1));
    if (!eoff || !ecnt || !enbr) return -1;
    for (i = 0; i < rtotal; ++i) eoff[rnbr[i] + 1]++;
    for (i = 0; i < dst->n; ++i) eoff[i + 1] += eoff[i];
    for (j = 0; j < src->n; ++j)
        for (i = roff[j]; i < roff[j + 1]; ++i) {
            int64_t k = rnbr[i];
            enbr[eoff[k] + ecnt[k]++] = j;
        }
    /* merge: forward row, then the candidates it does not hold yet */
    out->n_dst = dst->n;
    out->off = (int64_t *)malloc(sizeof(int64_t) * (dst->n + 1));
    out->nbr = (int64_t *)malloc(sizeof(int64_t) * (foff[dst->n] + rtotal + 1));
    if (!out->off || !out->nbr) return -1;
    out->off[0] = 0;
    for (i = 0; i < dst->n; ++i) {
        int64_t b = foff[i], e = foff[i + 1], t, u;
        for (t = b; t < e; ++t) out->nbr[total++] = fnbr[t];
        for (t = eoff[i]; t < eoff[i + 1]; ++t) {
            int present = 0;
            for (u = out->off[i]; u < total; ++u)
                if (out->nbr[u] == enbr[t]) { present = 1; break; }
            if (!present) out->nbr[total++] = enbr[t];
        }
        out->off[i + 1] = total;
    }
    free(foff); free(fnbr); free(roff); free(rnbr);
    free(eoff); free(ecnt); free(enbr);
    cells_free(&cs_src);
    cells_free(&cd);
    return 0;
}

void orc_nbrs_free(orc_nbrs *nb)
{
    free(nb->off);
    free(nb->nbr);
    nb->off = nb->nbr = NULL;
}

/* Flat CSR export for the Python side: returns total pairs; writes only when
 * cap suffices.  offsets must hold n_dst + 1 entries. */
int64_t orc_nnps_csr(const orc_arr *dst, const orc_arr *src, int same,
                     double cell_size, double radius_scale, int64_t *offsets,
                     int64_t *nbrs, int64_t cap, int varh)
{
    orc_nbrs nb;
    int64_t total, i;
    const orc_arr *s = same ? dst : src;
    if (varh ? orc_nbrs_build_varh(&nb, dst, s, cell_size, radius_scale)
             : orc_nbrs_build(&nb, dst, s, cell_size, radius_scale)) return -1;
    total = nb.off[dst->n];
    for (i = 0; i <= dst->n; ++i) offsets[i] = nb.off[i];
    if (total <= cap) memcpy(nbrs, nb.nbr, sizeof(int64_t) * total);
    orc_nbrs_free(&nb);
    return total;
}

/* Point query with an explicit cell size (NbrParticleLocatorBase.
 * get_nearest_particles_to_point, nnps.pyx:209-242): the reference's test
 * fixtures bin with a fixed cell size.  exclude_index < 0 keeps all. */
int64_t orc_point_query(const orc_arr *src, double cell_size, double px,
                        double py, double pz, double radius,
                        int64_t exclude_index, int64_t *out, int64_t cap)
{
    orc_cells c;
    int64_t len, i, k = 0;
    int64_t *tmp;
    if (cells_build(&c, src, cell_size)) return -1;
    len = query_one(&c, src, cell_size, px, py, pz, radius, NULL, 0, 0);
    tmp = (int64_t *)malloc(sizeof(int64_t) * (len + 1));
    query_one(&c, src, cell_size, px, py, pz, radius, tmp, 0, len);
    for (i = 0; i < len; ++i)
        if (tmp[i] != exclude_index) {
            if (k < cap) out[k] = tmp[i];
            ++k;
        }
    free(tmp);
    cells_free(&c);
    return k;
}

/* ------------------------------------------------------------------------ */
/* SPH functions, each evaluated the way SPHCalc.sph drives them:           */
/* outputs zeroed, then += one eval per source (sph_calc.pyx:223-276), and  */
/* a non-LocalReal destination written 0 (sph_func.pyx:211-234).            */
/* ------------------------------------------------------------------------ */
static int is_local(const orc_arr *a, int64_t i)
{
    return a->tag == NULL || a->tag[i] == 0;
}

/* SPHRho: density_funcs.pyx:38-84 (hks off, no corrections; CSPH with
 * dnr == 0 so no normalisation, sph_func.pyx:466-484). */
void orc_sphrho(int dim, const orc_arr *dst, int nsrc, const orc_arr *const *srcs,
                const orc_nbrs *const *nb, double *out)
{
    int64_t i;
#pragma omp parallel for schedule(dynamic, 256) num_threads(g_threads)
    for (i = 0; i < dst->n; ++i) {
        int s;
        out[i] = 0.0;
        for (s = 0; s < nsrc; ++s) {
            const orc_arr *src = srcs[s];
            double nr = 0.0;
            int64_t t;
            if (!is_local(dst, i)) { out[i] = 0; continue; }
            for (t = nb[s]->off[i]; t < nb[s]->off[i + 1]; ++t) {
                int64_t j = nb[s]->nbr[t];
                double hab = 0.5 * (dst->h[i] + src->h[j]);
                double w = orc_cubic_w(dst->x[i], dst->y[i], dst->z[i], src->x[j],
                                       src->y[j], src->z[j], hab, dim);
                nr += w * src->m[j];
            }
            out[i] += nr;
        }
    }
}

/* TaitEquation: eos_funcs.pyx:84-116 */
void orc_tait(const orc_arr *a, const double *rho, double co, double ro,
              double gamma, double *p, double *cs)
{
    int64_t i;
    double B = co * co * ro / gamma;
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (i = 0; i < a->n; ++i) {
        if (!is_local(a, i)) { p[i] = cs[i] = 0; continue; }
        {
            double ratio = rho[i] / ro;
            double gamma2 = 0.5 * (gamma - 1.0);
            double tmp = pow(ratio, gamma);
            p[i] = (tmp - 1.0) * B;
            cs[i] = pow(ratio, gamma2) * co;
        }
    }
}

/* MomentumEquation: pressure_funcs.pyx:154-245 (hks off).  num_outputs = dim
 * components are written; out3 may be NULL when dim < 3. */
void orc_momentum(int dim, const orc_arr *dst, int nsrc, const orc_arr *const *srcs,
                  const orc_nbrs *const *nb, double alpha, double beta,
                  double eta, double *o1, double *o2, double *o3)
{
    int64_t i;
#pragma omp parallel for schedule(dynamic, 256) num_threads(g_threads)
    for (i = 0; i < dst->n; ++i) {
        int s;
        o1[i] = 0; if (dim > 1) o2[i] = 0; if (dim > 2) o3[i] = 0;
        for (s = 0; s < nsrc; ++s) {
            const orc_arr *src = srcs[s];
            double nr[3] = {0, 0, 0};
            int64_t t;
            if (!is_local(dst, i)) {
                o1[i] = 0; if (dim > 1) o2[i] = 0; if (dim > 2) o3[i] = 0;
                continue;
            }
            for (t = nb[s]->off[i]; t < nb[s]->off[i + 1]; ++t) {
                int64_t j = nb[s]->nbr[t];
                double ha = dst->h[i], hb = src->h[j];
                double hab = 0.5 * (ha + hb);
                double ca = dst->cs[i], cb = src->cs[j];
                double rx = dst->x[i] - src->x[j], ry = dst->y[i] - src->y[j],
                       rz = dst->z[i] - src->z[j];
                double vx = dst->u[i] - src->u[j], vy = dst->v[i] - src->v[j],
                       vz = dst->w[i] - src->w[j];
                double dot = vx * rx + vy * ry + vz * rz; /* cPoint_dot */
                double Pa = dst->p[i], rhoa = dst->rho[i];
                double Pb = src->p[j], rhob = src->rho[j], mb = src->m[j];
                double tmp = Pa / (rhoa * rhoa) + Pb / (rhob * rhob);
                double piab = 0, g[3];
                if (dot < 0) {
                    double cab = 0.5 * (ca + cb);
                    double rhoab = 0.5 * (rhoa + rhob);
                    double mu = hab * dot;
                    mu /= ((rx * rx + ry * ry + rz * rz) + eta * eta * hab * hab);
                    piab = -alpha * cab * mu + beta * mu * mu;
                    piab /= rhoab;
                }
                tmp += piab;
                tmp *= -mb;
                orc_cubic_grad(dst->x[i], dst->y[i], dst->z[i], src->x[j],
                               src->y[j], src->z[j], hab, dim, g);
                nr[0] += tmp * g[0];
                if (dim > 1) {
                    nr[1] += tmp * g[1];
                    if (dim > 2) nr[2] += tmp * g[2];
                }
            }
            o1[i] += nr[0];
            if (dim > 1) o2[i] += nr[1];
            if (dim > 2) o3[i] += nr[2];
        }
    }
}

/* XSPHCorrection: xsph_funcs.pyx:31-88 (CSPH, dnr == 0). */
void orc_xsph(int dim, const orc_arr *dst, int nsrc, const orc_arr *const *srcs,
              const orc_nbrs *const *nb, double eps, double *o1, double *o2,
              double *o3)
{
    int64_t i;
#pragma omp parallel for schedule(dynamic, 256) num_threads(g_threads)
    for (i = 0; i < dst->n; ++i) {
        int s;
        o1[i] = 0; if (dim > 1) o2[i] = 0; if (dim > 2) o3[i] = 0;
        for (s = 0; s < nsrc; ++s) {
            const orc_arr *src = srcs[s];
            double nr[3] = {0, 0, 0};
            int64_t t;
            if (!is_local(dst, i)) {
                o1[i] = 0; if (dim > 1) o2[i] = 0; if (dim > 2) o3[i] = 0;
                continue;
            }
            for (t = nb[s]->off[i]; t < nb[s]->off[i + 1]; ++t) {
                int64_t j = nb[s]->nbr[t];
                double hab = 0.5 * (dst->h[i] + src->h[j]);
                double rhoab = 0.5 * (src->rho[j] + dst->rho[i]);
                double vbax = src->u[j] - dst->u[i], vbay = src->v[j] - dst->v[i],
                       vbaz = src->w[j] - dst->w[i];
                double mb = src->m[j];
                double w = orc_cubic_w(dst->x[i], dst->y[i], dst->z[i], src->x[j],
                                       src->y[j], src->z[j], hab, dim);
                double temp = mb * w / rhoab;
                nr[0] += temp * vbax * eps;
                if (dim > 1) {
                    nr[1] += temp * vbay * eps;
                    if (dim > 2) nr[2] += temp * vbaz * eps;
                }
            }
            o1[i] += nr[0];
            if (dim > 1) o2[i] += nr[1];
            if (dim > 2) o3[i] += nr[2];
        }
    }
}

/* SPHDensityRate: density_funcs.pyx:89-158 (hks off, no corrections):
 * drho_a/dt = sum_b m_b (v_a - v_b) . grad_a W(r_a - r_b, h_ab). */
void orc_density_rate(int dim, const orc_arr *dst, int nsrc, const orc_arr *const *srcs,
                      const orc_nbrs *const *nb, double *out)
{
    int64_t i;
#pragma omp parallel for schedule(dynamic, 256) num_threads(g_threads)
    for (i = 0; i < dst->n; ++i) {
        int s;
        out[i] = 0.0;
        if (!is_local(dst, i)) continue;
        for (s = 0; s < nsrc; ++s) {
            const orc_arr *src = srcs[s];
            double nr = 0.0;
            int64_t t;
            for (t = nb[s]->off[i]; t < nb[s]->off[i + 1]; ++t) {
                int64_t j = nb[s]->nbr[t];
                double hab = 0.5 * (dst->h[i] + src->h[j]);
                double vx = dst->u[i] - src->u[j], vy = dst->v[i] - src->v[j],
                       vz = dst->w[i] - src->w[j];
                double g[3];
                orc_cubic_grad(dst->x[i], dst->y[i], dst->z[i], src->x[j], src->y[j],
                               src->z[j], hab, dim, g);
                /* cPoint_dot(vel, grad) * m_b */
                nr += (vx * g[0] + vy * g[1] + vz * g[2]) * src->m[j];
            }
            out[i] += nr;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* The WCSPH scheme used by the elliptical-drop and dam-break set-ups, with */
/* the reference integrator's bookkeeping.  Operation order (one calc per   */
/* destination array, arrays in list order):                                */
/*   rho  : SPHRho on {F,S} from {F,S}              (non-integrating)       */
/*   eos  : TaitEquation on {F,S}                   (non-integrating)       */
/*   mom  : MomentumEquation on F from {F,S}        (integrating u,v[,w])   */
/*   grav : GravityForce on F  (optional)           (integrating u,v[,w])   */
/*   step : PositionStepping on F                   (integrating x,y[,z])   */
/*   xsph : XSPHCorrection on F from F (optional)   (integrating x,y[,z])   */
/* ------------------------------------------------------------------------ */
typedef struct {
    int32_t dim;
    int32_t integrator;  /* 0 Euler, 1 RK2 */
    double dt;
    double co, ro, gamma;
    double alpha, beta, eta;
    int32_t has_gravity;
    double g[3];
    int32_t has_xsph;
    double eps;
    double cell_size;    /* <= 0: max_radius_scale * max h (cell.pyx:906-912) */
    double radius_scale; /* kernel.radius() == 2 */
} orc_scheme;

#define ORC_MAXARR 8

typedef struct {
    /* per fluid array: k arrays per stage for mom(3), grav(3), step(3), xsph(3)
     * and the shared initial arrays _x_0.._w_0 */
    double *k[2][4][3];
    double *x0[3], *v0[3];
} fluid_work;

static double scheme_cell_size(const orc_scheme *sc, int narr, orc_arr *const *arrs)
{
    double mh = 0;
    int a;
    int64_t i;
    if (sc->cell_size > 0) return sc->cell_size;
    for (a = 0; a < narr; ++a)
        for (i = 0; i < arrs[a]->n; ++i)
            if (arrs[a]->h[i] > mh) mh = arrs[a]->h[i];
    mh *= sc->radius_scale;
    return mh == 0.0 ? 1.0 : mh;
}

static double *pos_comp(orc_arr *a, int c) { return c == 0 ? a->x : c == 1 ? a->y : a->z; }
static double *vel_comp(orc_arr *a, int c) { return c == 0 ? a->u : c == 1 ? a->v : a->w; }

/* Integrator.eval for the scheme at the current state: writes rho/p/cs on
 * every array and the k arrays of stage `st` on fluid arrays. */
static int scheme_eval(const orc_scheme *sc, int narr, orc_arr *const *arrs,
                       fluid_work *fw, int st)
{
    double cs = scheme_cell_size(sc, narr, arrs);
    int dim = sc->dim, a, b, c;
    orc_nbrs nb[ORC_MAXARR][ORC_MAXARR];
    memset(nb, 0, sizeof(nb));
    /* neighbour caches for every (src, dst) pair that an operation uses */
    for (a = 0; a < narr; ++a)
        for (b = 0; b < narr; ++b)
            if (orc_nbrs_build(&nb[a][b], arrs[a], arrs[b], cs, sc->radius_scale))
                return -1;
    /* rho on F,S from F,S */
    for (a = 0; a < narr; ++a) {
        const orc_arr *srcs[ORC_MAXARR];
        const orc_nbrs *nbs[ORC_MAXARR];
        for (b = 0; b < narr; ++b) { srcs[b] = arrs[b]; nbs[b] = &nb[a][b]; }
        orc_sphrho(dim, arrs[a], narr, srcs, nbs, arrs[a]->rho);
    }
    /* eos on F,S */
    for (a = 0; a < narr; ++a)
        orc_tait(arrs[a], arrs[a]->rho, sc->co, sc->ro, sc->gamma, arrs[a]->p,
                 arrs[a]->cs);
    /* integrating calcs on fluid arrays */
    for (a = 0; a < narr; ++a) {
        const orc_arr *srcs[ORC_MAXARR], *fsrcs[ORC_MAXARR];
        const orc_nbrs *nbs[ORC_MAXARR], *fnbs[ORC_MAXARR];
        int nf = 0;
        int64_t i;
        double **k;
        if (arrs[a]->type != 0) continue;
        for (b = 0; b < narr; ++b) {
            srcs[b] = arrs[b];
            nbs[b] = &nb[a][b];
            if (arrs[b]->type == 0) { fsrcs[nf] = arrs[b]; fnbs[nf++] = &nb[a][b]; }
        }
        k = fw[a].k[st][0];
        orc_momentum(dim, arrs[a], narr, srcs, nbs, sc->alpha, sc->beta, sc->eta,
                     k[0], k[1], k[2]);
        if (sc->has_gravity) {
            k = fw[a].k[st][1];
            for (i = 0; i < arrs[a]->n; ++i)
                for (c = 0; c < dim; ++c)
                    k[c][i] = is_local(arrs[a], i) ? sc->g[c] : 0.0;
        }
        k = fw[a].k[st][2];
        for (i = 0; i < arrs[a]->n; ++i)
            for (c = 0; c < dim; ++c)
                k[c][i] = is_local(arrs[a], i) ? vel_comp(arrs[a], c)[i] : 0.0;
        if (sc->has_xsph) {
            k = fw[a].k[st][3];
            orc_xsph(dim, arrs[a], nf, fsrcs, fnbs, sc->eps, k[0], k[1], k[2]);
        }
    }
    for (a = 0; a < narr; ++a)
        for (b = 0; b < narr; ++b) orc_nbrs_free(&nb[a][b]);
    return 0;
}

/* One integrator step of the scheme (EulerIntegrator.integrate,
 * integrator.py:561-573, or RK2Integrator.integrate, :621-645). */
int orc_scheme_step(const orc_scheme *sc, int narr, orc_arr *const *arrs)
{
    fluid_work fw[ORC_MAXARR];
    int a, st, c, calc, dim = sc->dim, rc = 0;
    int nst = sc->integrator == 1 ? 2 : 1;
    int64_t i;
    if (narr > ORC_MAXARR) return -1;
    memset(fw, 0, sizeof(fw));
    for (a = 0; a < narr; ++a) {
        int64_t n = arrs[a]->n;
        if (arrs[a]->type != 0) continue;
        for (st = 0; st < nst; ++st)
            for (calc = 0; calc < 4; ++calc)
                for (c = 0; c < 3; ++c)
                    fw[a].k[st][calc][c] = (double *)calloc((size_t)n + 1, sizeof(double));
        for (c = 0; c < 3; ++c) {
            fw[a].x0[c] = (double *)malloc(sizeof(double) * (n + 1));
            fw[a].v0[c] = (double *)malloc(sizeof(double) * (n + 1));
        }
        /* set_initial_arrays (integrator.py:339-367) */
        for (c = 0; c < dim; ++c) {
            memcpy(fw[a].x0[c], pos_comp(arrs[a], c), sizeof(double) * n);
            memcpy(fw[a].v0[c], vel_comp(arrs[a], c), sizeof(double) * n);
        }
    }
    if (scheme_eval(sc, narr, arrs, fw, 0)) { rc = -1; goto done; }
    if (nst == 1) {
        /* EulerIntegrator.final_step per integrating calc, in order:
         * p = p_0 + k1*dt; p_0 = p (integrator.py:534-551) */
        for (a = 0; a < narr; ++a) {
            if (arrs[a]->type != 0) continue;
            for (calc = 0; calc < 4; ++calc) {
                int isvel = calc < 2;
                if (calc == 1 && !sc->has_gravity) continue;
                if (calc == 3 && !sc->has_xsph) continue;
                for (c = 0; c < dim; ++c) {
                    double *p0 = isvel ? fw[a].v0[c] : fw[a].x0[c];
                    double *p = isvel ? vel_comp(arrs[a], c) : pos_comp(arrs[a], c);
                    double *k1 = fw[a].k[0][calc][c];
                    for (i = 0; i < arrs[a]->n; ++i) {
                        double up = p0[i] + k1[i] * sc->dt;
                        p[i] = up;
                        p0[i] = up;
                    }
                }
            }
        }
    } else {
        /* do_step: reset_current_arrays then p = p + k1*dt per calc
         * (integrator.py:369-392, :472-514) */
        for (a = 0; a < narr; ++a) {
            if (arrs[a]->type != 0) continue;
            for (c = 0; c < dim; ++c) {
                memcpy(pos_comp(arrs[a], c), fw[a].x0[c], sizeof(double) * arrs[a]->n);
                memcpy(vel_comp(arrs[a], c), fw[a].v0[c], sizeof(double) * arrs[a]->n);
            }
            for (calc = 0; calc < 4; ++calc) {
                int isvel = calc < 2;
                if (calc == 1 && !sc->has_gravity) continue;
                if (calc == 3 && !sc->has_xsph) continue;
                for (c = 0; c < dim; ++c) {
                    double *p = isvel ? vel_comp(arrs[a], c) : pos_comp(arrs[a], c);
                    double *k1 = fw[a].k[0][calc][c];
                    for (i = 0; i < arrs[a]->n; ++i) p[i] = p[i] + k1[i] * sc->dt;
                }
            }
        }
        if (scheme_eval(sc, narr, arrs, fw, 1)) { rc = -1; goto done; }
        /* RK2Integrator.final_step: p = p_0 + 0.5*dt*(k1+k2); p_0 = p
         * (integrator.py:597-619) */
        for (a = 0; a < narr; ++a) {
            if (arrs[a]->type != 0) continue;
            for (calc = 0; calc < 4; ++calc) {
                int isvel = calc < 2;
                if (calc == 1 && !sc->has_gravity) continue;
                if (calc == 3 && !sc->has_xsph) continue;
                for (c = 0; c < dim; ++c) {
                    double *p0 = isvel ? fw[a].v0[c] : fw[a].x0[c];
                    double *p = isvel ? vel_comp(arrs[a], c) : pos_comp(arrs[a], c);
                    double *k1 = fw[a].k[0][calc][c], *k2 = fw[a].k[1][calc][c];
                    double hdt = 0.5 * sc->dt;
                    for (i = 0; i < arrs[a]->n; ++i) {
                        double up = p0[i] + hdt * (k1[i] + k2[i]);
                        p[i] = up;
                        p0[i] = up;
                    }
                }
            }
        }
    }
done:
    for (a = 0; a < narr; ++a) {
        for (st = 0; st < 2; ++st)
            for (calc = 0; calc < 4; ++calc)
                for (c = 0; c < 3; ++c) free(fw[a].k[st][calc][c]);
        for (c = 0; c < 3; ++c) { free(fw[a].x0[c]); free(fw[a].v0[c]); }
    }
    return rc;
}

/* Evaluate only the right-hand side of the scheme at the current state:
 * rho/p/cs on every array and, for fluid array `fa`, the momentum
 * acceleration (a), the XSPH velocity correction (dv).  Used by the
 * per-stage parity tests. */
int orc_scheme_rhs(const orc_scheme *sc, int narr, orc_arr *const *arrs, int fa,
                   double *ax, double *ay, double *az, double *dvx, double *dvy,
                   double *dvz)
{
    fluid_work fw[ORC_MAXARR];
    int a, c, calc;
    int64_t n;
    if (narr > ORC_MAXARR || fa < 0 || fa >= narr) return -1;
    memset(fw, 0, sizeof(fw));
    for (a = 0; a < narr; ++a) {
        if (arrs[a]->type != 0) continue;
        for (calc = 0; calc < 4; ++calc)
            for (c = 0; c < 3; ++c)
                fw[a].k[0][calc][c] = (double *)calloc((size_t)arrs[a]->n + 1, sizeof(double));
    }
    if (scheme_eval(sc, narr, arrs, fw, 0)) return -1;
    n = arrs[fa]->n;
    memcpy(ax, fw[fa].k[0][0][0], sizeof(double) * n);
    memcpy(ay, fw[fa].k[0][0][1], sizeof(double) * n);
    memcpy(az, fw[fa].k[0][0][2], sizeof(double) * n);
    memcpy(dvx, fw[fa].k[0][3][0], sizeof(double) * n);
    memcpy(dvy, fw[fa].k[0][3][1], sizeof(double) * n);
    memcpy(dvz, fw[fa].k[0][3][2], sizeof(double) * n);
    for (a = 0; a < narr; ++a)
        for (calc = 0; calc < 4; ++calc)
            for (c = 0; c < 3; ++c) free(fw[a].k[0][calc][c]);
    return 0;
}

/* FP64 conservation diagnostics of a set of arrays (no reference
 * counterpart; the reference prints none).  out[0] mass, out[1..3] linear
 * momentum, out[4] kinetic energy, out[5] gravitational potential energy
 * -sum m g.x, out[6] Tait elastic energy sum m e(rho) with
 * e(rho) = B[(rho/ro)^(gamma-1)/((gamma-1) ro) + 1/rho - gamma/((gamma-1) ro)],
 * out[7] the total of 4..6.  Only LocalReal fluid particles count. */
void orc_conservation(int narr, orc_arr *const *arrs, const double g[3], double co,
                      double ro, double gamma, double *out)
{
    int a, c;
    int64_t i;
    double B = co * co * ro / gamma;
    for (c = 0; c < 8; ++c) out[c] = 0;
    for (a = 0; a < narr; ++a) {
        const orc_arr *A = arrs[a];
        if (A->type != 0) continue;
        for (i = 0; i < A->n; ++i) {
            double m = A->m[i], e;
            if (!is_local(A, i)) continue;
            out[0] += m;
            out[1] += m * A->u[i];
            out[2] += m * A->v[i];
            out[3] += m * A->w[i];
            out[4] += 0.5 * m * (A->u[i] * A->u[i] + A->v[i] * A->v[i] + A->w[i] * A->w[i]);
            out[5] -= m * (g[0] * A->x[i] + g[1] * A->y[i] + g[2] * A->z[i]);
            e = B * (pow(A->rho[i] / ro, gamma - 1.0) / ((gamma - 1.0) * ro) +
                     1.0 / A->rho[i] - gamma / ((gamma - 1.0) * ro));
            out[6] += m * e;
        }
    }
    out[7] = out[4] + out[5] + out[6];
}
