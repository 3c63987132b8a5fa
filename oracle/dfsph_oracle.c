/*
 * dfsph_oracle.c -- FP64 CPU restatement of the reference DFSPH time step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP path; never linked into
 * the product).  Every function cites the reference file:line it restates;
 * paths are relative to the reference repository root.
 *
 * Mode OR_MODE_GPU restates SPlisHSPlasH/DFSPH/DFSPH_core_cuda.cu, the
 * DFSPHCUDA::step() sequence of SPlisHSPlasH/DFSPH/DFSPH_CUDA.cpp:82-794 and
 * the time-step bookkeeping of DFSPH_c_arrays_structure.cpp:1100-1118.
 * Mode OR_MODE_CPU restates SPlisHSPlasH/DFSPH/TimeStepDFSPH.cpp:29-747
 * (warm-start formulas and accumulation points; constant dt assumed, as in
 * that file, whose CFL call is disabled at :63-65).
 */
#include "dfsph_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846       /* DFSPH_define_c.h:5 */
#endif
#define OR_EPS 1.0e-5                     /* m_eps, DFSPH_define_cuda.h:7 */
#define OR_TABLE_RES 10000u               /* FluidModel.h:43             */
#define OR_CELL_MARGIN 1.001              /* grid cell = 1.001 h_k        */

struct or_state {
    or_params p;
    double hk, hk2, k_coef, l_coef, W0;
    double* tabW;   /* [RES]   */
    double* tabG;   /* [RES+1] */
    double inv_step;

    uint32_t n, cap;
    double *x, *v, *m, *rho, *alpha, *radv, *kappa, *kappaV, *acc;
    uint32_t* id;

    uint32_t nbn;
    double *xb, *psi;

    /* rigid bodies: their particles follow the nbw wall particles in xb / psi
     * (vector_dynamic_bodies_data, DFSPH_c_arrays_structure.cpp:897-905) */
    uint32_t nbw, nbody, ns;
    struct or_body* body;
    double *x0, *psi0, *vb, *sF;    /* body frame (3 ns), psi (ns), boundary velocity (3 nbn), reactions (3 ns) */
    uint32_t* sbody;

    uint32_t *nf, *nbc, *nbr;
    uint32_t max_seen, overflow;

    double h, h_future, h_past, h_ratio_to_past, h_ratio_to_past2, time;
    double bbmin[3], bbmax[3];

    /* solver restriction (GPU mode): DFSPHCData::restriction_mode,
     * count_active, count_active_neighbors, true_particle_count */
    int rmode;
    uint32_t r_active, r_active_nbrs;
    long long true_count;
    uint32_t* tag;
};

typedef struct or_body {
    int type;
    double mass, I[3], c[3], v[3], q[4], w[3], F[3], T[3], c0[3], axis[3], amp, omega;
    double clock;              /* the body's time: the sum of the steps' h */
    uint32_t first, count;
} or_body;

#define OR_TAG_ACTIVE 10000000u            /* cuda_compilation_center.cu:6-9 */
#define OR_TAG_ACTIVE_NEIGHBORS 20000000u
#define OR_TAG_UNTAGGED 40000000u

/* the restriction tests of the kernels (DFSPH_core_cuda.cu:225-231 "source"
 * kernels: active + active neighbors; :500-506, 3052-3058, 3131-3137 "compute"
 * kernels: active only) */
static inline int r_compute(const or_state* s, uint32_t i)
{
    if (s->rmode == 1) return s->tag[i] == OR_TAG_ACTIVE;
    if (s->rmode == 2) return i < s->r_active;
    return 1;
}
static inline int r_source(const or_state* s, uint32_t i)
{
    if (s->rmode == 1) return s->tag[i] == OR_TAG_ACTIVE || s->tag[i] == OR_TAG_ACTIVE_NEIGHBORS;
    if (s->rmode == 2) return i < s->r_active + s->r_active_nbrs;
    return 1;
}
/* error average: true_particle_count when set (DFSPH_core_cuda.cu:791-795, 1339-1343) */
static inline double r_count(const or_state* s)
{
    return s->true_count >= 0 ? (double)s->true_count : (double)s->n;
}

/* ------------------------------------------------------------------ */
/* A3: cubic spline kernel (SPH_kernels.h:20-100, SPHKernels.h:13-92)  */
/* ------------------------------------------------------------------ */
static double cubic_W(double r, double hk)
{
    const double k = 8.0 / (M_PI * hk * hk * hk);
    const double q = r / hk;
    if (q <= 1.0) {
        if (q <= 0.5) {
            const double q2 = q * q;
            return k * (6.0 * q2 * q - 6.0 * q2 + 1.0);
        }
        const double f = 1.0 - q;
        return k * (2.0 * f * f * f);
    }
    return 0.0;
}

static void cubic_gradW(const double* r3, double hk, double* out)
{
    const double l = 48.0 / (M_PI * hk * hk * hk);
    const double rl = sqrt(r3[0] * r3[0] + r3[1] * r3[1] + r3[2] * r3[2]);
    const double q = rl / hk;
    out[0] = out[1] = out[2] = 0.0;
    if (q <= 1.0 && rl > 1.0e-6) {
        const double gq = 1.0 / (rl * hk);
        double c;
        if (q <= 0.5) c = l * q * (3.0 * q - 2.0) * gq;
        else { const double f = 1.0 - q; c = -l * f * f * gq; }
        out[0] = c * r3[0]; out[1] = c * r3[1]; out[2] = c * r3[2];
    }
}

/* PrecomputedKernel (SPHKernels.h:402-476).  The reference reads m_W[res]
 * one past its array when r == h exactly (:441 vs :405); here that entry
 * is taken as W(h) = 0. */
static void build_table(or_state* s)
{
    const double step = s->hk / (double)OR_TABLE_RES;
    s->inv_step = 1.0 / step;
    s->tabW = (double*)calloc(OR_TABLE_RES + 1, sizeof(double));
    s->tabG = (double*)calloc(OR_TABLE_RES + 1, sizeof(double));
    for (uint32_t i = 0; i < OR_TABLE_RES; ++i) {
        const double px = step * (double)i;
        s->tabW[i] = cubic_W(px, s->hk);
        if (px > 1.0e-9) {
            double r3[3] = {px, 0.0, 0.0}, g[3];
            cubic_gradW(r3, s->hk, g);
            s->tabG[i] = g[0] / px;
        } else s->tabG[i] = 0.0;
    }
    s->tabW[OR_TABLE_RES] = 0.0;
    s->tabG[OR_TABLE_RES] = 0.0;
}

static inline double kW(const or_state* s, const double* d)
{
    const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (!s->p.kernel_table) return cubic_W(sqrt(r2), s->hk);
    if (r2 <= s->hk2) {
        uint32_t pos = (uint32_t)(sqrt(r2) * s->inv_step);
        if (pos > OR_TABLE_RES) pos = OR_TABLE_RES;
        return s->tabW[pos];
    }
    return 0.0;
}

static inline void kG(const or_state* s, const double* d, double* g)
{
    if (!s->p.kernel_table) { cubic_gradW(d, s->hk, g); return; }
    const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    g[0] = g[1] = g[2] = 0.0;
    if (r2 <= s->hk2) {
        uint32_t pos = (uint32_t)(sqrt(r2) * s->inv_step);
        if (pos > OR_TABLE_RES) pos = OR_TABLE_RES;
        const double c = s->tabG[pos];
        g[0] = c * d[0]; g[1] = c * d[1]; g[2] = c * d[2];
    }
}

double or_kernel_W(double r, double hk, int table)
{
    if (!table) return cubic_W(r, hk);
    or_state s; memset(&s, 0, sizeof s);
    s.hk = hk; s.hk2 = hk * hk; s.p.kernel_table = 1;
    build_table(&s);
    double d[3] = {r, 0.0, 0.0};
    double w = kW(&s, d);
    free(s.tabW); free(s.tabG);
    return w;
}

void or_kernel_gradW(const double* r3, double hk, int table, double* out3)
{
    if (!table) { cubic_gradW(r3, hk, out3); return; }
    or_state s; memset(&s, 0, sizeof s);
    s.hk = hk; s.hk2 = hk * hk; s.p.kernel_table = 1;
    build_table(&s);
    kG(&s, r3, out3);
    free(s.tabW); free(s.tabG);
}

/* ------------------------------------------------------------------ */
/* uniform grid (conservative: cell edge 1.001 h_k, so any pair closer  */
/* than h_k lies in adjacent cells whatever the rounding of x / cell)   */
/* ------------------------------------------------------------------ */
typedef struct grid {
    double org[3], inv_cell;
    int dim[3];
    uint32_t ncell;
    uint32_t* start;  /* ncell+1 */
    uint32_t* idx;    /* sorted point indices */
} grid;

static void grid_bounds(const double* pts, uint32_t n, double* lo, double* hi)
{
    for (uint32_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) {
            const double v = pts[3 * i + c];
            if (v < lo[c]) lo[c] = v;
            if (v > hi[c]) hi[c] = v;
        }
}

static inline int grid_coord(const grid* g, double v, int c)
{
    long k = (long)floor((v - g->org[c]) * g->inv_cell);
    if (k < 0) k = 0;
    if (k >= g->dim[c]) k = g->dim[c] - 1;
    return (int)k;
}

static void grid_build(grid* g, const double* lo, const double* hi, double cell,
                       const double* pts, uint32_t n)
{
    g->inv_cell = 1.0 / cell;
    g->ncell = 1;
    for (int c = 0; c < 3; ++c) {
        g->org[c] = lo[c] - cell;
        g->dim[c] = (int)floor((hi[c] - g->org[c]) * g->inv_cell) + 2;
        g->ncell *= (uint32_t)g->dim[c];
    }
    g->start = (uint32_t*)calloc(g->ncell + 1, sizeof(uint32_t));
    g->idx = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));
    uint32_t* key = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));
    for (uint32_t i = 0; i < n; ++i) {
        const int cx = grid_coord(g, pts[3 * i], 0), cy = grid_coord(g, pts[3 * i + 1], 1),
                  cz = grid_coord(g, pts[3 * i + 2], 2);
        key[i] = (uint32_t)cx + (uint32_t)g->dim[0] * ((uint32_t)cz + (uint32_t)g->dim[2] * (uint32_t)cy);
        g->start[key[i] + 1]++;
    }
    for (uint32_t c = 0; c < g->ncell; ++c) g->start[c + 1] += g->start[c];
    uint32_t* fill = (uint32_t*)malloc(g->ncell * sizeof(uint32_t));
    memcpy(fill, g->start, g->ncell * sizeof(uint32_t));
    for (uint32_t i = 0; i < n; ++i) g->idx[fill[key[i]]++] = i;
    free(fill);
    free(key);
}

static void grid_free(grid* g) { free(g->start); free(g->idx); }

static inline int pred64(const double* a, const double* b, double hk2)
{
    const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    return dx * dx + dy * dy + dz * dz < hk2;
}

/* The GPU's FP32 neighbor test, restating the strict < of
 * DFSPH_macro_cuda.h:124,149 with a fixed evaluation order:
 * d2 = fma(dz, dz, fma(dy, dy, dx*dx)). */
static inline int pred32(const float* a, const float* b, float hk2)
{
    const float dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    const float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    return d2 < hk2;
}

static int cmp_u32(const void* a, const void* b)
{
    const uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
    return (x > y) - (x < y);
}

/* ------------------------------------------------------------------ */
/* A8: neighbor search, DFSPH_neighborsSearch_kernel<true>            */
/* (DFSPH_core_cuda.cu:2219-2531): fluid neighbors without self, then  */
/* boundary neighbors; strict < h^2.                                   */
/* ------------------------------------------------------------------ */
uint32_t or_neighbors_fp32(uint32_t n_f, const float* xf, uint32_t n_b, const float* xb,
                           float hk, int brute, uint32_t max,
                           uint32_t* nf, uint32_t* nb, uint32_t* list)
{
    const float hk2 = hk * hk;
    uint32_t maxseen = 0;
    double* pf = (double*)malloc((n_f ? n_f : 1) * 3 * sizeof(double));
    double* pb = (double*)malloc((n_b ? n_b : 1) * 3 * sizeof(double));
    for (uint32_t i = 0; i < 3 * n_f; ++i) pf[i] = xf[i];
    for (uint32_t i = 0; i < 3 * n_b; ++i) pb[i] = xb[i];
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    grid_bounds(pf, n_f, lo, hi);
    grid_bounds(pb, n_b, lo, hi);
    if (n_f + n_b == 0) { free(pf); free(pb); return 0; }
    grid gf, gb;
    const double cell = (double)hk * OR_CELL_MARGIN;
    if (!brute) { grid_build(&gf, lo, hi, cell, pf, n_f); grid_build(&gb, lo, hi, cell, pb, n_b); }

#pragma omp parallel for schedule(dynamic, 256) reduction(max : maxseen)
    for (long il = 0; il < (long)n_f; ++il) {
        const uint32_t i = (uint32_t)il;
        uint32_t* row = list + (size_t)i * max;
        uint32_t cf = 0, cb = 0;
        const float* xi = xf + 3 * (size_t)i;
        if (brute) {
            for (uint32_t j = 0; j < n_f; ++j)
                if (j != i && pred32(xi, xf + 3 * (size_t)j, hk2)) { if (cf < max) row[cf] = j; cf++; }
            for (uint32_t j = 0; j < n_b; ++j)
                if (pred32(xi, xb + 3 * (size_t)j, hk2)) { if (cf + cb < max) row[cf + cb] = j; cb++; }
        } else {
            const int cx = grid_coord(&gf, pf[3 * i], 0), cy = grid_coord(&gf, pf[3 * i + 1], 1),
                      cz = grid_coord(&gf, pf[3 * i + 2], 2);
            for (int pass = 0; pass < 2; ++pass) {
                const grid* g = pass ? &gb : &gf;
                const float* xs = pass ? xb : xf;
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dz = -1; dz <= 1; ++dz)
                        for (int dx = -1; dx <= 1; ++dx) {
                            const int X = cx + dx, Y = cy + dy, Z = cz + dz;
                            if (X < 0 || Y < 0 || Z < 0 || X >= g->dim[0] || Y >= g->dim[1] || Z >= g->dim[2]) continue;
                            const uint32_t c = (uint32_t)X + (uint32_t)g->dim[0] * ((uint32_t)Z + (uint32_t)g->dim[2] * (uint32_t)Y);
                            for (uint32_t q = g->start[c]; q < g->start[c + 1]; ++q) {
                                const uint32_t j = g->idx[q];
                                if (!pass && j == i) continue;
                                if (pred32(xi, xs + 3 * (size_t)j, hk2)) {
                                    if (cf + cb < max) row[cf + cb] = j;
                                    if (pass) cb++; else cf++;
                                }
                            }
                        }
            }
        }
        nf[i] = cf; nb[i] = cb;
        if (cf + cb > maxseen) maxseen = cf + cb;
    }
    if (!brute) { grid_free(&gf); grid_free(&gb); }
    free(pf); free(pb);
    return maxseen;
}

/* ------------------------------------------------------------------ */
/* state                                                               */
/* ------------------------------------------------------------------ */
void or_default_params(or_params* p)
{
    memset(p, 0, sizeof *p);
    p->particle_radius = 0.025;
    p->density0 = 1000.0;
    p->viscosity = 0.01;
    p->gravity[1] = -9.81;
    p->dt = 0.001;
    p->mode = OR_MODE_GPU;
    p->enable_divergence = 1;
    p->max_iter_v = 100; p->max_iter = 100;
    p->min_iter_v = 3;   p->min_iter = 2;     /* DFSPH_core_cuda.cu:828,1378 */
    p->max_error_v = 0.1; p->max_error = 0.01;
    p->cfl_factor = 1.0; p->dt_min = 1e-4; p->dt_max = 5e-3;
    p->max_neighbors = 90;                     /* DFSPH_define_c.h:23 */
}

or_state* or_create(const or_params* p)
{
    or_state* s = (or_state*)calloc(1, sizeof(or_state));
    s->p = *p;
    s->hk = 4.0 * p->particle_radius;           /* DFSPH_c_arrays_structure.cpp:857-869 */
    s->hk2 = s->hk * s->hk;
    s->k_coef = 8.0 / (M_PI * s->hk * s->hk * s->hk);
    s->l_coef = 48.0 / (M_PI * s->hk * s->hk * s->hk);
    if (p->kernel_table) build_table(s);
    {
        double z[3] = {0, 0, 0};
        s->W0 = kW(s, z);
    }
    /* DFSPHCData::reset (DFSPH_c_arrays_structure.cpp:975-1048) */
    s->h = s->h_future = s->h_past = p->dt;
    s->h_ratio_to_past = 1.0;
    s->h_ratio_to_past2 = 1.0;
    s->time = 0.0;
    s->rmode = 0;                               /* no restriction, N-particle averages */
    s->true_count = -1;
    return s;
}

static void free_fluid(or_state* s)
{
    free(s->x); free(s->v); free(s->m); free(s->rho); free(s->alpha); free(s->radv);
    free(s->kappa); free(s->kappaV); free(s->acc); free(s->id); free(s->tag);
    free(s->nf); free(s->nbc); free(s->nbr);
    s->x = s->v = s->m = s->rho = s->alpha = s->radv = s->kappa = s->kappaV = s->acc = NULL;
    s->id = s->nf = s->nbc = s->nbr = s->tag = NULL;
}

void or_destroy(or_state* s)
{
    if (!s) return;
    free_fluid(s);
    free(s->xb); free(s->psi); free(s->tabW); free(s->tabG);
    free(s->body); free(s->x0); free(s->psi0); free(s->vb); free(s->sF); free(s->sbody);
    free(s);
}

int or_set_fluid(or_state* s, uint32_t n, const double* x, const double* v,
                 const double* m, const uint32_t* id)
{
    free_fluid(s);
    s->n = s->cap = n;
    const size_t n1 = n ? n : 1;
    s->x = (double*)malloc(3 * n1 * sizeof(double));
    s->v = (double*)calloc(3 * n1, sizeof(double));
    s->acc = (double*)calloc(3 * n1, sizeof(double));
    s->m = (double*)malloc(n1 * sizeof(double));
    s->rho = (double*)calloc(n1, sizeof(double));
    s->alpha = (double*)calloc(n1, sizeof(double));
    s->radv = (double*)calloc(n1, sizeof(double));
    s->kappa = (double*)calloc(n1, sizeof(double));
    s->kappaV = (double*)calloc(n1, sizeof(double));
    s->id = (uint32_t*)malloc(n1 * sizeof(uint32_t));
    s->tag = (uint32_t*)malloc(n1 * sizeof(uint32_t));
    s->nf = (uint32_t*)calloc(n1, sizeof(uint32_t));
    s->nbc = (uint32_t*)calloc(n1, sizeof(uint32_t));
    s->nbr = (uint32_t*)calloc(n1 * s->p.max_neighbors, sizeof(uint32_t));
    memcpy(s->x, x, 3 * n * sizeof(double));
    if (v) memcpy(s->v, v, 3 * n * sizeof(double));
    /* FluidModel::initMasses: m = 0.8 (2r)^3 rho0 (FluidModel.cpp:94-108) */
    const double d = 2.0 * s->p.particle_radius;
    for (uint32_t i = 0; i < n; ++i) {
        s->m[i] = m ? m[i] : 0.8 * d * d * d * s->p.density0;
        s->id[i] = id ? id[i] : i;
        s->tag[i] = OR_TAG_UNTAGGED;
    }
    return 0;
}

int or_set_restriction(or_state* s, int mode, uint32_t count_active, uint32_t count_active_neighbors,
                       long long true_particle_count, const uint32_t* tags)
{
    if (mode < 0 || mode > 2) return -1;
    s->rmode = mode;
    s->r_active = count_active;
    s->r_active_nbrs = count_active_neighbors;
    s->true_count = true_particle_count;
    if (tags) memcpy(s->tag, tags, s->n * sizeof(uint32_t));
    return 0;
}

/* FluidModel::computeBoundaryPsi (FluidModel.cpp:233-258): psi = rho0 / (W0 + sum_j W) */
int or_set_boundary(or_state* s, uint32_t n, const double* x, const double* psi)
{
    free(s->xb); free(s->psi);
    free(s->body); free(s->x0); free(s->psi0); free(s->vb); free(s->sF); free(s->sbody);
    s->body = NULL; s->x0 = s->psi0 = s->vb = s->sF = NULL; s->sbody = NULL;
    s->nbody = s->ns = 0;
    s->nbw = n;
    s->nbn = n;
    s->xb = (double*)malloc(3 * (n ? n : 1) * sizeof(double));
    s->psi = (double*)malloc((n ? n : 1) * sizeof(double));
    memcpy(s->xb, x, 3 * n * sizeof(double));
    for (int c = 0; c < 3; ++c) { s->bbmin[c] = 1e300; s->bbmax[c] = -1e300; }
    grid_bounds(s->xb, n, s->bbmin, s->bbmax);
    if (psi) { memcpy(s->psi, psi, n * sizeof(double)); return 0; }
    if (n == 0) return 0;
    grid g;
    grid_build(&g, s->bbmin, s->bbmax, s->hk * OR_CELL_MARGIN, s->xb, n);
#pragma omp parallel for schedule(static)
    for (long il = 0; il < (long)n; ++il) {
        const uint32_t i = (uint32_t)il;
        const double* xi = s->xb + 3 * (size_t)i;
        const int cx = grid_coord(&g, xi[0], 0), cy = grid_coord(&g, xi[1], 1), cz = grid_coord(&g, xi[2], 2);
        double delta = s->W0;
        for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int X = cx + dx, Y = cy + dy, Z = cz + dz;
                    if (X < 0 || Y < 0 || Z < 0 || X >= g.dim[0] || Y >= g.dim[1] || Z >= g.dim[2]) continue;
                    const uint32_t c = (uint32_t)X + (uint32_t)g.dim[0] * ((uint32_t)Z + (uint32_t)g.dim[2] * (uint32_t)Y);
                    for (uint32_t q = g.start[c]; q < g.start[c + 1]; ++q) {
                        const uint32_t j = g.idx[q];
                        if (j == i) continue;
                        if (!pred64(xi, s->xb + 3 * (size_t)j, s->hk2)) continue;
                        double d[3] = {xi[0] - s->xb[3 * j], xi[1] - s->xb[3 * j + 1], xi[2] - s->xb[3 * j + 2]};
                        delta += kW(s, d);
                    }
                }
        s->psi[i] = s->p.density0 / delta;
    }
    grid_free(&g);
    return 0;
}

int or_get_boundary_psi(const or_state* s, double* out)
{
    memcpy(out, s->psi, s->nbn * sizeof(double));
    return 0;
}

uint32_t or_num_fluid(const or_state* s) { return s->n; }

static double* field_ptr(const or_state* s, int f, int* width)
{
    *width = 1;
    switch (f) {
    case OR_F_POS: *width = 3; return s->x;
    case OR_F_VEL: *width = 3; return s->v;
    case OR_F_MASS: return s->m;
    case OR_F_DENSITY: return s->rho;
    case OR_F_FACTOR: return s->alpha;
    case OR_F_DENSITY_ADV: return s->radv;
    case OR_F_KAPPA: return s->kappa;
    case OR_F_KAPPAV: return s->kappaV;
    case OR_F_ACC: *width = 3; return s->acc;
    default: return NULL;
    }
}

int or_get(const or_state* s, int field, double* out)
{
    if (field == OR_F_ID) { for (uint32_t i = 0; i < s->n; ++i) out[i] = s->id[i]; return 0; }
    int w; double* p = field_ptr(s, field, &w);
    if (!p) return -1;
    memcpy(out, p, (size_t)w * s->n * sizeof(double));
    return 0;
}

int or_set(or_state* s, int field, const double* in)
{
    int w; double* p = field_ptr(s, field, &w);
    if (!p) return -1;
    memcpy(p, in, (size_t)w * s->n * sizeof(double));
    return 0;
}

int or_get_ids(const or_state* s, uint32_t* out) { memcpy(out, s->id, s->n * sizeof(uint32_t)); return 0; }

void or_get_time(const or_state* s, double* h, double* hf, double* hp, double* t)
{
    *h = s->h; *hf = s->h_future; *hp = s->h_past; *t = s->time;
}

void or_set_time_step(or_state* s, double h)
{
    s->h = s->h_future = s->h_past = h;
    s->h_ratio_to_past = s->h_ratio_to_past2 = 1.0;
}

static void place_bodies(or_state* s);

int or_neighbor_search(or_state* s)
{
    if (s->ns) place_bodies(s);
    const uint32_t n = s->n, max = s->p.max_neighbors;
    s->max_seen = 0; s->overflow = 0;
    if (s->p.fp32_predicate) {
        float* xf = (float*)malloc(3 * (n ? n : 1) * sizeof(float));
        float* xb = (float*)malloc(3 * (s->nbn ? s->nbn : 1) * sizeof(float));
        for (uint32_t i = 0; i < 3 * n; ++i) xf[i] = (float)s->x[i];
        for (uint32_t i = 0; i < 3 * s->nbn; ++i) xb[i] = (float)s->xb[i];
        s->max_seen = or_neighbors_fp32(n, xf, s->nbn, xb, (float)s->hk, 0, max, s->nf, s->nbc, s->nbr);
        free(xf); free(xb);
    } else {
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        grid_bounds(s->x, n, lo, hi);
        grid_bounds(s->xb, s->nbn, lo, hi);
        if (n == 0) return 0;
        grid gf, gb;
        grid_build(&gf, lo, hi, s->hk * OR_CELL_MARGIN, s->x, n);
        grid_build(&gb, lo, hi, s->hk * OR_CELL_MARGIN, s->xb, s->nbn);
        uint32_t maxseen = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(max : maxseen)
        for (long il = 0; il < (long)n; ++il) {
            const uint32_t i = (uint32_t)il;
            uint32_t* row = s->nbr + (size_t)i * max;
            uint32_t cf = 0, cb = 0;
            const double* xi = s->x + 3 * (size_t)i;
            const int cx = grid_coord(&gf, xi[0], 0), cy = grid_coord(&gf, xi[1], 1), cz = grid_coord(&gf, xi[2], 2);
            for (int pass = 0; pass < 2; ++pass) {
                const grid* g = pass ? &gb : &gf;
                const double* xs = pass ? s->xb : s->x;
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dz = -1; dz <= 1; ++dz)
                        for (int dx = -1; dx <= 1; ++dx) {
                            const int X = cx + dx, Y = cy + dy, Z = cz + dz;
                            if (X < 0 || Y < 0 || Z < 0 || X >= g->dim[0] || Y >= g->dim[1] || Z >= g->dim[2]) continue;
                            const uint32_t c = (uint32_t)X + (uint32_t)g->dim[0] * ((uint32_t)Z + (uint32_t)g->dim[2] * (uint32_t)Y);
                            for (uint32_t q = g->start[c]; q < g->start[c + 1]; ++q) {
                                const uint32_t j = g->idx[q];
                                if (!pass && j == i) continue;
                                if (pred64(xi, xs + 3 * (size_t)j, s->hk2)) {
                                    if (cf + cb < max) row[cf + cb] = j;
                                    if (pass) cb++; else cf++;
                                }
                            }
                        }
            }
            s->nf[i] = cf; s->nbc[i] = cb;
            if (cf + cb > maxseen) maxseen = cf + cb;
        }
        s->max_seen = maxseen;
        grid_free(&gf); grid_free(&gb);
    }
    /* restricted-out particles get no lists (DFSPH_core_cuda.cu:2545-2552) */
    for (uint32_t i = 0; i < n; ++i)
        if (!r_source(s, i)) { s->nf[i] = 0; s->nbc[i] = 0; }
    /* sort each particle's fluid and boundary sub-lists so sums have a fixed order */
    for (uint32_t i = 0; i < n; ++i) {
        if (s->nf[i] + s->nbc[i] > max) {
            s->overflow++;
            /* truncate like a capped buffer: keep the first max entries */
            uint32_t tot = max, f = s->nf[i] < max ? s->nf[i] : max;
            s->nf[i] = f; s->nbc[i] = tot - f;
        }
        uint32_t* row = s->nbr + (size_t)i * max;
        qsort(row, s->nf[i], sizeof(uint32_t), cmp_u32);
        qsort(row + s->nf[i], s->nbc[i], sizeof(uint32_t), cmp_u32);
    }
    return s->overflow ? 1 : 0;
}

int or_get_neighbors(const or_state* s, uint32_t* nf, uint32_t* nb, uint32_t* list)
{
    memcpy(nf, s->nf, s->n * sizeof(uint32_t));
    memcpy(nb, s->nbc, s->n * sizeof(uint32_t));
    memcpy(list, s->nbr, (size_t)s->n * s->p.max_neighbors * sizeof(uint32_t));
    return 0;
}

/* helpers over one particle's lists */
#define ROW(i) (s->nbr + (size_t)(i) * s->p.max_neighbors)
#define XI(i) (s->x + 3 * (size_t)(i))
#define VI(i) (s->v + 3 * (size_t)(i))
#define XB(j) (s->xb + 3 * (size_t)(j))

static inline void sub3(const double* a, const double* b, double* o) { o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2]; }
static inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* Sum m_j (v_i - v_j) . gradW_ij over fluid and boundary (walls: v_j = 0;
 * rigid-body particles move).  computeDensityChange / computeDensityAdv
 * (DFSPH_core_cuda.cu:546-614, 1165-1215). */
static double density_change_sum(const or_state* s, uint32_t i)
{
    const uint32_t* row = ROW(i);
    const double* xi = XI(i);
    const double* vi = VI(i);
    double acc = 0.0, d[3], g[3], dv[3];
    for (uint32_t k = 0; k < s->nf[i]; ++k) {
        const uint32_t j = row[k];
        sub3(xi, XI(j), d); kG(s, d, g); sub3(vi, VI(j), dv);
        acc += s->m[j] * dot3(dv, g);
    }
    for (uint32_t k = 0; k < s->nbc[i]; ++k) {
        const uint32_t j = row[s->nf[i] + k];
        sub3(xi, XB(j), d); kG(s, d, g);
        if (s->ns) sub3(vi, s->vb + 3 * (size_t)j, dv);
        else { dv[0] = vi[0]; dv[1] = vi[1]; dv[2] = vi[2]; }
        acc += s->psi[j] * dot3(dv, g);     /* body particles: (v_i - v_j), DFSPH_core_cuda.cu:599-603 */
    }
    return acc;
}

/* divergence source with the particle-deficiency rule (n < 20 -> 0) */
static void compute_density_change(or_state* s, uint32_t i)
{
    const uint32_t nn = s->nf[i] + s->nbc[i];
    if (nn < 20) { s->radv[i] = 0.0; return; }
    const double dr = density_change_sum(s, i);
    s->radv[i] = dr > 0.0 ? dr : 0.0;
}

/* A9 fused density + factor + divergence source
 * (DFSPH_divergence_warmstart_init_kernel, DFSPH_core_cuda.cu:220-376;
 *  TimeStep::computeDensities TimeStep.cpp:64-105 and
 *  TimeStepDFSPH::computeDFSPHFactor TimeStepDFSPH.cpp:162-226). */
static void density_and_factor(or_state* s, uint32_t i)
{
    const uint32_t* row = ROW(i);
    const double* xi = XI(i);
    double rho = s->m[i] * s->W0, sum_g2 = 0.0, gi[3] = {0, 0, 0}, d[3], g[3];
    for (uint32_t k = 0; k < s->nf[i]; ++k) {
        const uint32_t j = row[k];
        sub3(xi, XI(j), d);
        rho += s->m[j] * kW(s, d);
        kG(s, d, g);
        const double gp[3] = {s->m[j] * g[0], s->m[j] * g[1], s->m[j] * g[2]};
        sum_g2 += dot3(gp, gp);
        gi[0] += gp[0]; gi[1] += gp[1]; gi[2] += gp[2];
    }
    for (uint32_t k = 0; k < s->nbc[i]; ++k) {
        const uint32_t j = row[s->nf[i] + k];
        sub3(xi, XB(j), d);
        rho += s->psi[j] * kW(s, d);
        kG(s, d, g);
        const double gp[3] = {s->psi[j] * g[0], s->psi[j] * g[1], s->psi[j] * g[2]};
        sum_g2 += dot3(gp, gp);
        gi[0] += gp[0]; gi[1] += gp[1]; gi[2] += gp[2];
    }
    sum_g2 += dot3(gi, gi);
    s->rho[i] = rho;
    s->alpha[i] = -(1.0 / s->h) / (sum_g2 > OR_EPS ? sum_g2 : OR_EPS);
}

/* Jacobi velocity update shared by both solvers
 * (divergenceSolveParticle DFSPH_core_cuda.cu:400-481,
 *  pressureSolveParticle :867-951).  k holds k_j for every fluid particle. */
static void jacobi_update(or_state* s, uint32_t i, const double* k, double step, double* dv_out)
{
    const uint32_t* row = ROW(i);
    const double* xi = XI(i);
    const double ki = k[i];
    double dv[3] = {0, 0, 0}, d[3], g[3];
    for (uint32_t q = 0; q < s->nf[i]; ++q) {
        const uint32_t j = row[q];
        const double ks = ki + k[j];
        if (fabs(ks) > OR_EPS) {
            sub3(xi, XI(j), d); kG(s, d, g);
            const double c = ks * s->m[j];
            dv[0] += c * g[0]; dv[1] += c * g[1]; dv[2] += c * g[2];
        }
    }
    if (fabs(ki) > OR_EPS) {
        for (uint32_t q = 0; q < s->nbc[i]; ++q) {
            const uint32_t j = row[s->nf[i] + q];
            sub3(xi, XB(j), d); kG(s, d, g);
            const double c = ki * s->psi[j];
            dv[0] += c * g[0]; dv[1] += c * g[1]; dv[2] += c * g[2];
        }
    }
    dv_out[0] = step * dv[0]; dv_out[1] = step * dv[1]; dv_out[2] = step * dv[2];
}

/* Reactions of one Jacobi update on the body particles: -m_i k_i psi_j gradW_ij
 * for every particle that was updated (gate: also a positive source, the
 * divergence warm start) and |k_i| > eps (DFSPH_core_cuda.cu:467-477, 936-946).
 * Serial, in particle order. */
static void body_reactions(or_state* s, const double* k, int gate_radv)
{
    if (!s->ns) return;
    double d[3], g[3];
    for (uint32_t i = 0; i < s->n; ++i) {
        if (!r_compute(s, i) || (gate_radv && !(s->radv[i] > 0.0))) continue;
        const double ki = k[i];
        if (!(fabs(ki) > OR_EPS)) continue;
        const uint32_t* row = ROW(i);
        for (uint32_t q = 0; q < s->nbc[i]; ++q) {
            const uint32_t j = row[s->nf[i] + q];
            if (j < s->nbw) continue;
            sub3(XI(i), XB(j), d); kG(s, d, g);
            const double c = -s->m[i] * ki * s->psi[j];
            double* f = s->sF + 3 * (size_t)(j - s->nbw);
            f[0] += c * g[0]; f[1] += c * g[1]; f[2] += c * g[2];
        }
    }
}

static double sum_radv(const or_state* s)
{
    double acc = 0.0;
    for (uint32_t i = 0; i < s->n; ++i) acc += s->radv[i];
    return acc;
}

/* A9-A12: cuda_divergenceSolve (DFSPH_core_cuda.cu:785-861) in mode GPU;
 * TimeStepDFSPH::divergenceSolve (TimeStepDFSPH.cpp:463-667) in mode CPU. */
uint32_t or_divergence_solve(or_state* s, double* err_out)
{
    const uint32_t n = s->n;
    const double h = s->h;
    double* k = (double*)malloc((n ? n : 1) * sizeof(double));
    double* dv = (double*)malloc(3 * (n ? n : 1) * sizeof(double));
    const int gpu = s->p.mode == OR_MODE_GPU;

    /* warm start init: kappaV rescale + fused density / factor / source */
#pragma omp parallel for schedule(static)
    for (long il = 0; il < (long)n; ++il) {
        const uint32_t i = (uint32_t)il;
        if (!r_source(s, i)) continue;
        if (gpu) {
            const double kv = s->kappaV[i] * s->h_ratio_to_past / 2.0;   /* core.cu:240 */
            s->kappaV[i] = kv > -0.5 ? kv : -0.5;
        } else {
            const double kv = s->kappaV[i] / h;                          /* TimeStepDFSPH.cpp:488 */
            s->kappaV[i] = 0.5 * (kv > -0.5 ? kv : -0.5);
        }
        density_and_factor(s, i);
        compute_density_change(s, i);
    }
    /* warm start velocity update, gated on a positive source (core.cu:507-511) */
#pragma omp parallel for schedule(static)
    for (long il = 0; il < (long)n; ++il) {
        const uint32_t i = (uint32_t)il;
        if (s->radv[i] > 0.0 && r_compute(s, i)) jacobi_update(s, i, s->kappaV, h, dv + 3 * i);
        else dv[3 * i] = dv[3 * i + 1] = dv[3 * i + 2] = 0.0;
    }
    body_reactions(s, s->kappaV, 1);
    for (uint32_t i = 0; i < 3 * n; ++i) s->v[i] += dv[i];
    /* divergence init: kappaV = 0, source from the warm-started velocities */
#pragma omp parallel for schedule(static)
    for (long il = 0; il < (long)n; ++il) {
        if (!r_source(s, (uint32_t)il)) continue;
        s->kappaV[il] = 0.0;
        compute_density_change(s, (uint32_t)il);
    }
    const double eta = s->p.max_error_v * 0.01 * s->p.density0 / h;
    double err = 0.0;
    uint32_t it = 0;
    while ((err > eta || it < s->p.min_iter_v) && it < s->p.max_iter_v) {
        for (uint32_t i = 0; i < n; ++i) {
            k[i] = s->radv[i] * s->alpha[i];
            if (r_compute(s, i)) s->kappaV[i] += k[i];
        }
#pragma omp parallel for schedule(static)
        for (long il = 0; il < (long)n; ++il) {
            if (r_compute(s, (uint32_t)il)) jacobi_update(s, (uint32_t)il, k, h, dv + 3 * il);
            else dv[3 * il] = dv[3 * il + 1] = dv[3 * il + 2] = 0.0;
        }
        body_reactions(s, k, 0);
        for (uint32_t i = 0; i < 3 * n; ++i) s->v[i] += dv[i];
#pragma omp parallel for schedule(static)
        for (long il = 0; il < (long)n; ++il)
            if (r_source(s, (uint32_t)il)) compute_density_change(s, (uint32_t)il);
        err = n ? sum_radv(s) / r_count(s) : 0.0;
        it++;
    }
    if (!gpu) for (uint32_t i = 0; i < n; ++i) s->kappaV[i] *= h;        /* TimeStepDFSPH.cpp:658-661 */
    free(k); free(dv);
    if (err_out) *err_out = err;
    return it;
}

/* A13: DFSPH_viscosityXSPH_kernel (DFSPH_core_cuda.cu:1423-1510),
 * Viscosity_XSPH::step (Viscosity_XSPH.cpp:15-62): a = g + XSPH. */
void or_external_forces(or_state* s)
{
    const double invH = 1.0 / s->h, nu = s->p.viscosity;
#pragma omp parallel for schedule(static)
    for (long il = 0; il < (long)s->n; ++il) {
        const uint32_t i = (uint32_t)il;
        if (!r_compute(s, i)) continue;          /* DFSPH_core_cuda.cu:1428-1434 */
        const uint32_t* row = ROW(i);
        const double* xi = XI(i);
        const double* vi = VI(i);
        double a[3] = {0, 0, 0}, d[3];
        for (uint32_t q = 0; q < s->nf[i]; ++q) {
            const uint32_t j = row[q];
            sub3(xi, XI(j), d);
            const double w = kW(s, d);
            const double c = invH * nu * (s->m[j] / s->rho[j]) * w;
            a[0] -= c * (vi[0] - s->v[3 * j]);
            a[1] -= c * (vi[1] - s->v[3 * j + 1]);
            a[2] -= c * (vi[2] - s->v[3 * j + 2]);
        }
        double* ai = s->acc + 3 * (size_t)i;
        ai[0] = s->p.gravity[0] + a[0];
        ai[1] = s->p.gravity[1] + a[1];
        ai[2] = s->p.gravity[2] + a[2];
    }
}

/* A17: cuda_CFL (DFSPH_core_cuda.cu:3321-3399), TimeStep::updateTimeStepSizeCFL
 * (TimeStep.cpp:124-165): h = cfl 0.4 (2r) / sqrt(max |v + a h|^2). */
double or_cfl(or_state* s)
{
    double mx = s->p.cfl_seed;
    for (uint32_t i = 0; i < s->n; ++i) {
        double w[3];
        for (int c = 0; c < 3; ++c) w[c] = s->v[3 * i + c] + s->acc[3 * i + c] * s->h;
        const double m2 = dot3(w, w);
        if (m2 > mx) mx = m2;
    }
    double h = s->p.cfl_factor * 0.4 * (2.0 * s->p.particle_radius / sqrt(mx));
    if (!(h < s->p.dt_max)) h = s->p.dt_max;
    if (h < s->p.dt_min) h = s->p.dt_min;
    return h;
}

/* DFSPHCData::updateTimeStep (DFSPH_c_arrays_structure.cpp:1100-1106) */
void or_update_time_step(or_state* s, double hf)
{
    s->h_future = hf;
    s->h_ratio_to_past = s->h / hf;
    s->h_ratio_to_past2 = (s->h * s->h) / (hf * hf);
}

/* A14: DFSPH_update_vel_kernel (DFSPH_core_cuda.cu:3047-3070) */
void or_update_vel(or_state* s)
{
    for (uint32_t i = 0; i < 3 * s->n; ++i)
        if (r_compute(s, i / 3)) s->v[i] += s->h * s->acc[i];
    if (s->p.mode == OR_MODE_GPU) {
        for (uint32_t i = 0; i < s->n; ++i) {
            if (!r_compute(s, i)) continue;
            const double kk = s->kappa[i] * s->h_ratio_to_past2;
            s->kappa[i] = kk > -0.5 ? kk : -0.5;
        }
    }
}

/* pressure source: rho_adv = max(rho + h_f sum - rho0, 0) (computeDensityAdv,
 * DFSPH_core_cuda.cu:1165-1215; TimeStepDFSPH.cpp:669-697) */
static void compute_density_adv(or_state* s, uint32_t i, double hf)
{
    const double r = s->rho[i] + hf * density_change_sum(s, i) - s->p.density0;
    s->radv[i] = r > 0.0 ? r : 0.0;
}

/* A15: cuda_pressureSolve (DFSPH_core_cuda.cu:1338-1419) in mode GPU;
 * TimeStepDFSPH::pressureSolve (TimeStepDFSPH.cpp:228-461) in mode CPU. */
uint32_t or_pressure_solve(or_state* s, double* err_out)
{
    const uint32_t n = s->n;
    const int gpu = s->p.mode == OR_MODE_GPU;
    const double hf = gpu ? s->h_future : s->h;
    double* k = (double*)malloc((n ? n : 1) * sizeof(double));
    double* dv = (double*)malloc(3 * (n ? n : 1) * sizeof(double));
    if (!gpu) {
        /* kappa = max(kappa_stored / h^2, -0.5) (TimeStepDFSPH.cpp:286) */
        for (uint32_t i = 0; i < n; ++i) {
            const double kk = s->kappa[i] / (hf * hf);
            s->kappa[i] = kk > -0.5 ? kk : -0.5;
        }
    }
    /* warm start: cuda_pressure_compute<true> (core.cu:965-1000), no gate */
#pragma omp parallel for schedule(static)
    for (long il = 0; il < (long)n; ++il) {
        if (r_compute(s, (uint32_t)il)) jacobi_update(s, (uint32_t)il, s->kappa, hf, dv + 3 * il);
        else dv[3 * il] = dv[3 * il + 1] = dv[3 * il + 2] = 0.0;
    }
    body_reactions(s, s->kappa, 0);
    for (uint32_t i = 0; i < 3 * n; ++i) s->v[i] += dv[i];
    /* init: kappa = 0, factor *= 1/h_f, rho_adv (core.cu:1217-1262) */
#pragma omp parallel for schedule(static)
    for (long il = 0; il < (long)n; ++il) {
        const uint32_t i = (uint32_t)il;
        if (!r_source(s, i)) continue;
        s->kappa[i] = 0.0;
        s->alpha[i] *= 1.0 / hf;
        compute_density_adv(s, i, hf);
        if (gpu) s->kappa[i] += s->radv[i] * s->alpha[i];
    }
    const double eta = s->p.max_error * 0.01 * s->p.density0;
    double err = 0.0;
    uint32_t it = 0;
    while ((err > eta || it < s->p.min_iter) && it < s->p.max_iter) {
        for (uint32_t i = 0; i < n; ++i) {
            k[i] = s->radv[i] * s->alpha[i];
            if (!gpu) s->kappa[i] += k[i];                   /* TimeStepDFSPH.cpp:378 */
        }
#pragma omp parallel for schedule(static)
        for (long il = 0; il < (long)n; ++il) {
            if (r_compute(s, (uint32_t)il)) jacobi_update(s, (uint32_t)il, k, hf, dv + 3 * il);
            else dv[3 * il] = dv[3 * il + 1] = dv[3 * il + 2] = 0.0;
        }
        body_reactions(s, k, 0);
        for (uint32_t i = 0; i < 3 * n; ++i) s->v[i] += dv[i];
#pragma omp parallel for schedule(static)
        for (long il = 0; il < (long)n; ++il) {
            const uint32_t i = (uint32_t)il;
            if (!r_source(s, i)) continue;
            compute_density_adv(s, i, hf);
            if (gpu) s->kappa[i] += s->radv[i] * s->alpha[i];   /* core.cu:1209-1212 */
        }
        err = n ? sum_radv(s) / r_count(s) : 0.0;
        it++;
    }
    if (!gpu) for (uint32_t i = 0; i < n; ++i) s->kappa[i] *= hf * hf;   /* TimeStepDFSPH.cpp:455-459 */
    free(k); free(dv);
    if (err_out) *err_out = err;
    return it;
}

/* A16: DFSPH_update_pos_kernel (DFSPH_core_cuda.cu:3127-3214) */
static void integrate_bodies(or_state* s);

void or_update_pos(or_state* s)
{
    for (uint32_t i = 0; i < 3 * s->n; ++i)
        if (r_compute(s, i / 3)) s->x[i] += s->h * s->v[i];
    integrate_bodies(s);
}

/* DFSPHCData::onSimulationStepEnd (DFSPH_c_arrays_structure.cpp:1108-1118)
 * followed by time += h (DFSPH_CUDA.cpp:441-452) */
void or_step_end(or_state* s)
{
    s->h_past = s->h;
    s->h = s->h_future;
    s->time += s->h;
}

/* A18: check_particles_positions_cuda (SPH_other_systems_cuda.cu:1244-1330):
 * a particle is outside when any coordinate leaves the boundary bounding
 * box; with remove, the others keep their order and the count shrinks. */
uint32_t or_check_positions(or_state* s, int remove)
{
    uint32_t w = 0, cnt = 0;
    for (uint32_t i = 0; i < s->n; ++i) {
        const double* xi = XI(i);
        const int out = xi[0] > s->bbmax[0] || xi[1] > s->bbmax[1] || xi[2] > s->bbmax[2] ||
                        xi[0] < s->bbmin[0] || xi[1] < s->bbmin[1] || xi[2] < s->bbmin[2];
        if (out) { cnt++; if (remove) continue; }
        if (remove && w != i) {
            memcpy(s->x + 3 * w, s->x + 3 * i, 3 * sizeof(double));
            memcpy(s->v + 3 * w, s->v + 3 * i, 3 * sizeof(double));
            memcpy(s->acc + 3 * w, s->acc + 3 * i, 3 * sizeof(double));
            s->m[w] = s->m[i]; s->rho[w] = s->rho[i]; s->alpha[w] = s->alpha[i];
            s->radv[w] = s->radv[i]; s->kappa[w] = s->kappa[i]; s->kappaV[w] = s->kappaV[i];
            s->id[w] = s->id[i];
            s->tag[w] = s->tag[i];
        }
        w++;
    }
    if (remove) s->n = w;
    return cnt;
}

/* DFSPHCUDA::step (DFSPH_CUDA.cpp:82-794) / TimeStepDFSPH::step (:29-160) */
int or_step(or_state* s, or_stats* st)
{
    or_stats loc; memset(&loc, 0, sizeof loc);
    const int ns = or_neighbor_search(s);
    loc.max_nbr = s->max_seen;
    loc.overflow = s->overflow;
    loc.dt_used = s->h;
    if (s->p.enable_divergence) loc.iters_v = or_divergence_solve(s, &loc.err_v);
    else {
        /* the factor and density are still needed by the pressure solve */
        for (uint32_t i = 0; i < s->n; ++i) density_and_factor(s, i);
    }
    or_external_forces(s);
    double hf = s->h;
    if (s->p.use_cfl) hf = or_cfl(s);
    or_update_time_step(s, hf);
    or_update_vel(s);
    loc.iters = or_pressure_solve(s, &loc.err);
    or_update_pos(s);
    or_step_end(s);
    loc.removed = or_check_positions(s, s->p.remove_escaped);
    loc.n_fluid = s->n;
    if (st) *st = loc;
    return ns;
}

/* ------------------------------------------------------------------ */
/* A19: open boundary (OpenBoundariesSimple.cu) and BufferFluidSurface  */
/* ------------------------------------------------------------------ */

/* BufferFluidSurface::isinside (BufferFluidSurface.cu:775-880).  margin: the
 * distance of the deciding quantity from its threshold (points with a tiny
 * margin can flip between precisions). */
/* MeshPerso::isInside (BufferFluidSurface.cu:224-270): parity of the
 * crossings of the ray q + t u, t > 0, with the triangles, each found by the
 * Moeller-Trumbore test (Moeller and Trumbore 1997: Cramer's rule on
 * [-u, e1, e2] for the barycentric coordinates a, b and t).  The ray
 * direction is the library's fixed one.  *margin: the smallest of |t| over
 * the crossings and of the distances, within a triangle's plane, from the
 * ray's hit point to the triangle's edges (small values mean FP32 and FP64
 * may count differently). */
static int mesh_inside(const or_surface* S, const double* q, double* margin)
{
    static const double u[3] = {0.31190428, 0.75361035, 0.57860795};
    uint32_t crossings = 0;
    double mg = 1e300;
    for (uint32_t t = 0; t < S->ntri; ++t) {
        const double* v0 = S->tri + 9 * (size_t)t;
        const double* v1 = v0 + 3;
        const double* v2 = v0 + 6;
        double e1[3], e2[3], s[3], pv[3], w[3];
        sub3(v1, v0, e1); sub3(v2, v0, e2); sub3(q, v0, s);
        pv[0] = u[1] * e2[2] - u[2] * e2[1]; pv[1] = u[2] * e2[0] - u[0] * e2[2]; pv[2] = u[0] * e2[1] - u[1] * e2[0];
        const double det = dot3(e1, pv);
        if (fabs(det) < 1e-12) continue;
        const double inv = 1.0 / det;
        const double a = inv * dot3(s, pv);
        w[0] = s[1] * e1[2] - s[2] * e1[1]; w[1] = s[2] * e1[0] - s[0] * e1[2]; w[2] = s[0] * e1[1] - s[1] * e1[0];
        const double b = inv * dot3(u, w);
        const double tt = inv * dot3(e2, w);
        const double c = 1.0 - a - b;
        const double size = sqrt(dot3(e1, e1));
        const double edge = fmin(fabs(a), fmin(fabs(b), fabs(c))) * size;
        const int in_tri = a >= 0.0 && b >= 0.0 && c >= 0.0;
        if (tt > -1e-3 * size) mg = fmin(mg, edge);
        if (in_tri) mg = fmin(mg, fabs(tt));
        if (in_tri && tt > 0.0) crossings++;
    }
    if (margin) *margin = mg;
    return (int)(crossings & 1u);
}

int or_surface_inside(const or_surface* S, const double* p, double* margin)
{
    int r = 0;
    double mg = 1e300;
    const double vx = p[0] - S->center[0], vy = p[1] - S->center[1], vz = p[2] - S->center[2];
    switch (S->type) {
    case 3: {
        const double q[3] = {vx, vy, vz};
        r = mesh_inside(S, q, &mg);
        break;
    }
    case 0: {
        const double d = (p[0] - S->o[0]) * S->n[0] + (p[1] - S->o[1]) * S->n[1] + (p[2] - S->o[2]) * S->n[2];
        r = !(d < 0.0); mg = fabs(d); break;
    }
    case 1: {
        const double a[3] = {fabs(vx), fabs(vy), fabs(vz)};
        r = a[0] < S->half[0] && a[1] < S->half[1] && a[2] < S->half[2];
        for (int c = 0; c < 3; ++c) mg = fmin(mg, fabs(S->half[c] - a[c]));
        break;
    }
    case 2: {
        const double l = sqrt(vx * vx + vz * vz);
        mg = fabs(fabs(vy) - S->half[1]);
        if (fabs(vy) > S->half[1]) { r = 0; break; }
        r = l < S->radius; mg = fmin(mg, fabs(l - S->radius)); break;
    }
    case 4: {
        const double l = sqrt(vx * vx + vy * vy + vz * vz);
        r = l < S->radius; mg = fabs(l - S->radius); break;
    }
    case 5: {
        mg = fabs(fabs(vy) - S->half[1]);
        if (fabs(vy) > S->half[1]) { r = 0; break; }
        const double l = sqrt(vx * vx + vz * vz);
        mg = fmin(mg, fmin(fabs(l - S->radius_internal), fabs(l - S->radius)));
        if (l < S->radius_internal) { r = 1; break; }
        if (l > S->radius) { r = 0; break; }
        /* angle to the first point folded into one period, then the border
         * segment between the outer vertex (angle 0) and the inner one
         * (half a period) or the reverse */
        const double alpha = acos((vx * S->dir[0] + vz * S->dir[2]) / l);
        const double a0 = 2.0 * M_PI / (double)S->points;
        double k = alpha / a0;
        while (k > 1.0) k -= 1.0;
        const double x = k * a0;
        double a1 = 0.0, r1 = S->radius, a2 = a0 / 2.0, r2 = S->radius_internal;
        if (x > a0 / 2.0) { r1 = S->radius_internal; r2 = S->radius; a1 = a2; a2 = a1 + a0 / 2.0; }
        const double x1 = r1 * cos(a1), y1 = r1 * sin(a1), x2 = r2 * cos(a2), y2 = r2 * sin(a2);
        const double m = (y2 - y1) / (x2 - x1);
        const double d = (-x1 * m + y1) / (-m * cos(x) + sin(x));
        r = l <= d; mg = fmin(mg, fabs(l - d));
        break;
    }
    default: r = 0;
    }
    if (margin) *margin = mg;
    return S->reversed ? !r : r;
}

/* BufferFluidSurface::getNormal (BufferFluidSurface.cu:730-772) */
static void surface_normal(const or_surface* S, const double* p, double* nrm)
{
    nrm[0] = nrm[1] = nrm[2] = 0.0;
    if (S->type == 2 || S->type == 4 || S->type == 5)
        for (int c = 0; c < 3; ++c) nrm[c] = S->center[c] - p[c];
    if (S->reversed) for (int c = 0; c < 3; ++c) nrm[c] = -nrm[c];
    const double l = sqrt(dot3(nrm, nrm));
    if (l != 0.0) for (int c = 0; c < 3; ++c) nrm[c] /= l;
}

/* inflow_with_predefined_positions_kernel (OpenBoundariesSimple.cu:360-508),
 * by brute force over the fluid and boundary (small cases).  For each
 * candidate: nearest fluid distance within h (100 when none); when above
 * min_dist, the (m W)-weighted fluid velocity (boundary: weight only) and
 * the density W0 m + sum m_j W + sum psi_j W decide.  Returns the number
 * accepted; vel (3 nc) is set for every candidate with a density estimate. */
uint32_t or_inflow(const or_state* s, const or_surface* boundary, uint32_t nc, const double* cand, double min_dist,
                   double max_density, double mass, uint8_t* accept, double* vel, double* dmin_out, double* dens_out)
{
    uint32_t cnt = 0;
    for (uint32_t c = 0; c < nc; ++c) {
        const double* p = cand + 3 * (size_t)c;
        double dmin = 100.0, d[3];
        for (uint32_t j = 0; j < s->n; ++j) {
            sub3(p, XI(j), d);
            const double r2 = dot3(d, d);
            if (r2 < s->hk2) dmin = fmin(dmin, sqrt(r2));
        }
        double v[3] = {0, 0, 0}, dens = s->W0 * mass, sw = 0.0;
        int ok = 0;
        if (dmin > min_dist) {
            for (uint32_t j = 0; j < s->n; ++j) {
                sub3(p, XI(j), d);
                if (dot3(d, d) < s->hk2) {
                    const double w = kW(s, d) * s->m[j];
                    dens += w; sw += w;
                    for (int q = 0; q < 3; ++q) v[q] += s->v[3 * (size_t)j + q] * w;
                }
            }
            for (uint32_t j = 0; j < s->nbn; ++j) {
                sub3(p, XB(j), d);
                if (dot3(d, d) < s->hk2) {
                    const double w = kW(s, d) * s->psi[j];
                    sw += w; dens += w;
                }
            }
            if (sw > 0.0) for (int q = 0; q < 3; ++q) v[q] /= sw;
            double nrm[3];
            surface_normal(boundary, p, nrm);
            ok = !(dens > max_density || dot3(v, nrm) < 0.0);
        }
        if (accept) accept[c] = (uint8_t)ok;
        if (vel) for (int q = 0; q < 3; ++q) vel[3 * (size_t)c + q] = v[q];
        if (dmin_out) dmin_out[c] = dmin;
        if (dens_out) dens_out[c] = dens;
        cnt += (uint32_t)ok;
    }
    return cnt;
}

/* outflow_basic_kernel (OpenBoundariesSimple.cu:517-534): keep[i] = 0 when
 * particle i is outside both the interior and the fluid-surface shape */
uint32_t or_outflow(const or_state* s, const or_surface* interior, const or_surface* surface, uint8_t* keep)
{
    uint32_t removed = 0;
    for (uint32_t i = 0; i < s->n; ++i) {
        const int out = !or_surface_inside(interior, XI(i), NULL) && !or_surface_inside(surface, XI(i), NULL);
        keep[i] = (uint8_t)!out;
        removed += (uint32_t)out;
    }
    return removed;
}

/* One density-extraction pass of the rest-fluid loader
 * (RestFLuidLoader::tagDataToSurface, RestFluidLoader.cu:2700-3130, its
 * selection particle_selection_rule_1_kernel, :773-890): densities on the
 * current lists (boundary included).  air (n flags, may be NULL) marks the
 * air particles, which only add density and stay.  For a candidate (any
 * other), keep[i] = 0 when the boundary alone gives it more than
 * boundary_limit, or when rho_i > limit and no listed candidate above limit
 * is denser (of equal densities the lower index goes).  Returns the count
 * removed; *avg: the mean density of the kept candidates; rho_out (n, may be
 * NULL): the densities.  The state keeps its particles. */
uint32_t or_rest_extract(or_state* s, double limit, double boundary_limit, const uint8_t* air, uint8_t* keep,
                         double* avg, double* rho_out)
{
    const uint32_t n = s->n;
    for (uint32_t i = 0; i < n; ++i) density_and_factor(s, i);
    uint32_t removed = 0, kept = 0;
    double sum = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        if (rho_out) rho_out[i] = s->rho[i];
        if (air && air[i]) { keep[i] = 1; continue; }
        const uint32_t* row = ROW(i);
        const double* xi = XI(i);
        double rb = 0.0, d[3];
        for (uint32_t k = 0; k < s->nbc[i]; ++k) {
            const uint32_t j = row[s->nf[i] + k];
            sub3(xi, XB(j), d);
            rb += s->psi[j] * kW(s, d);
        }
        const double ri = s->rho[i];
        int leave = rb > boundary_limit;
        if (!leave && ri > limit) {
            int densest = 1;
            for (uint32_t k = 0; k < s->nf[i]; ++k) {
                const uint32_t j = row[k];
                if (air && air[j]) continue;
                const double rj = s->rho[j];
                if (rj > limit && (rj > ri || (rj == ri && j < i))) densest = 0;
            }
            leave = densest;
        }
        keep[i] = (uint8_t)!leave;
        removed += (uint32_t)leave;
        if (!leave) { sum += ri; kept++; }
    }
    if (avg) *avg = kept ? sum / (double)kept : 0.0;
    return removed;
}

/* ------------------------------------------------------------------ */
/* Features and training targets of the learned warm start (no reference */
/* counterpart; inserted where the reference forms its warm starts,     */
/* DFSPH_core_cuda.cu:240 and :3063-3068).  Mode GPU state: kappa and    */
/* kappaV hold the last solves' increments (raw units).                  */
/* ------------------------------------------------------------------ */
void or_features(const or_state* s, double* out)
{
    const double hk = s->hk, rho0 = s->p.density0, diameter = 2.0 * s->p.particle_radius;
#pragma omp parallel for schedule(static)
    for (long il = 0; il < (long)s->n; ++il) {
        const uint32_t i = (uint32_t)il;
        const uint32_t* row = ROW(i);
        const double* xi = XI(i);
        const double* vi = VI(i);
        double dfl = 0.0, dbd = 0.0, off[3] = {0, 0, 0}, dist = 0.0, d[3];
        for (uint32_t k = 0; k < s->nf[i]; ++k) {
            const uint32_t j = row[k];
            sub3(xi, XI(j), d);
            dfl += s->m[j] * kW(s, d);
            off[0] -= d[0]; off[1] -= d[1]; off[2] -= d[2];
            dist += sqrt(dot3(d, d));
        }
        for (uint32_t k = 0; k < s->nbc[i]; ++k) {
            const uint32_t j = row[s->nf[i] + k];
            sub3(xi, XB(j), d);
            dbd += s->psi[j] * kW(s, d);
            off[0] -= d[0]; off[1] -= d[1]; off[2] -= d[2];
            dist += sqrt(dot3(d, d));
        }
        const double rho = s->m[i] * s->W0 + dfl + dbd;
        const double nn = (double)(s->nf[i] + s->nbc[i]);
        const double inv_n = nn > 0.0 ? 1.0 / nn : 0.0;
        double* f = out + (size_t)i * 12;
        f[0] = rho / rho0 - 1.0;                                 /* density ratio                     */
        f[1] = density_change_sum(s, i) * s->h / rho0;           /* divergence source, unclamped      */
        f[2] = nn / 40.0;                                        /* neighbor count                    */
        f[3] = sqrt(dot3(vi, vi)) * s->h / diameter;             /* speed in diameters per step       */
        f[4] = s->kappa[i] * s->h * s->h / (diameter * diameter); /* last density increments kappa h_f^2 / d^2 */
        f[5] = s->kappaV[i] * s->h_past / diameter;              /* last divergence increments kappaV h / d */
        f[6] = (double)s->nbc[i] * inv_n;                        /* boundary share of the neighbors   */
        f[7] = off[0] * inv_n / hk;                              /* mean offset to the neighbors / hk */
        f[8] = off[1] * inv_n / hk;
        f[9] = off[2] * inv_n / hk;
        f[10] = dist * inv_n / hk;                               /* mean neighbor distance / hk       */
        f[11] = dfl / rho;                                       /* fluid share of the density        */
    }
}

void or_warm_start_targets(const or_state* s, double* out)
{
    /* after a step's solves: the increments the reference warm start left to
     * the Jacobi iterations, in units free of the time step and of the
     * particle diameter d (kappaV h / d, kappa h_f^2 / d^2) */
    const double d = 2.0 * s->p.particle_radius;
    for (uint32_t i = 0; i < s->n; ++i) {
        out[2 * i + 0] = s->kappaV[i] * s->h / d;
        out[2 * i + 1] = s->kappa[i] * s->h_future * s->h_future / (d * d);
    }
}

/* ------------------------------------------------------------------ */
/* FP64 MLP of the learned warm start (no reference counterpart)       */
/* ------------------------------------------------------------------ */
#define MLP_IN 12
#define MLP_H 64
#define MLP_OUT 2

static inline double silu(double z) { return z / (1.0 + exp(-z)); }
static inline double dsilu(double z) { const double s = 1.0 / (1.0 + exp(-z)); return s * (1.0 + z * (1.0 - s)); }

typedef struct { const double *W1, *b1, *W2, *b2, *W3, *b3, *W4, *b4; } mlp_view;

static mlp_view mlp_split(const double* p)
{
    mlp_view v;
    v.W1 = p; p += MLP_H * MLP_IN;
    v.b1 = p; p += MLP_H;
    v.W2 = p; p += MLP_H * MLP_H;
    v.b2 = p; p += MLP_H;
    v.W3 = p; p += MLP_H * MLP_H;
    v.b3 = p; p += MLP_H;
    v.W4 = p; p += MLP_OUT * MLP_H;
    v.b4 = p;
    return v;
}

static void dense(const double* W, const double* b, const double* x, int nin, int nout, double* z)
{
    for (int o = 0; o < nout; ++o) {
        double acc = b[o];
        for (int k = 0; k < nin; ++k) acc += W[o * nin + k] * x[k];
        z[o] = acc;
    }
}

void or_mlp_forward(uint32_t n, const double* params, const double* x, double* y)
{
    const mlp_view v = mlp_split(params);
#pragma omp parallel for schedule(static)
    for (long s = 0; s < (long)n; ++s) {
        double z1[MLP_H], a1[MLP_H], z2[MLP_H], a2[MLP_H], z3[MLP_H], a3[MLP_H];
        dense(v.W1, v.b1, x + s * MLP_IN, MLP_IN, MLP_H, z1);
        for (int k = 0; k < MLP_H; ++k) a1[k] = silu(z1[k]);
        dense(v.W2, v.b2, a1, MLP_H, MLP_H, z2);
        for (int k = 0; k < MLP_H; ++k) a2[k] = silu(z2[k]);
        dense(v.W3, v.b3, a2, MLP_H, MLP_H, z3);
        for (int k = 0; k < MLP_H; ++k) a3[k] = silu(z3[k]);
        dense(v.W4, v.b4, a3, MLP_H, MLP_OUT, y + s * MLP_OUT);
    }
}

double or_mlp_loss_grad(uint32_t n, const double* params, const double* x,
                        const double* target, double* grad)
{
    const mlp_view v = mlp_split(params);
    const size_t np = MLP_H * MLP_IN + MLP_H + 2 * (MLP_H * MLP_H + MLP_H) + MLP_OUT * MLP_H + MLP_OUT;
    memset(grad, 0, np * sizeof(double));
    double* gW1 = grad;
    double* gb1 = gW1 + MLP_H * MLP_IN;
    double* gW2 = gb1 + MLP_H;
    double* gb2 = gW2 + MLP_H * MLP_H;
    double* gW3 = gb2 + MLP_H;
    double* gb3 = gW3 + MLP_H * MLP_H;
    double* gW4 = gb3 + MLP_H;
    double* gb4 = gW4 + MLP_OUT * MLP_H;
    double loss = 0.0;
    const double inv_n = n ? 1.0 / (double)n : 0.0;
    for (uint32_t s = 0; s < n; ++s) {
        const double* xs = x + (size_t)s * MLP_IN;
        double z1[MLP_H], a1[MLP_H], z2[MLP_H], a2[MLP_H], z3[MLP_H], a3[MLP_H], y[MLP_OUT];
        dense(v.W1, v.b1, xs, MLP_IN, MLP_H, z1);
        for (int k = 0; k < MLP_H; ++k) a1[k] = silu(z1[k]);
        dense(v.W2, v.b2, a1, MLP_H, MLP_H, z2);
        for (int k = 0; k < MLP_H; ++k) a2[k] = silu(z2[k]);
        dense(v.W3, v.b3, a2, MLP_H, MLP_H, z3);
        for (int k = 0; k < MLP_H; ++k) a3[k] = silu(z3[k]);
        dense(v.W4, v.b4, a3, MLP_H, MLP_OUT, y);
        double dy[MLP_OUT];
        for (int o = 0; o < MLP_OUT; ++o) {
            const double e = y[o] - target[(size_t)s * MLP_OUT + o];
            loss += 0.5 * e * e * inv_n;
            dy[o] = e * inv_n;
        }
        double d3[MLP_H], d2[MLP_H], d1[MLP_H];
        for (int o = 0; o < MLP_OUT; ++o) {
            gb4[o] += dy[o];
            for (int k = 0; k < MLP_H; ++k) gW4[o * MLP_H + k] += dy[o] * a3[k];
        }
        for (int k = 0; k < MLP_H; ++k) {
            double acc = 0.0;
            for (int o = 0; o < MLP_OUT; ++o) acc += v.W4[o * MLP_H + k] * dy[o];
            d3[k] = acc * dsilu(z3[k]);
        }
        for (int o = 0; o < MLP_H; ++o) {
            gb3[o] += d3[o];
            for (int k = 0; k < MLP_H; ++k) gW3[o * MLP_H + k] += d3[o] * a2[k];
        }
        for (int k = 0; k < MLP_H; ++k) {
            double acc = 0.0;
            for (int o = 0; o < MLP_H; ++o) acc += v.W3[o * MLP_H + k] * d3[o];
            d2[k] = acc * dsilu(z2[k]);
        }
        for (int o = 0; o < MLP_H; ++o) {
            gb2[o] += d2[o];
            for (int k = 0; k < MLP_H; ++k) gW2[o * MLP_H + k] += d2[o] * a1[k];
        }
        for (int k = 0; k < MLP_H; ++k) {
            double acc = 0.0;
            for (int o = 0; o < MLP_H; ++o) acc += v.W2[o * MLP_H + k] * d2[o];
            d1[k] = acc * dsilu(z1[k]);
        }
        for (int o = 0; o < MLP_H; ++o) {
            gb1[o] += d1[o];
            for (int k = 0; k < MLP_IN; ++k) gW1[o * MLP_IN + k] += d1[o] * xs[k];
        }
    }
    return loss;
}

/* ------------------------------------------------------------------ */
/* Rigid bodies (the dynamic bodies of DFSPHCData): body particles are  */
/* moving boundary particles; see dfsph_oracle.h                        */
/* ------------------------------------------------------------------ */

static void quat_mat(const double* q, double R[3][3])
{
    const double x = q[0], y = q[1], z = q[2], w = q[3];
    R[0][0] = 1.0 - 2.0 * (y * y + z * z);
    R[0][1] = 2.0 * (x * y - z * w);
    R[0][2] = 2.0 * (x * z + y * w);
    R[1][0] = 2.0 * (x * y + z * w);
    R[1][1] = 1.0 - 2.0 * (x * x + z * z);
    R[1][2] = 2.0 * (y * z - x * w);
    R[2][0] = 2.0 * (x * z - y * w);
    R[2][1] = 2.0 * (y * z + x * w);
    R[2][2] = 1.0 - 2.0 * (x * x + y * y);
}

/* body particles at the body state: x = R(q) x0 + c, v = w x (x - c) + v_c,
 * reactions cleared (DFSPH_updateDynamicObjectParticles_kernel,
 * DFSPH_core_cuda.cu:3407-3420) */
static void place_bodies(or_state* s)
{
    for (uint32_t k = 0; k < s->ns; ++k) {
        const or_body* b = &s->body[s->sbody[k]];
        double R[3][3], r[3];
        quat_mat(b->q, R);
        const double* l = s->x0 + 3 * (size_t)k;
        double* x = s->xb + 3 * (size_t)(s->nbw + k);
        double* v = s->vb + 3 * (size_t)(s->nbw + k);
        for (int a = 0; a < 3; ++a) {
            r[a] = R[a][0] * l[0] + R[a][1] * l[1] + R[a][2] * l[2];
            x[a] = r[a] + b->c[a];
        }
        v[0] = b->w[1] * r[2] - b->w[2] * r[1] + b->v[0];
        v[1] = b->w[2] * r[0] - b->w[0] * r[2] + b->v[1];
        v[2] = b->w[0] * r[1] - b->w[1] * r[0] + b->v[2];
        s->psi[s->nbw + k] = s->psi0[k];
        s->sF[3 * k] = s->sF[3 * k + 1] = s->sF[3 * k + 2] = 0.0;
    }
}

/* force and torque about the center of mass (compute_fluid_impact_on_dynamic_body_kernel,
 * DFSPH_memory_management_cuda.cu:688-705), then one step over h: semi-implicit
 * Euler for dynamic bodies (Euler's equations in the body frame), the
 * sinusoidal motion at the body's time after the step (its clock sums the
 * steps' h), nothing for kinematic ones */
static void integrate_bodies(or_state* s)
{
    const double h = s->h;
    for (uint32_t id = 0; id < s->nbody; ++id) {
        or_body* b = &s->body[id];
        double F[3] = {0, 0, 0}, T[3] = {0, 0, 0};
        for (uint32_t k = b->first; k < b->first + b->count; ++k) {
            const double* f = s->sF + 3 * (size_t)k;
            const double* x = s->xb + 3 * (size_t)(s->nbw + k);
            /* lever arm from the float positions the GPU holds */
            const double rx = (double)(float)x[0] - b->c[0], ry = (double)(float)x[1] - b->c[1],
                         rz = (double)(float)x[2] - b->c[2];
            F[0] += f[0]; F[1] += f[1]; F[2] += f[2];
            T[0] += ry * f[2] - rz * f[1];
            T[1] += rz * f[0] - rx * f[2];
            T[2] += rx * f[1] - ry * f[0];
        }
        for (int a = 0; a < 3; ++a) { b->F[a] = F[a]; b->T[a] = T[a]; }
        b->clock += h;
        if (b->type == 2) {
            const double tn = b->clock, sn = sin(b->omega * tn), cs = cos(b->omega * tn);
            for (int a = 0; a < 3; ++a) {
                b->c[a] = b->c0[a] + b->amp * sn * b->axis[a];
                b->v[a] = b->amp * b->omega * cs * b->axis[a];
                b->w[a] = 0.0;
            }
            continue;
        }
        if (b->type != 0) continue;
        for (int a = 0; a < 3; ++a) b->v[a] += h * (F[a] / b->mass + s->p.gravity[a]);
        double R[3][3], wb[3], Tb[3];
        quat_mat(b->q, R);
        for (int a = 0; a < 3; ++a) {
            wb[a] = R[0][a] * b->w[0] + R[1][a] * b->w[1] + R[2][a] * b->w[2];
            Tb[a] = R[0][a] * T[0] + R[1][a] * T[1] + R[2][a] * T[2];
        }
        const double Lb[3] = {b->I[0] * wb[0], b->I[1] * wb[1], b->I[2] * wb[2]};
        const double gy[3] = {wb[1] * Lb[2] - wb[2] * Lb[1], wb[2] * Lb[0] - wb[0] * Lb[2], wb[0] * Lb[1] - wb[1] * Lb[0]};
        for (int a = 0; a < 3; ++a) wb[a] += h * (Tb[a] - gy[a]) / b->I[a];
        for (int a = 0; a < 3; ++a) b->w[a] = R[a][0] * wb[0] + R[a][1] * wb[1] + R[a][2] * wb[2];
        for (int a = 0; a < 3; ++a) b->c[a] += h * b->v[a];
        const double* q = b->q;
        const double* w = b->w;
        const double dx = q[3] * w[0] + (w[1] * q[2] - w[2] * q[1]);
        const double dy = q[3] * w[1] + (w[2] * q[0] - w[0] * q[2]);
        const double dz = q[3] * w[2] + (w[0] * q[1] - w[1] * q[0]);
        const double dw = -(w[0] * q[0] + w[1] * q[1] + w[2] * q[2]);
        double nq[4] = {q[0] + 0.5 * h * dx, q[1] + 0.5 * h * dy, q[2] + 0.5 * h * dz, q[3] + 0.5 * h * dw};
        const double nn = sqrt(nq[0] * nq[0] + nq[1] * nq[1] + nq[2] * nq[2] + nq[3] * nq[3]);
        for (int a = 0; a < 4; ++a) b->q[a] = nq[a] / nn;
    }
}

static int body_from_desc(const or_body_desc* d, double t_now, or_body* b)
{
    if (d->type < 0 || d->type > 2) return -1;
    const double qn = sqrt(d->rotation[0] * d->rotation[0] + d->rotation[1] * d->rotation[1] +
                           d->rotation[2] * d->rotation[2] + d->rotation[3] * d->rotation[3]);
    if (!(qn > 0.0)) return -1;
    double an = 0.0;
    if (d->type == 2) {
        an = sqrt(d->axis[0] * d->axis[0] + d->axis[1] * d->axis[1] + d->axis[2] * d->axis[2]);
        if (!(an > 0.0) || !(d->period > 0.0)) return -1;
    }
    if (d->type == 0 && !(d->mass > 0.0 && d->inertia[0] > 0.0 && d->inertia[1] > 0.0 && d->inertia[2] > 0.0))
        return -1;
    memset(b, 0, sizeof *b);
    b->type = d->type;
    b->mass = d->mass;
    for (int a = 0; a < 3; ++a) {
        b->I[a] = d->inertia[a];
        b->c[a] = b->c0[a] = d->position[a];
        b->v[a] = d->velocity[a];
        b->w[a] = d->angular_velocity[a];
        b->axis[a] = an > 0.0 ? d->axis[a] / an : 0.0;
    }
    for (int a = 0; a < 4; ++a) b->q[a] = d->rotation[a] / qn;
    b->clock = t_now;
    if (d->type == 2) {
        b->amp = d->amplitude;
        b->omega = 2.0 * M_PI / d->period;
        const double sn = sin(b->omega * t_now), cs = cos(b->omega * t_now);
        for (int a = 0; a < 3; ++a) {
            b->c[a] = b->c0[a] + b->amp * sn * b->axis[a];
            b->v[a] = b->amp * b->omega * cs * b->axis[a];
            b->w[a] = 0.0;
        }
    }
    return 0;
}

int or_add_body(or_state* s, const or_body_desc* d, uint32_t n, const double* xb_body, const double* psi)
{
    if (!n || !xb_body) return -1;
    or_body b;
    if (body_from_desc(d, s->time, &b)) return -1;
    b.first = s->ns;
    b.count = n;
    const uint32_t ns = s->ns + n, nb = s->nbw + ns, first = s->ns;
    s->body = (or_body*)realloc(s->body, (s->nbody + 1) * sizeof(or_body));
    s->body[s->nbody++] = b;
    s->x0 = (double*)realloc(s->x0, 3 * (size_t)ns * sizeof(double));
    s->psi0 = (double*)realloc(s->psi0, (size_t)ns * sizeof(double));
    s->sF = (double*)realloc(s->sF, 3 * (size_t)ns * sizeof(double));
    s->sbody = (uint32_t*)realloc(s->sbody, (size_t)ns * sizeof(uint32_t));
    s->xb = (double*)realloc(s->xb, 3 * (size_t)nb * sizeof(double));
    s->psi = (double*)realloc(s->psi, (size_t)nb * sizeof(double));
    s->vb = (double*)realloc(s->vb, 3 * (size_t)nb * sizeof(double));
    if (first == 0) memset(s->vb, 0, 3 * (size_t)s->nbw * sizeof(double));
    memcpy(s->x0 + 3 * (size_t)first, xb_body, 3 * (size_t)n * sizeof(double));
    for (uint32_t k = first; k < ns; ++k) {
        s->sbody[k] = s->nbody - 1;
        s->psi0[k] = psi ? psi[k - first] : 0.0;
    }
    s->ns = ns;
    s->nbn = nb;
    place_bodies(s);
    if (!psi) {
        /* FluidModel::computeBoundaryPsi (FluidModel.cpp:233-258) over walls and
         * bodies at their current places, FP32 predicate like the GPU's */
        const float hk2 = (float)s->hk2;
        for (uint32_t k = first; k < ns; ++k) {
            const uint32_t i = s->nbw + k;
            const float xi[3] = {(float)s->xb[3 * i], (float)s->xb[3 * i + 1], (float)s->xb[3 * i + 2]};
            double delta = s->W0;
            for (uint32_t j = 0; j < nb; ++j) {
                if (j == i) continue;
                const float xj[3] = {(float)s->xb[3 * j], (float)s->xb[3 * j + 1], (float)s->xb[3 * j + 2]};
                if (!pred32(xi, xj, hk2)) continue;
                double dd[3] = {s->xb[3 * i] - s->xb[3 * j], s->xb[3 * i + 1] - s->xb[3 * j + 1],
                                s->xb[3 * i + 2] - s->xb[3 * j + 2]};
                delta += kW(s, dd);
            }
            s->psi0[k] = s->p.density0 / delta;
        }
        place_bodies(s);
    }
    return (int)(s->nbody - 1);
}

int or_set_body(or_state* s, uint32_t id, const or_body_desc* d)
{
    if (id >= s->nbody) return -1;
    or_body b;
    if (body_from_desc(d, s->time, &b)) return -1;
    b.first = s->body[id].first;
    b.count = s->body[id].count;
    s->body[id] = b;
    return 0;
}

int or_get_body(const or_state* s, uint32_t id, or_body_desc* d, double* F, double* T)
{
    if (id >= s->nbody) return -1;
    const or_body* b = &s->body[id];
    if (d) {
        memset(d, 0, sizeof *d);
        d->type = b->type;
        d->mass = b->mass;
        for (int a = 0; a < 3; ++a) {
            d->inertia[a] = b->I[a];
            d->position[a] = b->c[a];
            d->velocity[a] = b->v[a];
            d->angular_velocity[a] = b->w[a];
            d->axis[a] = b->axis[a];
        }
        for (int a = 0; a < 4; ++a) d->rotation[a] = b->q[a];
        d->amplitude = b->amp;
        d->period = b->omega > 0.0 ? 2.0 * M_PI / b->omega : 0.0;
    }
    for (int a = 0; a < 3; ++a) {
        if (F) F[a] = b->F[a];
        if (T) T[a] = b->T[a];
    }
    return 0;
}

int or_body_forces(const or_state* s, uint32_t id, double* f)
{
    if (id >= s->nbody) return -1;
    memcpy(f, s->sF + 3 * (size_t)s->body[id].first, 3 * (size_t)s->body[id].count * sizeof(double));
    return 0;
}

int or_body_psi(const or_state* s, uint32_t id, double* psi)
{
    if (id >= s->nbody) return -1;
    memcpy(psi, s->psi0 + s->body[id].first, (size_t)s->body[id].count * sizeof(double));
    return 0;
}
