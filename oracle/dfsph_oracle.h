/*
 * dfsph_oracle.h -- CPU restatement of the reference DFSPH time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product (libdfsph_hip.so, libdfsph_host.so) never
 * links or calls it.
 *
 * Precision: FP64 throughout (the reference CPU path uses Real=double,
 * SPlisHSPlasH/BasicTypes.h:4-12).  The only FP32 code is the neighbor
 * predicate of or_neighbors_fp32(), which restates the GPU's strict
 * |xi-xj|^2 < h^2 test in single precision (DFSPH_macro_cuda.h:124,149) so
 * that neighbor sets can be compared bit-exactly.
 *
 * Two semantic modes (the GPU/CPU differences of the reference):
 *   OR_MODE_GPU  follows SPlisHSPlasH/DFSPH/DFSPH_core_cuda.cu
 *   OR_MODE_CPU  follows SPlisHSPlasH/DFSPH/TimeStepDFSPH.cpp
 * See dfsph_oracle.c for the per-function citations.
 *
 * Parity pins (tests/test_oracle.py): closed-form kernel values, kernel
 * normalisation by quadrature, gradient antisymmetry, brute force vs grid
 * neighbor search, perfect-lattice density, zero divergence source for
 * uniform translation and rigid rotation, fluid momentum conservation of
 * a Jacobi update, zero XSPH for uniform velocity, a hydrostatic column
 * converging to the tolerance.  The exact FP32 results and iteration
 * counts of the reference CUDA build cannot be produced here (no CUDA
 * toolchain): that comparison is "parity unpinned".
 */
#ifndef DFSPH_ORACLE_H
#define DFSPH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_MODE_GPU = 0, OR_MODE_CPU = 1 };

typedef struct or_params {
    double particle_radius;   /* r; kernel support h_k = 4 r              */
    double density0;          /* rho_0                                     */
    double viscosity;         /* XSPH coefficient                          */
    double gravity[3];
    double dt;                /* initial / fixed time step                 */
    int    mode;              /* OR_MODE_GPU or OR_MODE_CPU                */
    int    kernel_table;      /* 0 analytic cubic, 1 10000-sample table    */
    int    fp32_predicate;    /* neighbor test on float-rounded positions  */
    int    enable_divergence;
    uint32_t max_iter_v, max_iter, min_iter_v, min_iter;
    double max_error_v, max_error;      /* percent                         */
    int    use_cfl;
    double cfl_factor, dt_min, dt_max;
    double cfl_seed;          /* max |v|^2 seed: 0 (GPU) or 0.1 (CPU)      */
    int    remove_escaped;    /* checkParticlesPositions(2) each step      */
    uint32_t max_neighbors;   /* per particle, fluid + boundary            */
} or_params;

typedef struct or_stats {
    uint32_t iters_v, iters;
    double   err_v, err;      /* final average errors (density units)      */
    double   dt_used;         /* h of the step                             */
    uint32_t removed;
    uint32_t n_fluid;
    uint32_t max_nbr;
    uint32_t overflow;
} or_stats;

typedef struct or_state or_state;

/* fields for or_get / or_set */
enum {
    OR_F_POS = 0, OR_F_VEL = 1, OR_F_MASS = 2, OR_F_DENSITY = 3,
    OR_F_FACTOR = 4, OR_F_DENSITY_ADV = 5, OR_F_KAPPA = 6, OR_F_KAPPAV = 7,
    OR_F_ACC = 8, OR_F_ID = 9
};

void      or_default_params(or_params* p);
or_state* or_create(const or_params* p);
void      or_destroy(or_state* s);

/* fluid: x,v are 3n doubles, m n doubles (NULL -> 0.8 (2r)^3 rho0), id n (NULL -> 0..n-1) */
int  or_set_fluid(or_state* s, uint32_t n, const double* x, const double* v,
                  const double* m, const uint32_t* id);
/* boundary: x 3n doubles; psi NULL -> computed (FluidModel::computeBoundaryPsi) */
int  or_set_boundary(or_state* s, uint32_t n, const double* x, const double* psi);
int  or_get_boundary_psi(const or_state* s, double* out);
uint32_t or_num_fluid(const or_state* s);
/* copy a field out (3n for vectors, n for scalars; id as uint32 through double) */
int  or_get(const or_state* s, int field, double* out);
int  or_set(or_state* s, int field, const double* in);
int  or_get_ids(const or_state* s, uint32_t* out);
void or_get_time(const or_state* s, double* h, double* h_future, double* h_past, double* t);
void or_set_time_step(or_state* s, double h);

/* phases of DFSPHCUDA::step() (DFSPH_CUDA.cpp:82-794) */
int      or_neighbor_search(or_state* s);
uint32_t or_divergence_solve(or_state* s, double* err_out);
void     or_external_forces(or_state* s);
double   or_cfl(or_state* s);
void     or_update_time_step(or_state* s, double h_future);
void     or_update_vel(or_state* s);
uint32_t or_pressure_solve(or_state* s, double* err_out);
void     or_update_pos(or_state* s);
void     or_step_end(or_state* s);
uint32_t or_check_positions(or_state* s, int remove);
int      or_step(or_state* s, or_stats* st);

/* neighbor lists of the last search, row-major [i*max_neighbors + k],
 * fluid neighbors first then boundary neighbors */
int  or_get_neighbors(const or_state* s, uint32_t* nf, uint32_t* nb, uint32_t* list);

/* stand-alone FP32-predicate neighbor search (bit-exact reference for the
 * GPU lists).  xf: 3*n_f floats, xb: 3*n_b floats.  brute != 0 -> O(N^2).
 * Lists row-major [i*max + k]; returns the largest total count seen (lists
 * are truncated at max, counts are not). */
uint32_t or_neighbors_fp32(uint32_t n_f, const float* xf, uint32_t n_b, const float* xb,
                           float hk, int brute, uint32_t max,
                           uint32_t* nf, uint32_t* nb, uint32_t* list);

/* solver restriction in GPU mode (restriction_mode, count_active,
 * count_active_neighbors, true_particle_count, DFSPH_c_arrays_structure.h:410-418);
 * tags (n, may be NULL) are the reference's tag values */
int or_set_restriction(or_state* s, int mode, uint32_t count_active, uint32_t count_active_neighbors,
                       long long true_particle_count, const uint32_t* tags);

/* Rigid bodies (the dynamic bodies of DFSPHCData, vector_dynamic_bodies_data):
 * their particles are moving boundary particles after the walls; v_j enters
 * the density sources, every Jacobi update pushes -m_i k_i psi_j gradW_ij onto
 * body particle j, and or_update_pos sums the reactions into force and torque
 * and moves the body (type 0 dynamic: semi-implicit Euler under gravity and
 * the fluid force; 1 kinematic: left as set; 2 sine: position + amplitude
 * sin(2 pi t / period) axis).  Positions in the body frame are relative to
 * the center of mass; rotation is a unit quaternion (x, y, z, w).  psi NULL:
 * computed over the walls and every body at its current place. */
typedef struct or_body_desc {
    int type;
    double mass, inertia[3], position[3], velocity[3], rotation[4], angular_velocity[3], axis[3];
    double amplitude, period;
} or_body_desc;
int or_add_body(or_state* s, const or_body_desc* b, uint32_t n, const double* x_body, const double* psi);
int or_set_body(or_state* s, uint32_t id, const or_body_desc* b);
int or_get_body(const or_state* s, uint32_t id, or_body_desc* b, double* force, double* torque);
int or_body_forces(const or_state* s, uint32_t id, double* f);   /* 3 n, last step */
int or_body_psi(const or_state* s, uint32_t id, double* psi);

/* A19: BufferFluidSurface and the open boundary */
typedef struct or_surface {
    int type, reversed;       /* 0 plane, 1 cuboid, 2 cylinder, 3 mesh, 4 sphere, 5 star */
    double o[3], n[3], center[3], half[3];   /* mesh: center is the offset of the vertices */
    double radius, radius_internal;
    uint32_t points;
    double dir[3];
    const double* tri;        /* mesh: 9 doubles per triangle (its three vertices) */
    uint32_t ntri;
} or_surface;
int or_surface_inside(const or_surface* S, const double* p, double* margin);
uint32_t or_inflow(const or_state* s, const or_surface* boundary, uint32_t nc, const double* cand, double min_dist,
                   double max_density, double mass, uint8_t* accept, double* vel, double* dmin_out,
                   double* dens_out);
uint32_t or_outflow(const or_state* s, const or_surface* interior, const or_surface* surface, uint8_t* keep);
/* one density-extraction pass of the rest-fluid loader (RestFluidLoader.cu:2700-3130,
 * :773-890) on the lists of the last or_neighbor_search: keep[i] = 0 for a
 * candidate (air[i] == 0 or air NULL) the boundary alone makes denser than
 * boundary_limit, or one denser than limit with no denser listed candidate
 * above limit; returns the count */
uint32_t or_rest_extract(or_state* s, double limit, double boundary_limit, const uint8_t* air, uint8_t* keep,
                         double* avg, double* rho_out);

/* kernel evaluation (A3): analytic or table per flag */
double or_kernel_W(double r, double hk, int table);
void   or_kernel_gradW(const double* r3, double hk, int table, double* out3);

/* Learned warm start (no reference counterpart), mode GPU state, after a
 * neighbor search and before the divergence solve: 12 features per fluid
 * particle (n*12):
 *   0 rho / rho0 - 1 (fluid + boundary)    1 sum m_j (v_i - v_j).gradW h / rho0
 *   2 (n_fluid + n_boundary) / 40          3 |v| h / (2r)
 *   4 kappa h^2 / d^2 (last density increments)  5 kappaV h_past / d (last divergence
 *   increments), d = 2r
 *   6 n_boundary / n                       7-9 mean (x_j - x_i) / hk
 *   10 mean |x_i - x_j| / hk              11 fluid share of the density */
void or_features(const or_state* s, double* out);
/* after a step's two solves, before the step end: the training targets, the
 * solves' increments in units free of dt and d = 2r (kappaV h / d,
 * kappa h_future^2 / d^2), n*2 */
void or_warm_start_targets(const or_state* s, double* out);

/* FP64 per-particle MLP used by the learned warm start (12-64-64-64-2, SiLU).
 * W layout: W1[64*12], b1[64], W2[64*64], b2[64], W3[64*64], b3[64], W4[2*64], b4[2]
 * (row-major, out x in).  x: n*12 -> y: n*2. */
void or_mlp_forward(uint32_t n, const double* params, const double* x, double* y);
/* mean squared error loss over n samples and 2 outputs, and its gradient
 * with respect to every parameter (same layout as params). */
double or_mlp_loss_grad(uint32_t n, const double* params, const double* x,
                        const double* target, double* grad);

#ifdef __cplusplus
}
#endif
#endif
