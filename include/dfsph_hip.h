/*
 * dfsph_hip.h -- C ABI of the MI355X (gfx950) DFSPH time-step library
 * (libdfsph_hip.so).
 *
 * One entry point per function of the reference's GPU DFSPH path; each
 * declaration cites the reference interface it replaces (paths relative to
 * the reference repository root).  The reference passes SPH::DFSPHCData& to
 * C++ free functions that exit() on error; here a dfsph_handle owns all
 * device memory and every call returns a status code.
 *
 * Conventions
 *   - every function returns DFSPH_OK (0) or a DFSPH_ERR_* code; the text of
 *     the last error of the calling thread is in dfsph_last_error();
 *   - plain pointers and sizes only; host arrays are float32 / uint32, xyz
 *     arrays are interleaved [x0 y0 z0 x1 ...];
 *   - one handle per host thread; work is queued on one HIP stream per
 *     handle and the calls that return host values synchronise it;
 *   - particles are kept in cell-sorted order on the device and reordered by
 *     every sorting neighbor search; DFSPH_FIELD_ID gives each slot's
 *     persistent particle id (the id passed to dfsph_upload_fluid).
 */
#ifndef DFSPH_HIP_H
#define DFSPH_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DFSPH_OK 0
#define DFSPH_ERR_HIP 1            /* a HIP runtime call failed             */
#define DFSPH_ERR_ARG 2            /* invalid argument                      */
#define DFSPH_ERR_CAPACITY 3       /* more particles than the capacity      */
#define DFSPH_ERR_NEIGHBOR_OVERFLOW 4 /* a particle has more than max_neighbors */
#define DFSPH_ERR_STATE 5          /* call out of order (e.g. no boundary)  */
#define DFSPH_ERR_NO_MODEL 6       /* learned warm start without weights    */

/* warm_start_mode */
#define DFSPH_WARMSTART_REFERENCE 0   /* kappa / kappaV rescaled from the last step */
#define DFSPH_WARMSTART_LEARNED 1     /* kappa / kappaV predicted by the MLP          */
#define DFSPH_WARMSTART_NONE 2        /* no warm start (cold Jacobi start)             */

typedef struct dfsph_params {
    double particle_radius;   /* r; kernel support 4r (DFSPH_c_arrays_structure.cpp:857-869) */
    double density0;          /* rest density rho0                                           */
    double viscosity;         /* XSPH coefficient (DFSPH_CUDA.cpp:98)                         */
    double gravity[3];        /* (DFSPH_c_arrays_structure.cpp:778)                          */
    double dt;                /* initial time step h                                         */
    uint32_t max_iter_v;      /* divergence solver iteration cap                              */
    uint32_t max_iter;        /* density solver iteration cap                                 */
    uint32_t min_iter_v;      /* 3 in the reference (DFSPH_core_cuda.cu:828)                  */
    uint32_t min_iter;        /* 2 in the reference (DFSPH_core_cuda.cu:1378)                 */
    double max_error_v;       /* percent, divergence                                         */
    double max_error;         /* percent, density                                            */
    int32_t use_cfl;          /* 1: h_future from dfsph_cfl inside dfsph_step                 */
    double cfl_factor;        /* h = cfl_factor * 0.4 * 2r / max|v + a h|                     */
    double dt_min, dt_max;
    int32_t enable_divergence; /* m_enableDivergenceSolver (DFSPH_CUDA.h:64)                  */
    int32_t check_positions_mode; /* checkParticlesPositions: -1 off, 0 count, 2 remove       */
    uint32_t max_neighbors;   /* per-particle list capacity, fluid + boundary (<= 255)       */
    uint32_t fluid_capacity;  /* 0: 1.3 x the first upload (DFSPH_c_arrays_structure.cpp:170-172) */
    int32_t warm_start_mode;  /* DFSPH_WARMSTART_*                                            */
    uint32_t check_interval;  /* solver iterations queued per convergence readback (0: 4)     */
    int32_t list_mode;        /* DFSPH_LISTS_*: neighbor-list form (results are identical)     */
    int32_t learned_gate;     /* DFSPH_LEARNED_*: when the learned warm start runs             */
} dfsph_params;

/* learned_gate (warm_start_mode DFSPH_WARMSTART_LEARNED only) */
#define DFSPH_LEARNED_ALWAYS 0       /* every step                                                  */
#define DFSPH_LEARNED_ABOVE_FLOORS 1 /* only after a step whose divergence or density solve ran
                                        more than its minimum iterations; otherwise the step uses
                                        the reference warm start and skips the features and the
                                        network (at the floors no warm start saves an iteration) */

/* list_mode: how sweep blocks (dfsph_info.sweep_block particles) read their fluid neighbors */
#define DFSPH_LISTS_AUTO 0       /* LDS-staged set when it fits, else 16-bit set ids, else 32-bit */
#define DFSPH_LISTS_NO_LDS 1     /* 16-bit set ids read through the set's global ids               */
#define DFSPH_LISTS_GLOBAL 2     /* 32-bit global ids only                                          */

typedef struct dfsph_step_stats {
    uint32_t iters_v, iters;  /* divergence / density Jacobi iterations */
    double err_v, err;        /* final average source: divergence (rho/s), density (rho) */
    double dt;                /* h used for the position update        */
    double time;              /* simulated time after the step         */
    uint32_t n_fluid;         /* active fluid particles after the step */
    uint32_t removed;         /* particles removed by the position check */
    uint32_t max_neighbors_seen;
    uint32_t overflow;        /* particles whose list was truncated    */
    uint64_t neighbors_fluid_total;    /* stored neighbors of this step's lists */
    uint64_t neighbors_boundary_total;
} dfsph_step_stats;

typedef struct dfsph_info {
    uint32_t n_fluid, n_boundary, fluid_capacity, max_neighbors;
    uint32_t grid_dim[3];
    double grid_origin[3];
    double cell_size;
    double h, h_future, h_past, time;
    double bbox_min[3], bbox_max[3];
    uint32_t max_neighbors_seen;   /* largest count of the last neighbor search */
    uint32_t overflow;             /* particles whose list was truncated        */
    uint32_t escaped;              /* outside the box after the last position update */
    uint32_t unstaged_blocks;      /* sweep blocks whose neighbor set exceeded
                                      the LDS stage (32-bit global lists)      */
    uint64_t neighbors_fluid_total;    /* stored fluid neighbors, all particles    */
    uint64_t neighbors_boundary_total; /* stored boundary neighbors, all particles */
    uint64_t neighbors_fluid_unstaged; /* of the fluid ones, in 32-bit global lists */
    uint32_t sweep_block;          /* particles per sweep block (owner of a neighbor set) */
} dfsph_info;

/* fields for dfsph_download / dfsph_upload_field */
#define DFSPH_FIELD_POS 0          /* float xyz (3n)   */
#define DFSPH_FIELD_VEL 1          /* float xyz (3n)   */
#define DFSPH_FIELD_MASS 2         /* float            */
#define DFSPH_FIELD_DENSITY 3      /* float            */
#define DFSPH_FIELD_FACTOR 4       /* float (alpha)    */
#define DFSPH_FIELD_DENSITY_ADV 5  /* float            */
#define DFSPH_FIELD_KAPPA 6        /* float            */
#define DFSPH_FIELD_KAPPAV 7       /* float            */
#define DFSPH_FIELD_ACC 8          /* float xyz (3n)   */
#define DFSPH_FIELD_ID 9           /* uint32           */
#define DFSPH_FIELD_NBR_FLUID 10   /* uint32 fluid neighbor count    */
#define DFSPH_FIELD_NBR_BOUNDARY 11 /* uint32 boundary neighbor count */
#define DFSPH_FIELD_BOUNDARY_POS 12 /* float xyz (3 n_boundary), cell-sorted */
#define DFSPH_FIELD_BOUNDARY_PSI 13 /* float (n_boundary), cell-sorted */
#define DFSPH_FIELD_TAG 14         /* uint32 dynamic-window / open-boundary tag (also uploadable) */
#define DFSPH_FIELD_KAPPA_TOTAL 15  /* float: last density solve's total, kappa h_f^2 (also uploadable) */
#define DFSPH_FIELD_KAPPAV_TOTAL 16 /* float: last divergence solve's total, kappaV h (also uploadable) */

void dfsph_default_params(dfsph_params* p);
const char* dfsph_last_error(void);

/* DFSPHCData::DFSPHCData(FluidModel*) + DFSPHCUDA::DFSPHCUDA
 * (DFSPH_c_arrays_structure.cpp:852-944, DFSPH_CUDA.cpp:38-75) */
int dfsph_create(const dfsph_params* p, void** handle);
/* DFSPHCData::~DFSPHCData (DFSPH_c_arrays_structure.cpp:946-973) */
int dfsph_destroy(void* handle);
/* the parameters the handle was created with (the DFSPHCData members,
 * DFSPH_c_arrays_structure.h:297-420, read back) */
int dfsph_get_params(void* handle, dfsph_params* out);
/* The HIP stream the handle queues its work on (a hipStream_t; NULL: the
 * handle's own non-blocking stream, made at create).  The reference queues
 * on the default stream; a caller's stream orders the solver against the
 * caller's own kernels and copies.  Setting it waits for the work already
 * queued; the caller keeps the stream alive while the handle uses it and
 * keeps it at the default priority. */
int dfsph_set_stream(void* handle, void* stream);
int dfsph_get_stream(void* handle, void** stream);

/* Static boundary set: allocate_UnifiedParticleSet_cuda + load_UnifiedParticleSet_cuda
 * (DFSPH_memory_management_cuda.cu:502-564, :566-650), the one-time boundary cell table
 * (DFSPH_c_arrays_structure.cpp:1016) and initGridOffset (:1050-1098).
 * psi == NULL: computed on the device as FluidModel::computeBoundaryPsi
 * (FluidModel.cpp:233-258). */
int dfsph_upload_boundary(void* handle, const float* xyz, const float* psi, uint32_t n);

/* Fluid set: UnifiedParticleSet::reset<FluidModel> + first cell sort
 * (DFSPH_c_arrays_structure.cpp:1035-1037).  mass == NULL: 0.8 (2r)^3 rho0
 * (FluidModel.cpp:94-108); vel == NULL: zero; id == NULL: 0..n-1.
 * Resets kappa, kappaV and the time-step history. */
int dfsph_upload_fluid(void* handle, const float* xyz, const float* vel, const float* mass,
                       const uint32_t* id, uint32_t n);

/* cuda_neighborsSearch (DFSPH_core_cuda.h:65, DFSPH_core_cuda.cu:2651-2770):
 * cell hash, deterministic counting sort, reorder, neighbor lists.  Queued
 * without a host round trip; a list overflow is reported by dfsph_get_info. */
int dfsph_neighbor_search(void* handle, int sort);

/* The density + factor part of cuda_divergence_warmstart_init (DFSPH_core_cuda.h:24,
 * DFSPH_core_cuda.cu:220-376) without the warm start, for steps that skip the
 * divergence solver (the reference then keeps stale values, DFSPH_CUDA.cpp:261-270). */
int dfsph_compute_density_factor(void* handle);

/* cuda_divergenceSolve (DFSPH_core_cuda.h:31, DFSPH_core_cuda.cu:785-861); also
 * computes density and factor (the fused warm-start init, :220-376). */
int dfsph_divergence_solve(void* handle, uint32_t max_iter, double max_error,
                           uint32_t* iters, double* err);

/* cuda_externalForces (DFSPH_core_cuda.h:55, DFSPH_core_cuda.cu:1613-1627): a = g + XSPH */
int dfsph_external_forces(void* handle);

/* cuda_CFL (DFSPH_core_cuda.h:77, DFSPH_core_cuda.cu:3321-3399); applies the new h
 * with DFSPHCData::updateTimeStep and returns it in dt_out (a host round trip).
 * dt_out == NULL: no round trip, the value stays on the device (dfsph_get_time). */
int dfsph_cfl(void* handle, double min_dt, double cfl_factor, double max_dt, double* dt_out);

/* DFSPHCData::h, h_future and the simulated time (DFSPH_c_arrays_structure.h:
 * 248-252) as the device holds them.  Right after dfsph_divergence_solve or
 * dfsph_pressure_solve this reads the state their last loop record carried
 * (no synchronisation); otherwise it waits for the stream. */
int dfsph_get_time(void* handle, double* h, double* h_future, double* time);

/* DFSPHCData::getFluidParticlesCount (DFSPH_c_arrays_structure.h:441): the
 * active fluid count, a host-side value (no synchronisation). */
int dfsph_fluid_count(void* handle, uint32_t* n);

/* DFSPHCData::updateTimeStep (DFSPH_c_arrays_structure.cpp:1100-1106) */
int dfsph_update_time_step(void* handle, double h_future);

/* cuda_update_vel (DFSPH_core_cuda.h:75, DFSPH_core_cuda.cu:3047-3070) */
int dfsph_update_vel(void* handle);

/* cuda_pressureSolve (DFSPH_core_cuda.h:48, DFSPH_core_cuda.cu:1338-1419) */
int dfsph_pressure_solve(void* handle, uint32_t max_iter, double max_error,
                         uint32_t* iters, double* err);

/* cuda_update_pos (DFSPH_core_cuda.h:76, DFSPH_core_cuda.cu:3127-3256) */
int dfsph_update_pos(void* handle);

/* DFSPHCData::onSimulationStepEnd (DFSPH_c_arrays_structure.cpp:1108-1118) and the
 * TimeManager time advance of DFSPHCUDA::step (DFSPH_CUDA.cpp:441-452) */
int dfsph_step_end(void* handle);

/* Restore the time-step history when a run resumes from a snapshot (the
 * reference reloads only dt from general_data.txt, DFSPH_CUDA.cpp:1887-1979):
 * h, h_past and the simulated time; h_future = h. */
int dfsph_set_time(void* handle, double h, double h_past, double time);

/* DFSPHCData::checkParticlesPositions -> check_particles_positions_cuda
 * (DFSPH_c_arrays_structure.h:486-492, SPH_other_systems_cuda.cu:1283-1330):
 * mode 0 counts, mode 2 removes the particles outside the boundary bounding box
 * (order of the others kept).  Called right after dfsph_step_end it first reads
 * the count dfsph_update_pos made, from the step end's record, and returns at
 * once when no particle left the box. */
int dfsph_check_particles_positions(void* handle, int mode, uint32_t* count);

/* remove_tagged_particles (SPH_other_systems_cuda.cu:1606-1641): remove the
 * particles whose slot has keep[slot] == 0 (device order), order kept. */
int dfsph_remove_particles(void* handle, const uint8_t* keep, uint32_t* n_out);

/* add_particles_cuda (DFSPH_memory_management_cuda.cu:1036-1062): append
 * particles after the active ones (kappa = kappaV = 0). */
int dfsph_add_particles(void* handle, const float* xyz, const float* vel, const float* mass,
                        const uint32_t* id, uint32_t n);

/* DFSPHCUDA::step (DFSPH_CUDA.cpp:82-794): the sequence above in the
 * reference order, with the library's kernel fusions. */
int dfsph_step(void* handle, dfsph_step_stats* stats);

/* The loading of DynamicWindow::loadDataToSimulation (DynamicWindow.cu:2624-3140)
 * by the rule of the open boundary's inflow (OpenBoundariesSimple.cu:360-508)
 * without its direction test: each candidate position (e.g. a background fluid
 * buffer at rest) inside the boundary's box is added when no fluid particle is
 * within min_distance and the SPH density there (fluid and boundary) is at most
 * max_density; velocity 0, or the neighbors' (m W)-weighted mean with
 * interpolate_velocity.  Candidates are decided against the fluid before the
 * call (not against each other).  New particles: mass of fluid particle 0,
 * kappa = kappaV = 0, fresh ids, untagged, appended in candidate order; the
 * capacity grows as needed. */
int dfsph_load_fluid(void* handle, const float* xyz, uint32_t n, float min_distance, float max_density,
                     int interpolate_velocity, uint32_t* added);

/* change_max_particle_number (DFSPH_memory_management_cuda.cu:962-1018): grow
 * the fluid capacity to max_particles (no change when it is not larger),
 * keeping the particles in their order; the next step rebuilds the neighbor
 * lists.  dfsph_open_boundary_apply calls it by itself (capacity doubled once
 * more than 3/4 is in use, OpenBoundariesSimple.cu:570-573). */
int dfsph_change_max_particle_number(void* handle, uint32_t max_particles);

/* read_UnifiedParticleSet_cuda (DFSPH_memory_management_cuda.cu:652-674):
 * copy a field of the active particles (device order) to host memory.
 * count must equal the number of elements of the field. */
int dfsph_download(void* handle, int field, void* dst, uint32_t count);
/* overwrite a fluid field (device order): POS, VEL, MASS, KAPPA, KAPPAV, TAG,
 * KAPPA_TOTAL, KAPPAV_TOTAL, and the derived DENSITY and DENSITY_ADV (a
 * restricted solve keeps them for the particles it leaves out and sums the
 * stored DENSITY_ADV of those into its average error).  The derived fields
 * are not carried by reorderings (sorting searches, removals, the active-first
 * order): they are valid from the next density computation of a particle. */
int dfsph_upload_field(void* handle, int field, const void* src, uint32_t count);

/* neighbor lists of the last search, device order, row-major [i*max + k]:
 * fluid neighbor slots first, then boundary slots (DFSPH_macro_cuda.h:208-287) */
int dfsph_get_neighbors(void* handle, uint32_t* n_fluid, uint32_t* n_boundary,
                        uint32_t* list, uint32_t max);

int dfsph_get_info(void* handle, dfsph_info* info);
int dfsph_synchronize(void* handle);
/* Statistics of the last neighbor search, from the per-particle counts that
 * DFSPH_neighborsSearch_kernel stores (DFSPH_core_cuda.cu:2513-2516): the
 * largest list, the particles whose list was truncated at max_neighbors and
 * the stored neighbor totals.  No stream synchronisation when a solve has
 * already brought them back after that search. */
int dfsph_neighbor_stats(void* handle, uint32_t* max_seen, uint32_t* overflow, uint64_t* fluid_total,
                         uint64_t* boundary_total);

/* Per-kernel HIP-event timing, the device-side counterpart of the phase timers
 * of DFSPHCUDA::step (DFSPH_CUDA.cpp:185-194, 480-495) and SegmentedTiming
 * (SPlisHSPlasH/Utilities/SegmentedTiming.h:22-84).  mask: one bit per
 * DFSPH_KERNEL_* id (0 turns timing off).  The single-kernel sweeps (ids 2-12)
 * carry an event pair in their own dispatch (hipExtLaunchKernel), so their
 * duration is the kernel's execution alone; the sort, the search (with its
 * one-block finalize) and the MLP are bracketed by event records on the
 * handle's stream.  Velocity sweeps that exit at once because their loop had
 * already converged are not counted. */
#define DFSPH_KERNEL_SORT 0          /* cell hash, scan, counting sort, reorder */
#define DFSPH_KERNEL_NEIGHBORS 1     /* neighbor list build                     */
#define DFSPH_KERNEL_DIV_WS_INIT 2   /* density + factor + source + warm start  */
#define DFSPH_KERNEL_DIV_INIT 3      /* divergence source (init)                */
#define DFSPH_KERNEL_DIV_JACOBI 4    /* divergence velocity sweep               */
#define DFSPH_KERNEL_DIV_LOOP_END 5  /* divergence source + FP64 reduction      */
#define DFSPH_KERNEL_XSPH 6          /* XSPH + gravity + CFL maximum            */
#define DFSPH_KERNEL_PRESSURE_WS 7   /* velocity update + pressure warm start   */
#define DFSPH_KERNEL_DENS_INIT 8     /* density source (init)                   */
#define DFSPH_KERNEL_DENS_JACOBI 9   /* density velocity sweep                  */
#define DFSPH_KERNEL_DENS_LOOP_END 10 /* density source + FP64 reduction        */
#define DFSPH_KERNEL_UPDATE_POS 11   /* position update + escape count          */
#define DFSPH_KERNEL_FEATURES 12     /* learned warm start: feature sweep       */
#define DFSPH_KERNEL_MLP_INFER 13    /* learned warm start: MLP inference       */
#define DFSPH_KERNEL_COUNT 14
int dfsph_profile_enable(void* handle, uint32_t mask);
int dfsph_profile_reset(void* handle);
/* launches and summed duration (ms) of one kernel id since the last reset */
int dfsph_profile_read(void* handle, uint32_t kernel, uint32_t* launches, double* total_ms);

/* ---- dynamic window: solver restriction and tags -------------------------
 * The reference restricts its solver to a window of the fluid while it
 * stabilises a newly loaded region (DFSPHCData::restriction_mode, count_active,
 * count_active_neighbors, true_particle_count, DFSPH_c_arrays_structure.h:410-418,
 * set at DynamicWindow.cu:3047-3048, 3336, 3736-3739).  Tags carry the
 * reference's values (cuda_compilation_center.cu:2-11) and move with the
 * particles through sorts and compactions. */
#define DFSPH_TAG_ACTIVE 10000000u
#define DFSPH_TAG_ACTIVE_NEIGHBORS 20000000u
#define DFSPH_TAG_UNTAGGED 40000000u
#define DFSPH_TAG_AIR 50000000u
#define DFSPH_TAG_REMOVAL 310000000u
/* mode 0: whole fluid.  mode 1: velocity / position updates on the particles
 * tagged ACTIVE, density / factor / source / neighbor lists also on those
 * tagged ACTIVE_NEIGHBORS (DFSPH_core_cuda.cu:225-231, 500-506).  mode 2: the
 * same on the index prefixes [0, count_active) and
 * [0, count_active + count_active_neighbors) (dfsph_step then searches without
 * reordering).  true_particle_count >= 0 replaces N in the average errors
 * (DFSPH_core_cuda.cu:791-795, 1339-1343).  Restricted-out particles keep
 * their stored rho_adv in the error sums, as the reference's reductions do.
 * Neighbor lists of an unrestricted search cover every particle and stay
 * valid under a new restriction; those of a restricted search do not. */
int dfsph_set_restriction(void* handle, int mode, uint32_t count_active, uint32_t count_active_neighbors,
                          int64_t true_particle_count);
/* tag_neighbors_of_tagged_kernel<untagged_only, true> (RestFluidLoader.cu:2070-2101):
 * the stored fluid neighbors of every particle tagged tag_in get tag_out (only
 * untagged ones if untagged_only); changed: particles whose tag changed.  The
 * tags read are those before the call, so the result is order independent. */
int dfsph_tag_neighbors(void* handle, uint32_t tag_in, uint32_t tag_out, int untagged_only, uint32_t* changed);
/* count_particle_with_tag_kernel (RestFluidLoader.cu:2238-2252) */
int dfsph_count_tagged(void* handle, uint32_t tag, uint32_t* count);
/* One density-extraction pass of the rest-fluid loader
 * (RestFLuidLoader::tagDataToSurface, RestFluidLoader.cu:2700-3130, selection
 * particle_selection_rule_1_kernel, :773-890), on the lists of the last
 * unrestricted search: the densities are computed (boundary included).
 * Particles tagged DFSPH_TAG_AIR (the reference's air particles above the
 * fluid surface) only add density and stay.  Any other particle is a
 * candidate and is removed (order kept) when the boundary alone gives it more
 * than boundary_limit, or when its density exceeds limit and no listed
 * candidate above limit is denser (of equal densities the lower index goes),
 * so of two crowded neighbors one goes per pass.  removed: the count;
 * avg_density: the mean density of the kept candidates, before the removal. */
int dfsph_remove_density_peaks(void* handle, float limit, float boundary_limit, uint32_t* removed,
                               double* avg_density);
/* Stable reorder: ACTIVE, then ACTIVE_NEIGHBORS, then the rest (the prefix
 * order of restriction mode 2; the reference radix-sorts by tag,
 * DynamicWindow.cu:2990-2998); returns both counts. */
int dfsph_order_active_first(void* handle, uint32_t* count_active, uint32_t* count_active_neighbors);
/* remove_tagged_particles (SPH_other_systems_cuda.cu:1606-1641) for one tag, order kept */
int dfsph_remove_tagged(void* handle, uint32_t tag, uint32_t* removed);

/* ---- open boundary ---------------------------------------------------------
 * OpenBoundariesSimple (OpenBoundariesSimple.cu:117-621) with the analytic
 * surfaces of BufferFluidSurface (BufferFluidSurface.cu:421-1070). */
#define DFSPH_SURFACE_PLANE 0      /* inside: (p - o) . n >= 0                     */
#define DFSPH_SURFACE_CUBOID 1     /* |p - center| < half per axis                 */
#define DFSPH_SURFACE_CYLINDER 2   /* axis y: |dy| <= half[1], radial dist < radius */
#define DFSPH_SURFACE_MESH 3       /* closed triangle mesh (dfsph_mesh_upload), moved by center */
#define DFSPH_SURFACE_SPHERE 4     /* |p - center| < radius                         */
#define DFSPH_SURFACE_STAR 5       /* prism along y, linear-sided star in x-z       */
typedef struct dfsph_surface {
    int32_t type;                  /* DFSPH_SURFACE_*                               */
    int32_t reversed;              /* exterior becomes interior                     */
    float o[3], n[3];              /* plane                                         */
    float center[3], half[3];      /* cuboid; cylinder / star use half[1]; mesh: offset */
    float radius, radius_internal; /* cylinder, sphere; star outer / inner radius   */
    uint32_t points;               /* star                                          */
    float direction[3];            /* star: unit direction of the first point (x-z) */
    uint32_t mesh;                 /* mesh: the id dfsph_mesh_upload returned        */
} dfsph_surface;
/* MeshPerso (BufferFluidSurface.cu:51-410): a closed triangle mesh for
 * DFSPH_SURFACE_MESH surfaces, kept on the device until the handle is
 * destroyed.  vertices: 3 nv floats; faces: 3 nf vertex indices.  A point is
 * inside when a ray from it in one fixed direction crosses the mesh an odd
 * number of times (MeshPerso::isInside, :224-270; Moeller-Trumbore
 * ray-triangle tests); points outside the mesh's bounding box are outside.
 * The mesh must be closed.  mesh: the id, 1 and up. */
int dfsph_mesh_upload(void* handle, const float* vertices, uint32_t nv, const uint32_t* faces, uint32_t nf,
                      uint32_t* mesh);
/* BufferFluidSurface::isinside on the device for n points (inside[i] = 0 / 1) */
int dfsph_surface_inside(void* handle, const dfsph_surface* s, const float* xyz, uint32_t n, uint8_t* inside);
/* OpenBoundariesSimple::init (OpenBoundariesSimple.cu:117-357): the surfaces and
 * the inflow positions, of which those inside the interior shape or outside
 * the fluid-surface shape are dropped; kept: the remaining positions. */
int dfsph_open_boundary_init(void* handle, const dfsph_surface* boundary, const dfsph_surface* interior,
                             const dfsph_surface* fluid_surface, const float* inflow_xyz, uint32_t n,
                             uint32_t* kept);
/* OpenBoundariesSimple::applyOpenBoundary (OpenBoundariesSimple.cu:541-621):
 * window_displacement (may be NULL): the window's total displacement, the
 * surfaces and inflow positions follow its change since the last call;
 * inflow: inflow_with_predefined_positions_kernel (:360-508) appends, in
 * inflow-position order, a particle at every position whose nearest fluid
 * particle is farther than allowed_new_distance, whose density estimate is at
 * most allowed_new_density and whose interpolated velocity does not point
 * against the boundary's normal (mass of fluid particle 0, kappa = 0, ids
 * after the largest ever used); outflow: outflow_basic_kernel (:517-534)
 * removes the particles outside both the interior and the fluid-surface
 * shape.  The reference's call passes 1.75 r and 700 (DFSPH_CUDA.cpp:3903-3906). */
int dfsph_open_boundary_apply(void* handle, const double* window_displacement, float allowed_new_distance,
                              float allowed_new_density, int use_inflow, int use_outflow, uint32_t* added,
                              uint32_t* removed);

/* ---- rigid bodies -------------------------------------------------------
 * The dynamic bodies of DFSPHCData (vector_dynamic_bodies_data,
 * DFSPH_c_arrays_structure.cpp:897-905) coupled two ways with the fluid:
 * body particles act as moving boundary particles (their velocity enters the
 * density sources, ITER_NEIGHBORS_SOLIDS, DFSPH_core_cuda.cu:318-327, 599-603,
 * 1202-1206) and take the reaction of every pressure update
 * (DFSPH_core_cuda.cu:467-477, 936-946).  The position update then moves the
 * bodies: the external PositionBasedDynamics step of the reference's demo is
 * replaced by a semi-implicit Euler step on the device.  The boundary set must
 * be uploaded first; bodies stay in the solver until it is destroyed. */
#define DFSPH_BODY_DYNAMIC 0     /* gravity + fluid force and torque              */
#define DFSPH_BODY_KINEMATIC 1   /* moved only by dfsph_set_rigid_body             */
#define DFSPH_BODY_SINE 2        /* translation position + amplitude sin(2 pi t / period) axis */
typedef struct dfsph_rigid_body {
    int32_t type;                /* DFSPH_BODY_*                                    */
    double mass;                 /* kg (DYNAMIC)                                    */
    double inertia[3];           /* principal moments in the body frame, kg m^2     */
    double position[3];          /* center of mass (SINE: center at rest), world     */
    double velocity[3];
    double rotation[4];          /* unit quaternion x, y, z, w: body frame -> world  */
    double angular_velocity[3];  /* world frame                                     */
    double axis[3];              /* SINE: unit direction                            */
    double amplitude, period;    /* SINE                                            */
} dfsph_rigid_body;
/* Add a body from its particles in the body frame (relative to the center of
 * mass).  psi == NULL: psi of each particle as FluidModel::computeBoundaryPsi
 * (FluidModel.cpp:233-258) over the walls, the bodies already added and this
 * one, at their current places.  id: the body's index (0, 1, ...). */
int dfsph_add_rigid_body(void* handle, const dfsph_rigid_body* body, const float* xyz_body, const float* psi,
                         uint32_t n, uint32_t* id);
/* overwrite a body's state and motion (UnifiedParticleSet::updateDynamicBodiesParticles,
 * DFSPH_c_arrays_structure.cpp:295-303); the particles follow at the next search */
int dfsph_set_rigid_body(void* handle, uint32_t id, const dfsph_rigid_body* body);
/* the body's state, and the fluid force and torque (about the center of mass)
 * of the last step (compute_fluid_impact_on_dynamic_body_cuda,
 * DFSPH_memory_management_cuda.cu:688-723); force / torque may be NULL */
int dfsph_get_rigid_body(void* handle, uint32_t id, dfsph_rigid_body* body, double* force, double* torque);
/* per-particle fluid forces of the last step, in the order the particles were
 * given (read_rigid_body_force_cuda, DFSPH_memory_management_cuda.cu:676-684):
 * 3 n floats */
int dfsph_rigid_body_forces(void* handle, uint32_t id, float* f_xyz, uint32_t count);
int dfsph_rigid_body_count(void* handle, uint32_t* n);

/* ---- learned warm start -------------------------------------------------
 * No reference counterpart: these entry points add a per-particle MLP
 * (12 -> 64 -> 64 -> 64 -> 2, SiLU; BF16 weights on the matrix cores, FP32
 * accumulate) that predicts the warm starts the reference forms at
 * DFSPH_core_cuda.cu:240 (kappaV) and :3063-3068 (kappa).  The network's two
 * outputs are corrections to those reference warm starts in dt-invariant units
 * (kappaV h, kappa h_f^2): the increments the plain solver's Jacobi
 * iterations still add after the reference warm start.  Features 4 and 5 are
 * the last solves' increments in the same units (kappa h_f^2, kappaV h).  A
 * sweep applies ref + y / (h or h_f^2), bounded between min(ref, prev) and
 * max(ref, 0), where ref is the reference warm start and prev the particle's
 * converged total of the last step (density solver) or half of it
 * (divergence solver), both in raw units and clamped at -0.5: no more
 * pressure than the larger of the two, and a network that outputs zero gives
 * the reference warm start exactly.
 * Parameters use one flat FP32 layout of DFSPH_MLP_NPARAM values:
 *   W1[64x12] b1[64] W2[64x64] b2[64] W3[64x64] b3[64] W4[2x64] b4[2]
 * (row-major, out x in); norm holds in_mean[12] in_std[12] out_mean[2] out_std[2]. */
#define DFSPH_MLP_NPARAM 9282
#define DFSPH_MLP_NNORM 28
#define DFSPH_MLP_NFEAT 12
/* set / read the parameters (norm == NULL: identity normalisation) */
int dfsph_mlp_set_params(void* handle, const float* params, const float* norm);
int dfsph_mlp_get_params(void* handle, float* params, float* norm);
/* y[n][2] = MLP(x[n][12]) on the GPU (normalise: apply the stored normalisation) */
int dfsph_mlp_forward(void* handle, const float* x, uint32_t n, float* y, int normalise);
/* No reference counterpart.  dfsph_mlp_forward without any device allocation,
 * for callers that cannot allocate device memory (code around HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): the
 * features and outputs are staged in a static device array of
 * DFSPH_MLP_NOALLOC_MAX samples (3.7 MB per device, held until the process
 * ends); n above that returns DFSPH_ERR_ARG.  Needs parameters set first
 * (dfsph_mlp_set_params), else DFSPH_ERR_NO_MODEL.  Every call on a device
 * shares that array: calls that overlap on one device (from several threads
 * or handles) race. */
#define DFSPH_MLP_NOALLOC_MAX 65536u
int dfsph_mlp_forward_noalloc(void* handle, const float* x, uint32_t n, float* y, int normalise);
/* loss = sum 0.5 |MLP(x) - t|^2 / n and its gradient (DFSPH_MLP_NPARAM floats),
 * computed by the training kernels (forward + backward on the matrix cores) */
int dfsph_mlp_loss_grad(void* handle, const float* x, const float* t, uint32_t n, float* grad, double* loss,
                        int normalise);
/* Training set harvest: while capacity is not reached, every step (dfsph_step or
 * the phase calls) with the reference warm start appends each stride-th
 * particle's 12 features (computed after the neighbor search) and the
 * network's targets: the increments of the step's two solves (kappaV h,
 * kappa h_f^2).  A pressure solve in another warm-start mode while harvesting
 * returns DFSPH_ERR_STATE.  capacity 0 stops harvesting. */
int dfsph_mlp_harvest(void* handle, uint32_t capacity, uint32_t stride);
int dfsph_mlp_dataset_size(void* handle, uint32_t* n);
int dfsph_mlp_dataset_download(void* handle, float* x, float* t, uint32_t n);
/* replace the training set (x[n][12], t[n][2]) */
int dfsph_mlp_dataset_upload(void* handle, const float* x, const float* t, uint32_t n);
/* normalisation = mean / std of the training set's features and targets */
int dfsph_mlp_fit_normalization(void* handle);
/* mini-batch Adam (batch a multiple of 64) on the GPU's training kernels */
int dfsph_mlp_train(void* handle, uint32_t steps, uint32_t batch, float lr, uint32_t seed, float* loss_first,
                    float* loss_last);
/* features and predictions of the current step (device order, n particles);
 * the predictions are the network's corrections (kappaV h, kappa h_f^2) that
 * the sweeps add to the reference warm starts */
int dfsph_mlp_step_features(void* handle, float* feat, float* pred_v, float* pred_p, uint32_t n);

#ifdef __cplusplus
}
#endif
#endif
