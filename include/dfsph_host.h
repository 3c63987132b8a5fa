/*
 * dfsph_host.h -- C entry points of libdfsph_host.so, the C++ host layer
 * (SPH::DFSPHCUDA, splishsplash_for_cuda_amd/csrc/DFSPH_CUDA.h).
 *
 * They let a non-C++ caller drive DFSPHCUDA::step() exactly as the
 * reference's demo loop does (Demos/StaticBoundaryDemo/main.cpp:164-168):
 * create the solver from a boundary and a fluid set, call step() repeatedly,
 * read the state back through the solver's dfsph_hip.h handle.
 */
#ifndef DFSPH_HOST_H
#define DFSPH_HOST_H

#include <stdint.h>

#include "dfsph_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dfsph_cuda_stats {
    uint32_t iters_v, iters;
    double err_v, err;        /* final average source: divergence (rho/s), density (rho) */
    double h, time;           /* step size used by this step, simulated time after it  */
    uint32_t n_fluid;
    int32_t removed;
    uint32_t inflow_added, outflow_removed;   /* open boundaries, at the start of the step */
    uint32_t max_neighbors_seen;  /* largest neighbor list of the step's search */
    uint32_t overflow;            /* particles whose list was truncated; the step then
                                     returns DFSPH_ERR_NEIGHBOR_OVERFLOW after it completed */
} dfsph_cuda_stats;

/* new SPH::DFSPHCUDA(params, boundary, fluid); NULL on failure (message in err) */
void* dfsph_cuda_new(const dfsph_params* p, const float* boundary_xyz, uint32_t n_boundary,
                     const float* fluid_xyz, const float* fluid_vel, uint32_t n_fluid,
                     char* err, uint32_t err_len);
void dfsph_cuda_delete(void* solver);
/* SPH::DFSPHCUDA::step(); returns a DFSPH_* status */
int dfsph_cuda_step(void* solver, dfsph_cuda_stats* stats);
int dfsph_cuda_reset(void* solver);
/* DFSPHCUDA::applyOpenBoundaries (DFSPH_CUDA.cpp:3889-3924): OpenBoundariesSimple::init
 * now (dfsph_open_boundary_init), then every dfsph_cuda_step starts with
 * dfsph_open_boundary_apply(allowed_new_distance (0: 1.75 r), allowed_new_density) */
int dfsph_cuda_set_open_boundaries(void* solver, const dfsph_surface* boundary, const dfsph_surface* interior,
                                   const dfsph_surface* fluid_surface, const float* inflow_xyz, uint32_t n,
                                   float allowed_new_distance, float allowed_new_density, int use_inflow,
                                   int use_outflow);
/* DFSPHCUDA::applyDynamicWindow (DFSPH_CUDA.cpp:3988-4139) over
 * DynamicWindowInterface (DynamicWindow.cu:150-237): every later
 * dfsph_cuda_step starts by moving the simulated region (`simulation`: a
 * cuboid, cylinder or sphere at zero displacement) toward rigid body
 * `tracked_body` in x and z once it is more than `trigger` away (0: 2.5 h),
 * by at most `max_displacement` (0: 3.5 h): the region's walls (kinematic
 * body `wall_body`, -1: none) move with it, the existing fluid is kept
 * inside the region less 3 r of before the move and more than 1.5 h inside
 * the moved region (the rest is removed), the background buffer (n points at
 * rest, moving with the region) is loaded into the band of width `band`
 * (0: max_displacement + 1.5 h) inside the moved region less 3 r (x and z
 * faces of a cuboid, the radius of a cylinder or sphere) and under `fluid`
 * where no fluid is within
 * 1.75 r and the density is at most 0.7 rho0 (dfsph_load_fluid, the inflow's
 * values), and the loaded particles are settled by 10
 * density-solver iterations of 3 ms restricted to them and their neighbors,
 * with a velocity damping of 0.8 -> 0.2 (DynamicWindow::stabilizeFluid; they
 * leave at rest, the warm starts are cleared, the time and the rigid bodies
 * are restored afterwards); an open boundary set with
 * dfsph_cuda_set_open_boundaries follows the region. */
int dfsph_cuda_set_dynamic_window(void* solver, const dfsph_surface* simulation, const dfsph_surface* fluid,
                                  const float* background, uint32_t n, int tracked_body, int wall_body,
                                  double trigger, double max_displacement, double band);
/* RestFLuidLoaderInterface::initializeFluidToSurface + stabilizeFluid
 * (RestFluidLoader.h:31-309, RestFluidLoader.cu:127-207): the solver's fluid
 * becomes the n background points (a fluid at rest, e.g. a lattice of spacing
 * 2 r 0.8^(1/3), larger than the volume) inside `fluid`, less those the density
 * extraction passes remove, the background points outside `fluid` within one
 * kernel radius serving as air particles (dfsph_remove_density_peaks at a limit stepping
 * from density_start down to density_end by step_density; 0: 1900, 1000 and
 * 50), settled by stab_iters damped solver iterations of 3 ms (0: 10) and left
 * at rest (velocities and warm starts zero, the time and rigid bodies
 * restored).  loaded: the particle count; passes: extraction passes run;
 * avg_density: the mean density after the last pass. */
int dfsph_cuda_load_rest_fluid(void* solver, const dfsph_surface* fluid, const float* background, uint32_t n,
                               float density_start, float density_end, float step_density, uint32_t stab_iters,
                               uint32_t* loaded, uint32_t* passes, double* avg_density);
/* the window's total displacement (3 doubles), its moves so far, and the
 * particles loaded and removed by the last move */
int dfsph_cuda_dynamic_window_state(void* solver, double* displacement, uint32_t* moves, uint32_t* loaded,
                                    uint32_t* removed);
void dfsph_cuda_set_divergence(void* solver, int enable);
/* the libdfsph_hip.so handle owned by the solver (for dfsph_download etc.) */
void* dfsph_cuda_handle(void* solver);
int dfsph_cuda_save(void* solver, const char* path);
int dfsph_cuda_load(void* solver, const char* path);
const char* dfsph_cuda_last_error(void);

/* DFSPHCUDA::handleSimulationSave(save_liquid, false, save_boundaries)
 * (DFSPH_CUDA.cpp:1887-1919): folder/fluid_file.txt and general_data.txt (h),
 * folder/boundaries_file.txt, as UnifiedParticleSet::write_to_file
 * (DFSPH_c_arrays_structure.cpp:424-478) writes them. */
int dfsph_cuda_save_text(void* solver, const char* folder, int save_liquid, int save_boundaries);
/* DFSPHCUDA::handleSimulationLoad(true, load_liquid_velocities, false, false,
 * false, false) (DFSPH_CUDA.cpp:1921-1979): replace the fluid with
 * folder/fluid_file.txt, set h = h_past = h_future from general_data.txt. */
int dfsph_cuda_load_text(void* solver, const char* folder, int load_liquid_velocities);
/* a solver whose boundary set is folder/boundaries_file.txt (mass column = psi,
 * read_boundaries_from_file, DFSPH_c_arrays_structure.cpp:1294-1323) and whose
 * fluid and h come from the fluid files; NULL on failure (dfsph_cuda_last_error) */
void* dfsph_cuda_new_from_text(const dfsph_params* p, const char* folder, int load_liquid_velocities);

/* SceneLoader::readScene (Utilities/SceneLoader.cpp:13-300), the fluid of
 * DemoBase::createFluidBlocks (Demos/Common/DemoBase.cpp:539-600) and the
 * boundary particles of the RigidBodies (OBJ meshes, paths relative to
 * base_path; NULL: the scene file's folder).  NULL on failure. */
void* dfsph_scene_read(const char* path, const char* base_path);
void dfsph_scene_free(void* scene);
/* library defaults overridden by the scene's Configuration (DemoBase.cpp:440-455) */
void dfsph_scene_params(const void* scene, dfsph_params* p);
uint32_t dfsph_scene_fluid_count(const void* scene);
void dfsph_scene_fluid(const void* scene, float* xyz, float* vel);   /* 3 n floats each */
uint32_t dfsph_scene_boundary_count(const void* scene);
void dfsph_scene_boundary(const void* scene, float* xyz);
/* the scene's dynamic rigid bodies ("isDynamic"): mass and principal
 * moments of the closed mesh at the body's "density", center of mass and
 * principal axes as position and rotation; returns the particle count and
 * fills xyz_body (3 n floats, body frame) when not NULL */
uint32_t dfsph_scene_body_count(const void* scene);
uint32_t dfsph_scene_body(const void* scene, uint32_t i, dfsph_rigid_body* body, float* xyz_body);
/* the solver of the scene with its dynamic bodies (p == NULL: dfsph_scene_params); NULL on failure */
void* dfsph_cuda_new_from_scene(const void* scene, const dfsph_params* p);

/* UnifiedParticleSet::write_to_file / load_from_file (DFSPH_c_arrays_structure.cpp:
 * 424-560) for one particle set; vel == NULL writes zeros.  The reader returns
 * the particle count (arrays filled only when it is <= capacity), -1 on error. */
int dfsph_particles_write_text(const char* path, const float* xyz, const float* vel, const float* mass,
                               uint32_t n, int has_factor_computation, int velocity_impacted_by_fluid_solver);
int64_t dfsph_particles_read_text(const char* path, float* xyz, float* vel, float* mass, uint32_t capacity);

#ifdef __cplusplus
}
#endif
#endif
