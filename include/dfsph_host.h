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
    double err_v, err;
    double h, time;
    uint32_t n_fluid;
    int32_t removed;
} dfsph_cuda_stats;

/* new SPH::DFSPHCUDA(params, boundary, fluid); NULL on failure (message in err) */
void* dfsph_cuda_new(const dfsph_params* p, const float* boundary_xyz, uint32_t n_boundary,
                     const float* fluid_xyz, const float* fluid_vel, uint32_t n_fluid,
                     char* err, uint32_t err_len);
void dfsph_cuda_delete(void* solver);
/* SPH::DFSPHCUDA::step(); returns a DFSPH_* status */
int dfsph_cuda_step(void* solver, dfsph_cuda_stats* stats);
int dfsph_cuda_reset(void* solver);
void dfsph_cuda_set_divergence(void* solver, int enable);
/* the libdfsph_hip.so handle owned by the solver (for dfsph_download etc.) */
void* dfsph_cuda_handle(void* solver);
int dfsph_cuda_save(void* solver, const char* path);
int dfsph_cuda_load(void* solver, const char* path);
const char* dfsph_cuda_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
