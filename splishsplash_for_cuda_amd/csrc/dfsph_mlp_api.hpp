// dfsph_mlp_api.hpp -- internal interface between the solver handle and the
// learned warm start (dfsph_mlp.hip).  No reference counterpart.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dfsph_device.hpp"

namespace dfsph {

struct MlpState;

// Solver-handle access used by dfsph_mlp.hip.
struct MlpView {
    hipStream_t s;
    MlpState** mlp;
};

MlpView handle_view(void* handle);
void handle_set_error(const char* msg);

bool mlp_loaded(const MlpState* m);
bool mlp_harvesting(const MlpState* m);
// per-particle buffers for up to cap particles
int mlp_ensure_particles(MlpState* m, uint32_t cap);
float* mlp_feat(MlpState* m);            // [cap][12] features of the current step
// predictions from the features of the first n particles
int mlp_infer_particles(MlpState* m, hipStream_t s, uint32_t n);
// the sweeps' inputs: corrections and last totals, kappaV h and kappa h_f^2
WsLearned mlp_ws_v(const MlpState* m, float diameter);
WsLearned mlp_ws_p(const MlpState* m, float diameter);
// where the feature sweep copies the last step's totals
float* mlp_prev_p(MlpState* m);
float* mlp_prev_v(MlpState* m);
// append (features, increments of the step's solves) to the training set
int mlp_harvest_step(MlpState* m, hipStream_t s, const DevTime* t, uint32_t n, const float* kapv, const float* kap,
                     float diameter);
void mlp_destroy(MlpState* m);

}  // namespace dfsph
