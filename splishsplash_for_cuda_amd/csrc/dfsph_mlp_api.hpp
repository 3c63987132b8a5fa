// dfsph_mlp_api.hpp -- internal interface between the solver handle and the
// learned warm start (dfsph_mlp.hip).  No reference counterpart.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dfsph_device.hpp"

namespace dfsph {

struct MlpState;

// Views of the solver buffers the learned warm start reads (current order).
struct MlpView {
    hipStream_t s;
    SimConst c;
    const DevTime* tm;
    uint32_t n;
    const float4* pos;
    const float4* bpos;
    const uint32_t* nbr;
    const uint32_t* cnt;
    const float4* vel;
    const float* rho;
    const float* alpha;
    const float* radv;
    float* kappa;
    float* kappaV;
    MlpState** mlp;
};

MlpView handle_view(void* handle);
void handle_set_error(const char* msg);

// Predict the divergence warm start for every particle (features from the
// current state, then the MLP); *kws_out receives the device array of
// predicted kappaV values (already in the units the warm-start sweep uses).
int mlp_predict_divergence(MlpState* m, hipStream_t s, const SimConst& c, const DevTime* tm, uint32_t n,
                           const float4* pos, const float4* bpos, const uint32_t* nbr, const uint32_t* cnt,
                           const float4* vel, const float* kappa, const float* kappaV, const float** kws_out);
// The pressure warm start predicted by the same pass.
int mlp_predict_pressure(MlpState* m, hipStream_t s, const float** kws_out);
void mlp_destroy(MlpState* m);

}  // namespace dfsph
