// dfsph_mlp.hip -- learned warm start (placeholder until the MLP kernels land).
#include "dfsph_mlp_api.hpp"
#include "../../include/dfsph_hip.h"

namespace dfsph {

struct MlpState {
    int dummy;
};

int mlp_predict_divergence(MlpState*, hipStream_t, const SimConst&, const DevTime*, uint32_t, const float4*,
                           const float4*, const uint32_t*, const uint32_t*, const float4*, const float*,
                           const float*, const float**)
{
    handle_set_error("learned warm start not available");
    return DFSPH_ERR_NO_MODEL;
}

int mlp_predict_pressure(MlpState*, hipStream_t, const float**)
{
    handle_set_error("learned warm start not available");
    return DFSPH_ERR_NO_MODEL;
}

void mlp_destroy(MlpState* m) { delete m; }

}  // namespace dfsph
