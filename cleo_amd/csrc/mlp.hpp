// mlp.hpp -- learned warm start for the condensation Newton solve (mlp.hip).
#pragma once
#include "ctx.hpp"
#include "mlp_device.hpp"

namespace cleo {
hipError_t mlp_create(cleo_ctx* c);
void mlp_destroy(cleo_ctx* c);
// computes the learned guess z = r_new^2 for every active drop into a device array
int mlp_enqueue_guess(cleo_ctx* c, const cleo_cond_params& p, const double** warm);
int mlp_check_ready(cleo_ctx* c);
// device weights of the context's network (for kernels that evaluate the guess themselves)
MLPWeights mlp_weights(cleo_ctx* c);
// device array of one learned guess per drop (capacity maxn)
double* mlp_warm_buffer(cleo_ctx* c);
}  // namespace cleo
