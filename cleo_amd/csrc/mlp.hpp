// mlp.hpp -- learned warm start for the condensation Newton solve (mlp.hip).
#pragma once
#include "ctx.hpp"

namespace cleo {
hipError_t mlp_create(cleo_ctx* c);
void mlp_destroy(cleo_ctx* c);
// computes the learned guess z = r_new^2 for every active drop into a device array
int mlp_enqueue_guess(cleo_ctx* c, const cleo_cond_params& p, const double** warm);
int mlp_check_ready(cleo_ctx* c);
}  // namespace cleo
