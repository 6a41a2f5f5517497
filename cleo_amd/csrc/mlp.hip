// mlp.hip -- learned warm start for the condensation Newton solve.
#include <hip/hip_runtime.h>

#include "mlp.hpp"

namespace cleo {
struct MLPDev {
  bool ready = false;
};
hipError_t mlp_create(cleo_ctx* c) {
  c->mlp = new MLPDev();
  return hipSuccess;
}
void mlp_destroy(cleo_ctx* c) {
  delete c->mlp;
  c->mlp = nullptr;
}
int mlp_check_ready(cleo_ctx* c) {
  if (!c->mlp || !c->mlp->ready) {
    c->msg = "MLP weights not set";
    return CLEO_E_STATE;
  }
  return CLEO_OK;
}
int mlp_enqueue_guess(cleo_ctx* c, const cleo_cond_params&, const double** warm) {
  *warm = nullptr;
  return mlp_check_ready(c);
}
}  // namespace cleo
