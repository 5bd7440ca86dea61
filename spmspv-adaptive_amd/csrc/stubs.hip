// Entry points whose kernels are not built yet: they fail loudly with
// SPMSPV_NOT_IMPLEMENTED instead of falling back to anything else.
#include "handle.hpp"

using namespace spmspv;

extern "C" {

int spmspv_run_batched(spmspv_handle_t h, int, const int*, const int*, const float*,
                       const long long*, int*, float*, int*) {
  return set_error(h, SPMSPV_NOT_IMPLEMENTED, "run_batched is not built yet");
}

}  // extern "C"

extern "C" {
int spmspv_bfs(spmspv_handle_t h, int, int*, int, int, double*) {
  return set_error(h, SPMSPV_NOT_IMPLEMENTED, "bfs is not built yet");
}
int spmspv_bfs_release(spmspv_handle_t) { return SPMSPV_SUCCESS; }
int spmspv_mlp_train(const float*, const float*, int, int, int, float, unsigned, float*, float*) {
  return set_error(nullptr, SPMSPV_NOT_IMPLEMENTED, "mlp training is not built yet");
}
int spmspv_mlp_loss_grad(const float*, const float*, const float*, int, float*, float*) {
  return set_error(nullptr, SPMSPV_NOT_IMPLEMENTED, "mlp gradient is not built yet");
}
int spmspv_mlp_forward(const float*, const float*, int, float*) {
  return set_error(nullptr, SPMSPV_NOT_IMPLEMENTED, "mlp forward is not built yet");
}
}
