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

