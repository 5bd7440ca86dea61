// Batched multi-vector SpMSpV: nvec sparse vectors against one matrix with
// two host synchronisations for the whole batch instead of two per call.
//   1. one selector pass: k_prep with the selector runs for every vector
//      back to back and stashes {bin_len, choice} per vector; one readback;
//   2. every vector's chosen pipeline is queued back to back with its
//      bin_len already known (no readback inside), results at the caller's
//      offsets, each vector's y_nnz copied to d_ynnz[v] on the device;
//   3. one final synchronisation.
// No reference counterpart (the reference runs one vector per call).
#include <vector>

#include "handle.hpp"

using namespace spmspv;

namespace {

__global__ void k_copy_ynnz(const Ctl* ctl, int* dst) { *dst = ctl->ynnz; }

int ensure_batch_stash(spmspv_handle_t h, int nvec) {
  if (h->batch_cap >= nvec) return SPMSPV_SUCCESS;
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  if (h->batch_stash) HIP_TRY(h, hipFree(h->batch_stash));
  h->batch_stash = nullptr;
  h->batch_cap = 0;
  HIP_TRY(h, hipMalloc(&h->batch_stash, sizeof(long long) * (size_t)nvec));
  h->batch_cap = nvec;
  return SPMSPV_SUCCESS;
}

}  // namespace

extern "C" {

int spmspv_run_batched(spmspv_handle_t h, int nvec, const int* d_vptr, const int* d_vkey,
                       const float* d_vval, const long long* d_yoff, int* d_ykey, float* d_yval,
                       int* d_ynnz, int* kernel_used) {
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  if (nvec < 0 || (nvec > 0 && (!d_vptr || !d_yoff || !d_ykey || !d_yval || !d_ynnz)))
    return set_error(h, SPMSPV_INVALID_ARGUMENT, "run_batched: bad arguments");
  if (nvec == 0) return SPMSPV_SUCCESS;
  int rc;
  if ((rc = ensure_arena(h)) || (rc = need_csc(h)) || (rc = ensure_mat_features(h))) return rc;
  if ((rc = ensure_batch_stash(h, nvec))) return rc;
  std::vector<int> vptr(nvec + 1);
  std::vector<long long> yoff(nvec + 1);
  HIP_TRY(h, hipMemcpyAsync(vptr.data(), d_vptr, sizeof(int) * (nvec + 1), hipMemcpyDeviceToHost,
                            h->stream));
  HIP_TRY(h, hipMemcpyAsync(yoff.data(), d_yoff, sizeof(long long) * (nvec + 1),
                            hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  for (int v = 0; v < nvec; ++v) {
    const int xn = vptr[v + 1] - vptr[v];
    if (vptr[v] < 0 || xn < 0 || xn > h->n)
      return set_error(h, SPMSPV_INVALID_ARGUMENT, "run_batched: vector offsets out of range");
    if (xn > 0 && (!d_vkey || !d_vval))
      return set_error(h, SPMSPV_INVALID_ARGUMENT, "run_batched: null vector arrays");
    if (yoff[v] < 0 || yoff[v + 1] < yoff[v])
      return set_error(h, SPMSPV_INVALID_ARGUMENT, "run_batched: output offsets not ascending");
  }
  // 1. selector pass
  for (int v = 0; v < nvec; ++v) {
    const int xn = vptr[v + 1] - vptr[v];
    if ((rc = spmspv_set_sparse_x(h, const_cast<int*>(d_vkey) + vptr[v],
                                  const_cast<float*>(d_vval) + vptr[v], xn)))
      return rc;
    SelectorArgs sel{h->have_weights ? h->weights_dev : nullptr, h->mat, h->sel_mode, h->sel_fixed,
                     h->batch_stash + v, 0, -1, 0.f, h->bucket_lo, h->bucket_hi};
    // features only: the chosen pipeline runs the column scan itself when
    // its slot needs one (the scan buffers hold one vector at a time)
    launch_select_feat(h->stream, h->csc_colptr, h->x_key, h->x_val, h->n, h->xnnz, h->sel_part, h->ctl,
                       sel, nullptr);
    h->launches++;
    HIP_TRY(h, hipGetLastError());
  }
  std::vector<long long> st((size_t)nvec);
  HIP_TRY(h, hipMemcpyAsync(st.data(), h->batch_stash, sizeof(long long) * nvec,
                            hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  for (int v = 0; v < nvec; ++v) {
    const long long need = std::min<long long>(h->m, sel_word_binlen(st[v]));
    if (yoff[v + 1] - yoff[v] < need)
      return set_error(h, SPMSPV_INVALID_ARGUMENT,
                       "run_batched: output capacity of vector " + std::to_string(v) + " is " +
                           std::to_string(yoff[v + 1] - yoff[v]) + ", needs " + std::to_string(need));
    const int kid = sel_word_choice(st[v]);
    if (kid < 0 || kid > SPMSPV_K_COL_BUCKET)
      return set_error(h, SPMSPV_DEVICE_ERROR, "run_batched: selector returned no slot");
  }
  // 2. queued pipelines
  for (int v = 0; v < nvec; ++v) {
    const int xn = vptr[v + 1] - vptr[v];
    const int kid = sel_word_choice(st[v]);
    if (kernel_used) kernel_used[v] = kid;
    if ((rc = spmspv_set_sparse_x(h, const_cast<int*>(d_vkey) + vptr[v],
                                  const_cast<float*>(d_vval) + vptr[v], xn)))
      return rc;
    h->binlen_host = sel_word_binlen(st[v]);
    if ((rc = run_kernel(h, kid, nullptr, d_ykey + yoff[v], d_yval + yoff[v], true))) return rc;
    hipLaunchKernelGGL(k_copy_ynnz, dim3(1), dim3(1), 0, h->stream, h->ctl, d_ynnz + v);
  }
  // 3. the batch is complete when this returns; x is left unset
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  h->sparse_valid = false;
  h->dense_valid = false;
  h->bits_valid = false;
  h->enc_valid = h->enc5_valid = false;
  h->prep_valid = false;
  h->binlen_host = -1;
  h->vec_type = -1;
  h->xnnz = 0;
  return SPMSPV_SUCCESS;
}

}  // extern "C"
