// PageRank-style driver: residual-push PageRank whose every iteration is
// one selector-dispatched SpMSpV. The reference has no such driver; its
// harnesses replay frontiers recorded from a PageRank run, one sparse x
// per line of pr_x_<matrix>.txt (la/test/test_application.cu:206-210,
// 252-256; la/test/new_test_app.cu:293-299). This produces those frontiers
// on the device and runs them.
//
// Transition matrix: column u of the CSC holds u's out-edges with weights
// a_vu; the walk moves from u to v with probability a_vu / w_u, where
// w_u = sum_v a_vu (unit weights give the usual PageRank; columns with no
// entries keep their mass). With teleport (1 - d) / n the ranks are the
// solution of (I - d A W^-1) p = (1 - d) / n.
//
// Residual push: p = 0, r = (1 - d) / n everywhere, every vertex active.
// Per iteration, for the active set S (sorted keys):
//   take:  p[u] += r[u]; x_u = r[u] / w_u; r[u] = 0          (u in S)
//   y = A x                    (the selector's slot, sorted output)
//   push:  r[v] += d * y_v; v joins the next S if r[v] > eps * (1 - d) / n
// (only vertices that received a push can cross the threshold: every
// other vertex above it was in S and is now at 0). At the end p += r.
// With every residual at most eps (1 - d) / n, sum|p - p*| <= eps * d.
#include <chrono>

#include "handle.hpp"

using namespace spmspv;

struct PrState {
  void* alloc = nullptr;
  float* w = nullptr;         // n: column weight sums
  float* r = nullptr;         // n: residuals
  int* fk[2] = {nullptr, nullptr};  // n: active keys (sorted)
  float* fv[2] = {nullptr, nullptr};  // n: scratch values of the compaction
  float* xv = nullptr;        // n: x values of the active set
  int* yk = nullptr;          // m
  float* yv = nullptr;        // m
  unsigned* bits = nullptr;   // ceil(n/32)+2: vertices crossing the threshold
};

namespace {

__global__ void k_col_weights(const int* __restrict__ colptr, const float* __restrict__ cval, int n,
                              float* __restrict__ w) {
  // one wave per column, lanes stride its entries (hub columns stay wide)
  const int wid = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int nw = gridDim.x * blockDim.x / kWave;
  for (int c = wid; c < n; c += nw) {
    const int b = colptr[c], e = colptr[c + 1];
    float s = 0.f;
    for (int i = b + lane(); i < e; i += kWave) s += cval[i];
    s = wave_sum(s);
    if (lane() == 0) w[c] = s;
  }
}

__global__ void k_pr_init(int n, float r0, float* __restrict__ p, float* __restrict__ r,
                          int* __restrict__ fk) {
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    p[v] = 0.f;
    r[v] = r0;
    fk[v] = v;
  }
}

__global__ void k_pr_take(const int* __restrict__ fk, int cnt, const float* __restrict__ w,
                          float* __restrict__ p, float* __restrict__ r, float* __restrict__ xv) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += gridDim.x * blockDim.x) {
    const int u = fk[i];
    const float ru = r[u];
    p[u] += ru;
    r[u] = 0.f;
    const float wu = w[u];
    xv[i] = wu > 0.f ? ru / wu : 0.f;
  }
}

__global__ void k_pr_push(const int* __restrict__ yk, const float* __restrict__ yv,
                          const Ctl* __restrict__ ctl, float d, float thr, float* __restrict__ r,
                          unsigned* __restrict__ bits) {
  const int cnt = ctl->ynnz;  // y's keys are distinct: one writer per r[v]
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += gridDim.x * blockDim.x) {
    const int v = yk[i];
    const float nr = r[v] + d * yv[i];
    r[v] = nr;
    if (nr > thr) atomicOr(&bits[v >> 5], 1u << (v & 31));
  }
}

__global__ void k_pr_finish(int n, float* __restrict__ p, const float* __restrict__ r) {
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) p[v] += r[v];
}

int grid_of(long long items) {
  return (int)std::max<long long>(1, std::min<long long>(4096, (items + 255) / 256));
}

}  // namespace

namespace spmspv {
// PageRank buffers, sized by m and n: allocated once per handle.
int pr_alloc(spmspv_handle_t h) {
  if (h->pr) return SPMSPV_SUCCESS;
  auto* st = new PrState();
  const size_t n = h->n, m = h->m, words = (n + 31) / 32 + 2;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off = (off + bytes + 255) / 256 * 256;
    return o;
  };
  const size_t ow = take(n * 4), orr = take(n * 4), of0 = take(n * 4), of1 = take(n * 4);
  const size_t ov0 = take(n * 4), ov1 = take(n * 4), oxv = take(n * 4);
  const size_t oyk = take(m * 4), oyv = take(m * 4), ob = take(words * 4);
  if (hipMalloc(&st->alloc, off) != hipSuccess) {
    delete st;
    return set_error(h, SPMSPV_OUT_OF_MEMORY, "pagerank buffers");
  }
  char* p = static_cast<char*>(st->alloc);
  st->w = reinterpret_cast<float*>(p + ow);
  st->r = reinterpret_cast<float*>(p + orr);
  st->fk[0] = reinterpret_cast<int*>(p + of0);
  st->fk[1] = reinterpret_cast<int*>(p + of1);
  st->fv[0] = reinterpret_cast<float*>(p + ov0);
  st->fv[1] = reinterpret_cast<float*>(p + ov1);
  st->xv = reinterpret_cast<float*>(p + oxv);
  st->yk = reinterpret_cast<int*>(p + oyk);
  st->yv = reinterpret_cast<float*>(p + oyv);
  st->bits = reinterpret_cast<unsigned*>(p + ob);
  launch_fill_i32(h->stream, reinterpret_cast<int*>(st->bits), (long long)words, 0);
  if (hipStreamSynchronize(h->stream) != hipSuccess) {
    (void)hipFree(st->alloc);
    delete st;
    return set_error(h, SPMSPV_DEVICE_ERROR, "pagerank init");
  }
  h->pr = st;
  return SPMSPV_SUCCESS;
}

void pr_release(spmspv_handle_t h) {
  if (!h || !h->pr) return;
  if (h->pr->alloc) (void)hipFree(h->pr->alloc);
  delete h->pr;
  h->pr = nullptr;
}
}  // namespace spmspv

extern "C" {

int spmspv_pagerank(spmspv_handle_t h, float damping, float eps, int max_iter, int policy,
                    float* d_rank, double stats[4], int* sizes) {
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  if (!d_rank || !(damping > 0.f && damping < 1.f) || !(eps > 0.f) || max_iter < 1 || policy < -1 ||
      policy >= SPMSPV_NUM_SLOTS)
    return set_error(h, SPMSPV_INVALID_ARGUMENT, "pagerank: bad damping, eps, max_iter, policy or rank");
  if (h->m != h->n) return set_error(h, SPMSPV_INVALID_ARGUMENT, "pagerank needs a square matrix");
  int rc;
  if ((rc = ensure_arena(h)) || (rc = need_csc(h)) || (rc = ensure_mat_features(h)) || (rc = pr_alloc(h)))
    return rc;
  PrState* st = h->pr;
  hipStream_t s = h->stream;
  const int n = h->n;
  const int saved_mode = h->sel_mode, saved_fixed = h->sel_fixed;
  if (policy >= 0) {
    h->sel_mode = 2;
    h->sel_fixed = policy;
  }
  struct Restore {
    spmspv_handle_t h;
    int mode, fixed;
    ~Restore() {
      h->sel_mode = mode;
      h->sel_fixed = fixed;
    }
  } restore{h, saved_mode, saved_fixed};
  const auto t0 = std::chrono::steady_clock::now();
  hipLaunchKernelGGL(k_col_weights, dim3(grid_of((long long)n * kWave)), dim3(256), 0, s, h->csc_colptr,
                     h->csc_val, n, st->w);
  const float r0 = (1.f - damping) / (float)n;
  const float thr = eps * r0;
  hipLaunchKernelGGL(k_pr_init, dim3(grid_of(n)), dim3(256), 0, s, n, r0, d_rank, st->r, st->fk[0]);
  HIP_TRY(h, hipGetLastError());
  int cur = 0, cnt = n, it = 0;
  double pushes = 0.0;
  while (cnt > 0 && it < max_iter) {
    if (sizes) sizes[it] = cnt;
    pushes += cnt;
    hipLaunchKernelGGL(k_pr_take, dim3(grid_of(cnt)), dim3(256), 0, s, st->fk[cur], cnt, st->w, d_rank,
                       st->r, st->xv);
    HIP_TRY(h, hipGetLastError());
    if ((rc = spmspv_set_sparse_x(h, st->fk[cur], st->xv, cnt))) return rc;
    int ynnz = 0, used = -1;
    if ((rc = spmspv_run_auto(h, 1.f, nullptr, st->yk, st->yv, &ynnz, &used))) return rc;
    // the residual update and the threshold test, then the crossing
    // vertices as the next sorted active set (the compaction clears the bits)
    hipLaunchKernelGGL(k_pr_push, dim3(grid_of(ynnz)), dim3(256), 0, s, st->yk, st->yv, h->ctl, damping,
                       thr, st->r, st->bits);
    launch_compact(s, st->bits, nullptr, n, st->r, 0, st->fk[cur ^ 1], st->fv[cur ^ 1], h->scan_state,
                   h->ctl, next_ep(h, 1));
    HIP_TRY(h, hipGetLastError());
    if ((rc = readback_ynnz(h, &cnt))) return rc;
    cur ^= 1;
    ++it;
  }
  hipLaunchKernelGGL(k_pr_finish, dim3(grid_of(n)), dim3(256), 0, s, n, d_rank, st->r);
  HIP_TRY(h, hipGetLastError());
  HIP_TRY(h, hipStreamSynchronize(s));
  // x was the driver's own buffer: leave none set
  h->sparse_valid = h->dense_valid = h->bits_valid = false;
  h->enc_valid = h->enc5_valid = false;
  h->prep_valid = false;
  h->binlen_host = -1;
  h->vec_type = -1;
  h->xnnz = 0;
  if (stats) {
    stats[0] = it;
    stats[1] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    stats[2] = pushes;
    stats[3] = cnt;  // vertices still above the threshold (0 when converged)
  }
  return SPMSPV_SUCCESS;
}

}  // extern "C"
