// The reference's adaptive-framework flow (la/test/new_test_app.cu:133-501)
// written against the C++ handle: CSR in, ToCSC, allocPreBuffer, then per
// sparse x the learned selector picks a slot and the result is checked
// against the reference's own known-answer pattern (all values 1, x = 1 on
// [0, k): y[i] counts the entries of row i in columns < k,
// la/test/generate_model_data.cu:639-643, 942-944).
// Build: make -C spmspv-adaptive_amd examples (links lib/libspmspv_amd.so)

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "spmspv_handle.hpp"

int main(int argc, char** argv) {
  const int scale = argc > 1 ? std::atoi(argv[1]) : 14;
  long long nnz = 0;
  if (spmspv_gen_rmat_csr(scale, 16, 0.57, 0.19, 0.19, 1, &nnz, nullptr, nullptr, nullptr)) return 1;
  const int n = 1 << scale;
  std::vector<int> rp(n + 1), col(nnz);
  std::vector<float> val(nnz, 1.0f);
  spmspv_gen_rmat_csr(scale, 16, 0.57, 0.19, 0.19, 1, &nnz, rp.data(), col.data(), val.data());
  for (auto& v : val) v = 1.0f;  // the reference forces every value to 1 (new_test_app.cu:572-574)
  int *d_rp, *d_col, *d_xk, *d_yk;
  float *d_val, *d_xv, *d_yv;
  spmspv_malloc((void**)&d_rp, sizeof(int) * (n + 1));
  spmspv_malloc((void**)&d_col, sizeof(int) * nnz);
  spmspv_malloc((void**)&d_val, sizeof(float) * nnz);
  spmspv_malloc((void**)&d_xk, sizeof(int) * n);
  spmspv_malloc((void**)&d_xv, sizeof(float) * n);
  spmspv_malloc((void**)&d_yk, sizeof(int) * n);
  spmspv_malloc((void**)&d_yv, sizeof(float) * n);
  spmspv_memcpy_h2d(d_rp, rp.data(), sizeof(int) * (n + 1));
  spmspv_memcpy_h2d(d_col, col.data(), sizeof(int) * nnz);
  spmspv_memcpy_h2d(d_val, val.data(), sizeof(float) * nnz);
  SpmspvHandle<int, unsigned int, float> A(n, n, (int)nnz);
  A.InputCSR((int)nnz, d_rp, d_col, d_val);
  A.holaPreprocess();
  A.ToCSC();
  A.allocPreBuffer();
  int bad = 0;
  for (int k : {1, 10, 100, 1000, n / 4, n}) {
    std::vector<int> key(k);
    std::vector<float> one(k, 1.0f);
    for (int i = 0; i < k; ++i) key[i] = i;
    spmspv_memcpy_h2d(d_xk, key.data(), sizeof(int) * k);
    spmspv_memcpy_h2d(d_xv, one.data(), sizeof(float) * k);
    A.set_sparse_x(d_xk, d_xv, k);
    int ynnz = -1, used = -1;
    if (A.run_auto(nullptr, d_yk, d_yv, &ynnz, &used)) {
      std::printf("error: %s\n", A.last_error());
      return 1;
    }
    std::vector<int> yk(ynnz);
    std::vector<float> yv(ynnz);
    spmspv_memcpy_d2h(yk.data(), d_yk, sizeof(int) * ynnz);
    spmspv_memcpy_d2h(yv.data(), d_yv, sizeof(float) * ynnz);
    // expected: count of entries < k per row (exact in FP32)
    int j = 0;
    for (int r = 0; r < n; ++r) {
      int c = 0;
      for (int e = rp[r]; e < rp[r + 1]; ++e) c += col[e] < k;
      if (!c) continue;
      if (j >= ynnz || yk[j] != r || yv[j] != (float)c) ++bad;
      ++j;
    }
    if (j != ynnz) ++bad;
    std::printf("x_nnz %d -> slot %d, y_nnz %d, %s\n", k, used, ynnz, bad ? "MISMATCH" : "RESULT OK");
  }
  return bad ? 1 : 0;
}
