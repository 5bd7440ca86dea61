/*
 * CPU oracle for the adaptive SpMV/SpMSpV hot path -- TEST INFRASTRUCTURE.
 * Loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg. Each function cites the reference text it restates (oracle.h).
 */
#include "oracle.h"

#include <float.h>
#include <limits.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

void oracle_csr_to_csc(int m, int n, int nnz, const int* row_ptr,
                       const int* col, const float* val, int* col_ptr,
                       int* row_idx, float* cval) {
  (void)nnz;
  memset(col_ptr, 0, sizeof(int) * (size_t)(n + 1));
  for (int i = 0; i < m; ++i)
    for (int k = row_ptr[i]; k < row_ptr[i + 1]; ++k) col_ptr[col[k] + 1]++;
  for (int j = 0; j < n; ++j) col_ptr[j + 1] += col_ptr[j];
  int* next = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  for (int j = 0; j < n; ++j) next[j] = col_ptr[j];
  /* rows visited in ascending order -> rows ascending inside each column */
  for (int i = 0; i < m; ++i)
    for (int k = row_ptr[i]; k < row_ptr[i + 1]; ++k) {
      int p = next[col[k]]++;
      row_idx[p] = i;
      cval[p] = val[k];
    }
  free(next);
}

void oracle_serialspmv_f32(int m, const int* row_ptr, const int* col,
                           const float* val, const float* x_dense, float alpha,
                           float* y) {
  /* test_application.cu:107-116: sum += x[col] * val * alpha, row order */
  for (int i = 0; i < m; ++i) {
    float sum = 0.0f;
    for (int j = row_ptr[i]; j < row_ptr[i + 1]; ++j)
      sum += x_dense[col[j]] * val[j] * alpha;
    y[i] = sum;
  }
}

int oracle_spmspv_csc_host_f32(int m, const int* col_ptr, const int* row_idx,
                               const float* cval, int xnnz, const int* x_key,
                               const float* x_val, float* y_dense, int* y_key,
                               float* y_val) {
  /* test_sort_spmspv.cu:336-351 (x value kept as float) */
  memset(y_dense, 0, sizeof(float) * (size_t)m);
  for (int i = 0; i < xnnz; ++i) {
    int c = x_key[i];
    float xv = x_val[i];
    for (int j = col_ptr[c]; j < col_ptr[c + 1]; ++j)
      y_dense[row_idx[j]] += cval[j] * xv;
  }
  /* :354-372: keep |y| >= FLT_EPSILON, ascending */
  int idx = 0;
  for (int i = 0; i < m; ++i) {
    if (!(fabsf(y_dense[i]) < FLT_EPSILON)) {
      if (y_key) y_key[idx] = i;
      if (y_val) y_val[idx] = y_dense[i];
      idx++;
    }
  }
  return idx;
}

int oracle_spmspv_csc_host_f32_omp(int m, const int* col_ptr, const int* row_idx,
                                   const float* cval, int xnnz, const int* x_key,
                                   const float* x_val, float* y_dense, int* y_key,
                                   float* y_val, int threads) {
  /* the loop of test_sort_spmspv.cu:336-351 with the rows split into one
   * contiguous range per thread: every thread walks all of x's columns and
   * adds only the products of its own rows (found by binary search, rows
   * ascending in a column), in x's order, so each row's sum is the serial
   * one. Then the |y| >= FLT_EPSILON filter of :354-372, in parallel. */
  if (threads < 1) threads = 1;
  int* cnt = (int*)calloc((size_t)threads + 1, sizeof(int));
  int used = 1; /* the team size the runtime gave */
#pragma omp parallel num_threads(threads)
  {
    const int t = omp_get_thread_num(), T = omp_get_num_threads();
    const int r0 = (int)((long long)m * t / T), r1 = (int)((long long)m * (t + 1) / T);
    for (int i = r0; i < r1; ++i) y_dense[i] = 0.f;
    for (int i = 0; i < xnnz; ++i) {
      const int c = x_key[i];
      const float xv = x_val[i];
      int lo = col_ptr[c], hi = col_ptr[c + 1];
      while (lo < hi) { /* first entry with row >= r0 */
        const int mid = lo + (hi - lo) / 2;
        if (row_idx[mid] < r0) lo = mid + 1; else hi = mid;
      }
      for (int j = lo; j < col_ptr[c + 1] && row_idx[j] < r1; ++j) y_dense[row_idx[j]] += cval[j] * xv;
    }
    int k = 0;
    for (int i = r0; i < r1; ++i) k += !(fabsf(y_dense[i]) < FLT_EPSILON);
    cnt[t + 1] = k;
#pragma omp barrier
#pragma omp single
    {
      used = T;
      for (int u = 0; u < T; ++u) cnt[u + 1] += cnt[u];
    }
    int o = cnt[t];
    for (int i = r0; i < r1; ++i) {
      if (!(fabsf(y_dense[i]) < FLT_EPSILON)) {
        if (y_key) y_key[o] = i;
        if (y_val) y_val[o] = y_dense[i];
        ++o;
      }
    }
  }
  const int total = cnt[used];
  free(cnt);
  return total;
}

int oracle_spmspv_f64(int m, const int* col_ptr, const int* row_idx,
                      const float* cval, int xnnz, const int* x_key,
                      const float* x_val, double* y64, double* bound,
                      int* y_key, double* y_val, double* y_bound) {
  /* Column-driven accumulation of SpmSpvCSCHost (test_sort_spmspv.cu:342-351)
   * in FP64. bound[i] = sum |a_ij x_j| doubles as the structural mark: it
   * is set to -1 before any touch, so a touched row has bound >= 0 even when
   * every product is 0, or NaN when a product is NaN. */
  for (int i = 0; i < m; ++i) {
    y64[i] = 0.0;
    bound[i] = -1.0;
  }
  for (int t = 0; t < xnnz; ++t) {
    int c = x_key[t];
    double xv = (double)x_val[t];
    for (int j = col_ptr[c]; j < col_ptr[c + 1]; ++j) {
      int r = row_idx[j];
      double p = (double)cval[j] * xv;
      if (bound[r] < 0.0) bound[r] = 0.0;
      y64[r] += p;
      bound[r] += fabs(p);
    }
  }
  int idx = 0;
  for (int i = 0; i < m; ++i) {
    /* touched: bound >= 0, or NaN once a NaN product came in (not < 0) */
    if (!(bound[i] < 0.0)) {
      if (y_key) y_key[idx] = i;
      if (y_val) y_val[idx] = y64[i];
      if (y_bound) y_bound[idx] = bound[i];
      idx++;
    }
  }
  return idx;
}

void oracle_spmv_f64(int m, const int* row_ptr, const int* col,
                     const float* val, const float* x_dense, double* y,
                     double* bound) {
  for (int i = 0; i < m; ++i) {
    double s = 0.0, b = 0.0;
    for (int j = row_ptr[i]; j < row_ptr[i + 1]; ++j) {
      double p = (double)val[j] * (double)x_dense[col[j]];
      s += p;
      b += fabs(p);
    }
    y[i] = s;
    if (bound) bound[i] = b;
  }
}

int oracle_bfs(int n, const int* col_ptr, const int* row_idx, int src,
               int* levels, int* queue) {
  for (int i = 0; i < n; ++i) levels[i] = -1;
  if (src < 0 || src >= n) return 0;
  int head = 0, tail = 0, maxlev = 0;
  levels[src] = 0;
  queue[tail++] = src;
  while (head < tail) {
    int u = queue[head++];
    int lu = levels[u];
    for (int j = col_ptr[u]; j < col_ptr[u + 1]; ++j) {
      int v = row_idx[j];
      if (levels[v] < 0) {
        levels[v] = lu + 1;
        if (lu + 1 > maxlev) maxlev = lu + 1;
        queue[tail++] = v;
      }
    }
  }
  return maxlev + 1;
}

static int cmp_int(const void* a, const void* b) {
  int x = *(const int*)a, y = *(const int*)b;
  return (x > y) - (x < y);
}

void oracle_mat_features(int m, int n, int nnz, const int* row_ptr,
                         int* deg, double out[7]) {
  /* spmspv.h:2274-2334 */
  int mx = 0, mn = INT_MAX;
  double avg = (double)nnz / (double)m;
  double sum_sq = 0.0, sum_eq = 0.0;
  for (int i = 0; i < m; ++i) {
    int d = row_ptr[i + 1] - row_ptr[i];
    deg[i] = d;
    sum_sq += (d - avg) * (d - avg);
    if (d != 0) {
      double p = d / (1.0 * nnz);
      sum_eq += -p * log(p);
    }
    if (d > mx) mx = d;
    if (d < mn) mn = d;
  }
  double stdv = sqrt(sum_sq / m);
  double equality = sum_eq * (1.0 / log(m * 1.0));
  qsort(deg, (size_t)m, sizeof(int), cmp_int);
  /* :2318-2324: j runs 1..m-1 over deg[j]; S2 runs over deg[0..m-2] */
  double s1 = 0.0, s2 = 0.0;
  for (int j = 1; j < m; ++j) {
    s1 += (double)j * deg[j];
    s2 += (double)deg[j - 1];
  }
  double gini = (2.0 * s1) / (m * s2) - (m + 1) / (double)m;
  if (gini < 0) gini = -gini;
  out[0] = mx;
  out[1] = mn;
  out[2] = avg;
  out[3] = (mx - mn) / (1.0 * n);
  out[4] = stdv;
  out[5] = equality;
  out[6] = gini;
}

double oracle_std_row_trunc(int m, int nnz, const int* row_ptr) {
  int avg = nnz / m; /* spmspv.h:929 int mean */
  double s = 0.0;
  for (int i = 0; i < m; ++i) {
    double d = (double)(row_ptr[i + 1] - row_ptr[i]) - avg;
    s += d * d;
  }
  return sqrt(s / m);
}

void oracle_vec_features(const int* col_ptr, int xnnz, const int* x_key,
                         int64_t out[3]) {
  /* spmspv.h:2346-2356 */
  int64_t sum = 0;
  int mx = 0, mn = INT_MAX;
  for (int i = 0; i < xnnz; ++i) {
    int c = x_key[i];
    int num = col_ptr[c + 1] - col_ptr[c];
    sum += num;
    if (num > mx) mx = num;
    if (num < mn) mn = num;
  }
  out[0] = sum;
  out[1] = mx;
  out[2] = mn;
}

void oracle_gm(int64_t m, int64_t nnz, int64_t xnnz, int64_t bin_len,
               double out[3]) {
  /* spmspv.h:2724-2753 with sizeof(int) = sizeof(float) = 4 */
  out[0] = (double)(xnnz * 4 + bin_len * 4 + bin_len * 4 + xnnz * 4 +
                    xnnz * 4 + bin_len * 4);
  out[1] = (double)((m + 1) * 4 + bin_len * 4 + nnz * 4 + bin_len * 4 + m * 4);
  out[2] = (double)((m + 1) * 4 + nnz * 4 + nnz * 4 + nnz * 4 + m * 4);
}

int64_t oracle_scan_col_lengths(const int* col_ptr, int xnnz, const int* x_key,
                                int64_t* ptr) {
  int64_t s = 0;
  ptr[0] = 0;
  for (int i = 0; i < xnnz; ++i) {
    int c = x_key[i];
    s += col_ptr[c + 1] - col_ptr[c];
    ptr[i + 1] = s;
  }
  return s;
}

/* ---------------- PageRank, FP64 ---------------- */

int oracle_pagerank_f64(int n, const int* col_ptr, const int* row_idx, const float* cval, double damping,
                        double eps, int max_iter, double* rank, int* sizes) {
  double* w = (double*)malloc(sizeof(double) * (size_t)n);
  double* r = (double*)malloc(sizeof(double) * (size_t)n);
  double* y = (double*)calloc((size_t)n, sizeof(double));
  double* x = (double*)malloc(sizeof(double) * (size_t)n);
  int* act = (int*)malloc(sizeof(int) * (size_t)n);
  unsigned char* hit = (unsigned char*)calloc((size_t)n, 1);
  for (int u = 0; u < n; ++u) {
    double s = 0.0;
    for (int k = col_ptr[u]; k < col_ptr[u + 1]; ++k) s += cval[k];
    w[u] = s;
    rank[u] = 0.0;
    r[u] = (1.0 - damping) / n;
    act[u] = u;
  }
  const double thr = eps * (1.0 - damping) / n;
  int cnt = n, it = 0;
  while (cnt > 0 && it < max_iter) {
    if (sizes) sizes[it] = cnt;
    for (int i = 0; i < cnt; ++i) {
      const int u = act[i];
      rank[u] += r[u];
      x[i] = w[u] > 0.0 ? r[u] / w[u] : 0.0;
      r[u] = 0.0;
    }
    for (int i = 0; i < cnt; ++i) {
      const int u = act[i];
      for (int k = col_ptr[u]; k < col_ptr[u + 1]; ++k) {
        y[row_idx[k]] += (double)cval[k] * x[i];
        hit[row_idx[k]] = 1;
      }
    }
    int nc = 0;
    for (int v = 0; v < n; ++v) {
      if (!hit[v]) continue;
      hit[v] = 0;
      r[v] += damping * y[v];
      y[v] = 0.0;
      if (r[v] > thr) act[nc++] = v;
    }
    cnt = nc;
    ++it;
  }
  for (int v = 0; v < n; ++v) rank[v] += r[v];
  free(w);
  free(r);
  free(y);
  free(x);
  free(act);
  free(hit);
  return it;
}

/* ---------------- selector MLP, FP64 ---------------- */

static void mlp_offsets(int F, int H1, int H2, int C, size_t off[6]) {
  off[0] = 0;                               /* W1 */
  off[1] = off[0] + (size_t)H1 * F;         /* b1 */
  off[2] = off[1] + (size_t)H1;             /* W2 */
  off[3] = off[2] + (size_t)H2 * H1;        /* b2 */
  off[4] = off[3] + (size_t)H2;             /* W3 */
  off[5] = off[4] + (size_t)C * H2;         /* b3 */
}

static void dense_fwd(int in, int out, const double* W, const double* b,
                      const double* x, double* y, int relu) {
  for (int o = 0; o < out; ++o) {
    double s = b[o];
    for (int i = 0; i < in; ++i) s += W[(size_t)o * in + i] * x[i];
    y[o] = (relu && s < 0.0) ? 0.0 : s;
  }
}

void oracle_mlp_forward(int F, int H1, int H2, int C, int B,
                        const double* params, const double* X,
                        double* logits) {
  size_t off[6];
  mlp_offsets(F, H1, H2, C, off);
  double* h1 = (double*)malloc(sizeof(double) * (size_t)H1);
  double* h2 = (double*)malloc(sizeof(double) * (size_t)H2);
  for (int s = 0; s < B; ++s) {
    dense_fwd(F, H1, params + off[0], params + off[1], X + (size_t)s * F, h1, 1);
    dense_fwd(H1, H2, params + off[2], params + off[3], h1, h2, 1);
    dense_fwd(H2, C, params + off[4], params + off[5], h2,
              logits + (size_t)s * C, 0);
  }
  free(h1);
  free(h2);
}

double oracle_mlp_loss_grad(int F, int H1, int H2, int C, int B,
                            const double* params, const double* X,
                            const int* labels, const double* soft,
                            double* grad) {
  size_t off[6];
  mlp_offsets(F, H1, H2, C, off);
  size_t np = off[5] + (size_t)C;
  if (grad) memset(grad, 0, sizeof(double) * np);
  double* h1 = (double*)malloc(sizeof(double) * (size_t)H1);
  double* h2 = (double*)malloc(sizeof(double) * (size_t)H2);
  double* z = (double*)malloc(sizeof(double) * (size_t)C);
  double* dz = (double*)malloc(sizeof(double) * (size_t)C);
  double* d2 = (double*)malloc(sizeof(double) * (size_t)H2);
  double* d1 = (double*)malloc(sizeof(double) * (size_t)H1);
  double loss = 0.0;
  for (int s = 0; s < B; ++s) {
    const double* x = X + (size_t)s * F;
    dense_fwd(F, H1, params + off[0], params + off[1], x, h1, 1);
    dense_fwd(H1, H2, params + off[2], params + off[3], h1, h2, 1);
    dense_fwd(H2, C, params + off[4], params + off[5], h2, z, 0);
    double zmax = z[0];
    for (int c = 1; c < C; ++c) zmax = z[c] > zmax ? z[c] : zmax;
    double se = 0.0;
    for (int c = 0; c < C; ++c) se += exp(z[c] - zmax);
    double lse = zmax + log(se);
    for (int c = 0; c < C; ++c) {
      double p = exp(z[c] - lse);
      double t = soft ? soft[(size_t)s * C + c] : (labels[s] == c ? 1.0 : 0.0);
      loss += -t * (z[c] - lse);
      dz[c] = (p - t) / B;
    }
    if (!grad) continue;
    /* layer 3 */
    for (int c = 0; c < C; ++c) {
      for (int k = 0; k < H2; ++k) grad[off[4] + (size_t)c * H2 + k] += dz[c] * h2[k];
      grad[off[5] + c] += dz[c];
    }
    for (int k = 0; k < H2; ++k) {
      double a = 0.0;
      for (int c = 0; c < C; ++c) a += params[off[4] + (size_t)c * H2 + k] * dz[c];
      d2[k] = h2[k] > 0.0 ? a : 0.0;
    }
    /* layer 2 */
    for (int k = 0; k < H2; ++k) {
      for (int j = 0; j < H1; ++j) grad[off[2] + (size_t)k * H1 + j] += d2[k] * h1[j];
      grad[off[3] + k] += d2[k];
    }
    for (int j = 0; j < H1; ++j) {
      double a = 0.0;
      for (int k = 0; k < H2; ++k) a += params[off[2] + (size_t)k * H1 + j] * d2[k];
      d1[j] = h1[j] > 0.0 ? a : 0.0;
    }
    /* layer 1 */
    for (int j = 0; j < H1; ++j) {
      for (int i = 0; i < F; ++i) grad[off[0] + (size_t)j * F + i] += d1[j] * x[i];
      grad[off[1] + j] += d1[j];
    }
  }
  free(h1);
  free(h2);
  free(z);
  free(dz);
  free(d2);
  free(d1);
  return loss / B;
}
