// Synthetic inputs for tests, benchmarks and the selector corpus (the
// reference read MatrixMarket files instead, la/include/spmspv/mtx.hpp).
// Every random draw is a hash of (seed, stream, index), so results are
// identical for any thread count.
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <unordered_set>
#include <vector>

#include "../../include/spmspv_amd.h"

namespace {

inline uint64_t mix64(uint64_t z) {  // splitmix64 finaliser
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t h3(uint64_t seed, uint64_t a, uint64_t b) {
  return mix64(mix64(mix64(seed) ^ a) ^ (b * 0xD6E8FEB86659FD93ull));
}
inline double u01(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
inline float u_pm1(uint64_t r) { return (float)(2.0 * u01(r) - 1.0); }                      // [-1,1)
inline float u_0_1(uint64_t r) { return (float)(1.0 - u01(r)); }                            // (0,1]

struct Rng {  // sequential stream for per-row/column sampling
  uint64_t s;
  uint64_t next() { return s = mix64(s); }
  uint64_t below(uint64_t n) { return n ? next() % n : 0; }
};

// Floyd's algorithm: k distinct values in [0, n), sorted.
void floyd_sorted(Rng& r, int n, int k, int* out) {
  std::vector<int> pick;
  pick.reserve(k);
  if (k * 4 > n) {  // dense: selection sampling (Knuth S)
    int need = k;
    for (int i = 0; i < n && need > 0; ++i) {
      if ((double)r.below((uint64_t)(n - i)) < need) {
        pick.push_back(i);
        --need;
      }
    }
  } else {
    std::unordered_set<int> seen;
    seen.reserve(k * 2);
    for (int j = n - k; j < n; ++j) {
      int t = (int)r.below((uint64_t)j + 1);
      if (!seen.insert(t).second) {
        seen.insert(j);
        t = j;
      }
      pick.push_back(t);
    }
    std::sort(pick.begin(), pick.end());
  }
  std::copy(pick.begin(), pick.end(), out);
}

struct RmatCache {
  std::mutex mu;
  int scale = -1, ef = -1;
  double a = 0, b = 0, c = 0;
  uint64_t seed = 0;
  std::vector<int> row_ptr, col;
  std::vector<float> val;
} g_rmat;

void build_rmat(int scale, int ef, double a, double b, double c, uint64_t seed,
                std::vector<int>& row_ptr, std::vector<int>& col, std::vector<float>& val) {
  const int64_t n = 1ll << scale;
  const int64_t E = (int64_t)ef * n;
  const double ab = a + b, abc = a + b + c;
  // pass 1: endpoints of every edge (skip self loops), count both directions
  std::vector<std::pair<int, int>> edges((size_t)E);
  std::vector<std::atomic<int>> cnt((size_t)n);
  for (auto& x : cnt) x.store(0, std::memory_order_relaxed);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < E; ++e) {
    int64_t i = 0, j = 0;
    for (int l = 0; l < scale; ++l) {
      const double u = u01(h3(seed, (uint64_t)e, (uint64_t)l));
      const int64_t bit = 1ll << (scale - 1 - l);
      if (u < a) {
      } else if (u < ab) {
        j |= bit;
      } else if (u < abc) {
        i |= bit;
      } else {
        i |= bit;
        j |= bit;
      }
    }
    edges[(size_t)e] = {(int)i, (int)j};
    if (i != j) {
      cnt[(size_t)i].fetch_add(1, std::memory_order_relaxed);
      cnt[(size_t)j].fetch_add(1, std::memory_order_relaxed);
    }
  }
  std::vector<int64_t> start((size_t)n + 1, 0);
  for (int64_t v = 0; v < n; ++v) start[(size_t)v + 1] = start[(size_t)v] + cnt[(size_t)v].load();
  std::vector<int> adj((size_t)start[(size_t)n]);
  std::vector<std::atomic<int64_t>> pos((size_t)n);
  for (int64_t v = 0; v < n; ++v) pos[(size_t)v].store(start[(size_t)v], std::memory_order_relaxed);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < E; ++e) {
    const int i = edges[(size_t)e].first, j = edges[(size_t)e].second;
    if (i == j) continue;
    adj[(size_t)pos[(size_t)i].fetch_add(1, std::memory_order_relaxed)] = j;
    adj[(size_t)pos[(size_t)j].fetch_add(1, std::memory_order_relaxed)] = i;
  }
  edges.clear();
  edges.shrink_to_fit();
  // pass 2: sort + dedup each row
  std::vector<int> deg((size_t)n);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t v = 0; v < n; ++v) {
    int* b0 = adj.data() + start[(size_t)v];
    int* e0 = adj.data() + start[(size_t)v + 1];
    std::sort(b0, e0);
    deg[(size_t)v] = (int)(std::unique(b0, e0) - b0);
  }
  row_ptr.assign((size_t)n + 1, 0);
  for (int64_t v = 0; v < n; ++v) row_ptr[(size_t)v + 1] = row_ptr[(size_t)v] + deg[(size_t)v];
  const int64_t nnz = row_ptr[(size_t)n];
  col.resize((size_t)nnz);
  val.resize((size_t)nnz);
  const uint64_t vseed = mix64(seed ^ 0x5A5A5A5A5A5A5A5Aull);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t v = 0; v < n; ++v) {
    const int* src = adj.data() + start[(size_t)v];
    int64_t o = row_ptr[(size_t)v];
    for (int k = 0; k < deg[(size_t)v]; ++k) {
      const int j = src[k];
      col[(size_t)(o + k)] = j;
      const uint64_t lo = (uint64_t)std::min<int64_t>(v, j), hi = (uint64_t)std::max<int64_t>(v, j);
      val[(size_t)(o + k)] = u_0_1(h3(vseed, lo, hi));  // same value on (i,j) and (j,i)
    }
  }
}

}  // namespace

extern "C" {

int spmspv_gen_uniform_csc(int n, int k, uint64_t seed, int* col_ptr, int* row, float* val) {
  if (n <= 0 || k < 0 || k > n || !col_ptr || (k > 0 && (!row || !val))) return SPMSPV_INVALID_ARGUMENT;
  for (int j = 0; j <= n; ++j) col_ptr[j] = (int)std::min<int64_t>((int64_t)j * k, INT32_MAX);
#pragma omp parallel for schedule(dynamic, 256)
  for (int j = 0; j < n; ++j) {
    Rng r{h3(seed, 1, (uint64_t)j)};
    floyd_sorted(r, n, k, row + (int64_t)j * k);
    for (int t = 0; t < k; ++t) val[(int64_t)j * k + t] = u_pm1(h3(seed, 2, (uint64_t)j * k + t));
  }
  return SPMSPV_SUCCESS;
}

int spmspv_gen_rmat_csr(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                        long long* nnz, int* row_ptr, int* col, float* val) {
  if (scale < 1 || scale > 30 || edge_factor < 1 || !nnz) return SPMSPV_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(g_rmat.mu);
  if (!(g_rmat.scale == scale && g_rmat.ef == edge_factor && g_rmat.a == a && g_rmat.b == b &&
        g_rmat.c == c && g_rmat.seed == seed)) {
    g_rmat.scale = -1;
    build_rmat(scale, edge_factor, a, b, c, seed, g_rmat.row_ptr, g_rmat.col, g_rmat.val);
    g_rmat.scale = scale;
    g_rmat.ef = edge_factor;
    g_rmat.a = a;
    g_rmat.b = b;
    g_rmat.c = c;
    g_rmat.seed = seed;
  }
  *nnz = (long long)g_rmat.col.size();
  if (row_ptr) {
    std::memcpy(row_ptr, g_rmat.row_ptr.data(), sizeof(int) * g_rmat.row_ptr.size());
    if (col) std::memcpy(col, g_rmat.col.data(), sizeof(int) * g_rmat.col.size());
    if (val) std::memcpy(val, g_rmat.val.data(), sizeof(float) * g_rmat.val.size());
    // release the cache once copied out (large graphs)
    std::vector<int>().swap(g_rmat.row_ptr);
    std::vector<int>().swap(g_rmat.col);
    std::vector<float>().swap(g_rmat.val);
    g_rmat.scale = -1;
  }
  return SPMSPV_SUCCESS;
}

int spmspv_gen_sparse_vector(int n, int k, uint64_t seed, float lo, float hi, int* key, float* val) {
  if (n <= 0 || k < 0 || k > n || (k > 0 && !key)) return SPMSPV_INVALID_ARGUMENT;
  Rng r{h3(seed, 3, 0)};
  floyd_sorted(r, n, k, key);
  if (val)
    for (int t = 0; t < k; ++t) val[t] = (float)(lo + (hi - lo) * u01(h3(seed, 4, (uint64_t)t)));
  return SPMSPV_SUCCESS;
}

int spmspv_gen_banded_csr(int n, int bw, uint64_t seed, long long* nnz, int* row_ptr, int* col,
                          float* val) {
  if (n <= 0 || bw < 0 || !nnz) return SPMSPV_INVALID_ARGUMENT;
  long long tot = 0;
  for (int i = 0; i < n; ++i) tot += std::min(n - 1, i + bw) - std::max(0, i - bw) + 1;
  *nnz = tot;
  if (!row_ptr) return SPMSPV_SUCCESS;
  if (tot > INT32_MAX) return SPMSPV_INVALID_ARGUMENT;
  row_ptr[0] = 0;
  for (int i = 0; i < n; ++i)
    row_ptr[i + 1] = row_ptr[i] + std::min(n - 1, i + bw) - std::max(0, i - bw) + 1;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    int o = row_ptr[i];
    for (int j = std::max(0, i - bw); j <= std::min(n - 1, i + bw); ++j, ++o) {
      col[o] = j;
      val[o] = u_0_1(h3(seed, (uint64_t)i, (uint64_t)j));
    }
  }
  return SPMSPV_SUCCESS;
}

int spmspv_gen_blockdiag_csr(int n, int bs, uint64_t seed, long long* nnz, int* row_ptr, int* col,
                             float* val) {
  if (n <= 0 || bs <= 0 || !nnz) return SPMSPV_INVALID_ARGUMENT;
  long long tot = 0;
  for (int i = 0; i < n; ++i) {
    const int b0 = (i / bs) * bs;
    tot += std::min(n, b0 + bs) - b0;
  }
  *nnz = tot;
  if (!row_ptr) return SPMSPV_SUCCESS;
  if (tot > INT32_MAX) return SPMSPV_INVALID_ARGUMENT;
  row_ptr[0] = 0;
  for (int i = 0; i < n; ++i) {
    const int b0 = (i / bs) * bs;
    row_ptr[i + 1] = row_ptr[i] + std::min(n, b0 + bs) - b0;
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const int b0 = (i / bs) * bs;
    int o = row_ptr[i];
    for (int j = b0; j < std::min(n, b0 + bs); ++j, ++o) {
      col[o] = j;
      val[o] = u_0_1(h3(seed, (uint64_t)i, (uint64_t)j));
    }
  }
  return SPMSPV_SUCCESS;
}

int spmspv_gen_random_csr(int m, int n, double mean, uint64_t seed, long long* nnz, int* row_ptr,
                          int* col, float* val) {
  if (m <= 0 || n <= 0 || mean < 0 || !nnz) return SPMSPV_INVALID_ARGUMENT;
  // per-row count ~ Poisson(mean) (normal approximation above 30), capped at n
  std::vector<int> cntv((size_t)m);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < m; ++i) {
    int c;
    if (mean <= 30.0) {
      const double L = std::exp(-mean);
      double p = 1.0;
      c = -1;
      uint64_t t = 0;
      do {
        ++c;
        p *= u01(h3(seed, 5, (uint64_t)i * 1024 + (t++)));
      } while (p > L && c < n);
    } else {
      const double u1 = std::max(1e-12, u01(h3(seed, 6, (uint64_t)i)));
      const double u2 = u01(h3(seed, 7, (uint64_t)i));
      const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
      c = (int)std::lround(mean + z * std::sqrt(mean));
    }
    cntv[(size_t)i] = std::max(0, std::min(n, c));
  }
  long long tot = 0;
  for (int i = 0; i < m; ++i) tot += cntv[(size_t)i];
  *nnz = tot;
  if (!row_ptr) return SPMSPV_SUCCESS;
  if (tot > INT32_MAX) return SPMSPV_INVALID_ARGUMENT;
  row_ptr[0] = 0;
  for (int i = 0; i < m; ++i) row_ptr[i + 1] = row_ptr[i] + cntv[(size_t)i];
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < m; ++i) {
    Rng r{h3(seed, 8, (uint64_t)i)};
    floyd_sorted(r, n, cntv[(size_t)i], col + row_ptr[i]);
    for (int t = 0; t < cntv[(size_t)i]; ++t)
      val[row_ptr[i] + t] = u_0_1(h3(seed, 9, (uint64_t)row_ptr[i] + t));
  }
  return SPMSPV_SUCCESS;
}

}  // extern "C"
