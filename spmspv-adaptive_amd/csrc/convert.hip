// Matrix preparation and matrix features on gfx950:
//  * CSR -> CSC transpose (ToCSC, spmspv.h:383-419, which called
//    cusparseScsr2csc): expand row ids, stable radix sort of (col, nnz id)
//    with the in-house sort, gather rows/values, and column pointers from
//    the sorted column ids. Stability keeps rows ascending in each column.
//  * Row-degree features (computeMatFeture_serial, spmspv.h:2274-2334, and
//    the int-mean stdRow of extractMatFeatureParallel, spmspv.h:920-976):
//    one grid-stride pass with per-workgroup partials, a degree histogram,
//    and the gini sums taken from the histogram instead of a sort (the
//    position sum of a run of equal degrees is an arithmetic series).
#include <cfloat>
#include <climits>

#include "kernels.hpp"

namespace spmspv {


__global__ void k_expand_rows(const int* __restrict__ row_ptr, int m, int* __restrict__ rowid) {
  for (long long r = (long long)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; r < m;
       r += (long long)gridDim.x * (blockDim.x / kWave)) {
    const int b = row_ptr[r], e = row_ptr[r + 1];
    for (int j = b + lane(); j < e; j += kWave) rowid[j] = (int)r;
  }
}

__global__ void k_keys_from_cols(const int* __restrict__ col, int nnz, unsigned* __restrict__ key,
                                 unsigned* __restrict__ idx) {
  for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < nnz;
       j += (long long)gridDim.x * blockDim.x) {
    key[j] = (unsigned)col[j];
    idx[j] = (unsigned)j;
  }
}

__global__ void k_gather_csc(const unsigned* __restrict__ skey, const unsigned* __restrict__ sidx,
                             const int* __restrict__ rowid, const float* __restrict__ val, int nnz,
                             int n, int* __restrict__ crow, float* __restrict__ cval,
                             int* __restrict__ colptr) {
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < nnz;
       p += (long long)gridDim.x * blockDim.x) {
    const unsigned j = sidx[p];
    crow[p] = rowid[j];
    cval[p] = val[j];
    const int c = (int)skey[p];
    const int cprev = p == 0 ? -1 : (int)skey[p - 1];
    for (int q = cprev + 1; q <= c; ++q) colptr[q] = (int)p;  // columns (cprev, c] start at p
    if (p == nnz - 1)
      for (int q = c + 1; q <= n; ++q) colptr[q] = nnz;
  }
}

__global__ void k_set_ll(long long* p, long long v) { *p = v; }

hipError_t launch_csr_to_csc_sorted(hipStream_t s, const int* row_ptr, const int* col, const float* val,
                              int m, int n, int nnz, int* colptr, int* crow, float* cval,
                              const SortBufs& b, long long* n_dev, Ctl* ctl, Ep ep) {
  if (nnz == 0) {
    launch_fill_i32(s, colptr, (long long)n + 1, 0);
    return hipGetLastError();
  }
  const int g = (int)std::min<long long>(8192, ((long long)nnz + 255) / 256);
  // row id of every nonzero, parked in crow until the gather below
  hipLaunchKernelGGL(k_expand_rows, dim3(std::min(8192, (m + 3) / 4 + 1)), dim3(256), 0, s, row_ptr,
                     m, crow);
  hipLaunchKernelGGL(k_keys_from_cols, dim3(g), dim3(256), 0, s, col, nnz, b.keyA,
                     reinterpret_cast<unsigned*>(b.valA));
  hipLaunchKernelGGL(k_set_ll, dim3(1), dim3(1), 0, s, n_dev, (long long)nnz);
  const int fb = launch_radix_sort(s, b, n, n_dev, ctl, nnz, nnz, false, ep);
  const unsigned* sk = fb ? b.keyB : b.keyA;
  const unsigned* si = reinterpret_cast<const unsigned*>(fb ? b.valB : b.valA);
  // gather into the free key buffer, then move the rows into crow
  int* rows_out = reinterpret_cast<int*>(fb ? b.keyA : b.keyB);
  hipLaunchKernelGGL(k_gather_csc, dim3(g), dim3(256), 0, s, sk, si, crow, val, nnz, n, rows_out,
                     cval, colptr);
  return hipMemcpyAsync(crow, rows_out, sizeof(int) * (size_t)nnz, hipMemcpyDeviceToDevice, s);
}

// ------------------------------------------------------- mat features
template <int NT>
__global__ __launch_bounds__(NT) void k_mat_stats(const int* __restrict__ row_ptr, int m, int n,
                                                  int nnz, int* hist, double* partials, MatFeat* out,
                                                  Ctl* ctl) {
  __shared__ double s_d[NT / kWave];
  __shared__ int s_i[NT / kWave];
  __shared__ unsigned s_flag;
  // degree histogram: small degrees in LDS (flushed once per workgroup),
  // the rare large ones straight to memory
  constexpr int HB = 2048;
  __shared__ int s_hist[HB];
  for (int k = threadIdx.x; k < HB; k += NT) s_hist[k] = 0;
  __syncthreads();
  const double avg = (double)nnz / (double)m;
  const int avg_trunc = nnz / m;
  int mx = 0, mn = INT_MAX;
  double sq = 0.0, eq = 0.0, sqt = 0.0;
  for (long long r = (long long)blockIdx.x * NT + threadIdx.x; r < m; r += (long long)gridDim.x * NT) {
    const int d = row_ptr[r + 1] - row_ptr[r];
    mx = max(mx, d);
    mn = min(mn, d);
    sq += (d - avg) * (d - avg);
    const double dt = (double)(d - avg_trunc);
    sqt += dt * dt;
    if (d != 0) {
      const double p = d / (1.0 * nnz);
      eq += -p * log(p);
    }
    const int b = min(d, n);  // rows with duplicate columns stay in range
    if (b < HB) atomicAdd(&s_hist[b], 1);
    else atomicAdd(hist + b, 1);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < HB && k <= n; k += NT)
    if (s_hist[k]) atomicAdd(hist + k, s_hist[k]);
  mx = block_max<NT, int>(mx, s_i);
  mn = block_min<NT, int>(mn, s_i);
  sq = block_sum<NT, double>(sq, s_d);
  eq = block_sum<NT, double>(eq, s_d);
  sqt = block_sum<NT, double>(sqt, s_d);
  if (threadIdx.x == 0) {
    double* pp = partials + 5 * blockIdx.x;
    pp[0] = mx;
    pp[1] = mn;
    pp[2] = sq;
    pp[3] = eq;
    pp[4] = sqt;
  }
  const bool last = finish_persistent(ctl, kSlotMat, &s_flag);
  if (!last) return;
  double gmx = 0, gmn = 1e300, gsq = 0, geq = 0, gsqt = 0;
  if (threadIdx.x == 0) {
    for (unsigned b = 0; b < gridDim.x; ++b) {  // fixed order: deterministic sums
      const double* pp = partials + 5 * b;
      gmx = fmax(gmx, pp[0]);
      gmn = fmin(gmn, pp[1]);
      gsq += pp[2];
      geq += pp[3];
      gsqt += pp[4];
    }
    out->max_row = gmx;
    out->min_row = gmn;
    out->avg_row = avg;
    out->range = (gmx - gmn) / (1.0 * n);
    out->std_row = sqrt(gsq / m);
    out->equality = geq * (1.0 / log(m * 1.0));
    out->std_trunc = sqrt(gsqt / m);
    out->m = m;
    out->n = n;
    out->nnz = nnz;
  }
}

// gini from the degree histogram: ascending degrees d_(0..m-1);
// S1 = sum_{j=1}^{m-1} j*d_(j), S2 = sum_{j=0}^{m-2} d_(j) = nnz - d_(m-1)
// (spmspv.h:2318-2324). A run of c equal degrees v at positions
// [pos, pos+c) adds v * (c*pos + c*(c-1)/2) to S1.
template <int NT>
__global__ __launch_bounds__(NT) void k_gini(const int* __restrict__ hist, int m, int n, int nnz,
                                             MatFeat* out) {
  __shared__ long long s_l[NT / kWave];
  __shared__ double s_d[NT / kWave];
  const int maxd = min((int)out->max_row, n);
  long long pos = 0;
  double s1 = 0.0;
  for (int base = 0; base <= maxd; base += NT) {
    const int v = base + threadIdx.x;
    const long long c = v <= maxd ? hist[v] : 0;
    long long tot;
    const long long ex = block_excl_sum<NT, long long>(c, s_l, &tot);
    const double p0 = (double)(pos + ex);
    double contrib = (double)v * ((double)c * p0 + (double)c * (double)(c - 1) * 0.5);
    if (c == 0) contrib = 0.0;
    s1 += contrib;
    pos += tot;
  }
  s1 = block_sum<NT, double>(s1, s_d);
  if (threadIdx.x == 0) {
    const double s2 = (double)nnz - (double)maxd;
    double g = (2.0 * s1) / ((double)m * s2) - (m + 1) / (double)m;
    if (g < 0) g = -g;
    out->gini = g;
  }
}

void launch_mat_features(hipStream_t s, const int* row_ptr, int m, int n, int nnz, int* hist,
                         double* partials, MatFeat* out, Ctl* ctl) {
  launch_fill_i32(s, hist, (long long)n + 1, 0);
  const int grid = std::max(1, std::min(256, (m + 255) / 256));
  hipLaunchKernelGGL(k_mat_stats<256>, dim3(grid), dim3(256), 0, s, row_ptr, m, n, nnz, hist,
                     partials, out, ctl);
  hipLaunchKernelGGL(k_gini<1024>, dim3(1), dim3(1024), 0, s, hist, m, n, nnz, out);
}

}  // namespace spmspv
