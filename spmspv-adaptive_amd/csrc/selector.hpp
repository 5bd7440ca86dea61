// Selector features and in-workgroup inference (fused into the last
// workgroup of k_prep, so one launch does scan + features + MLP).
//
// The reference's selector inputs are the 23 columns of its .info files
// (model/1-test_dtree.py:63): m, n, nnz, maxRow, minRow, avgRow, range,
// stdRow, edge_equality, gini, x_nnz, bin_len, max, min, xnnz/n, bin_len/nnz,
// xnnz_range, GM1, GM2, GM3 and their ratios. The 16 used here are the
// scale-free (log2 / ratio) forms of the same quantities.
#pragma once

#include "kernels.hpp"

namespace spmspv {

constexpr int kF = 16, kH1 = 64, kH2 = 64, kC = 8;
constexpr int kNParam = kH1 * kF + kH1 + kH2 * kH1 + kH2 + kC * kH2 + kC;
constexpr int kNWeights = kNParam + 2 * kF;  // + the feature normaliser (mean, 1/std)
static_assert((kF & (kF - 1)) == 0 && (kH1 & (kH1 - 1)) == 0 && (kH2 & (kH2 - 1)) == 0,
              "layer widths are powers of two (rotated row reads)");

__host__ __device__ inline void build_features(double m, double n, double nnz, double max_row,
                                               double avg_row, double std_row, double equality,
                                               double gini, double xnnz, double bin_len,
                                               double max_len, float* f) {
  const double gm1 = 4.0 * (3.0 * xnnz + 3.0 * bin_len);                   // spmspv.h:2727-2731
  const double gm2 = 4.0 * ((m + 1) + 2.0 * bin_len + nnz + m);            // spmspv.h:2737-2740
  const double gm3 = 4.0 * ((m + 1) + 3.0 * nnz + m);                      // spmspv.h:2746-2749
  f[0] = (float)log2(m > 1 ? m : 1.0);
  f[1] = (float)log2(n > 1 ? n : 1.0);
  f[2] = (float)log2(1.0 + nnz);
  f[3] = (float)log2(1.0 + max_row);
  f[4] = (float)log2(1.0 + avg_row);
  f[5] = (float)(std_row / (avg_row + 1e-9));
  f[6] = (float)equality;
  f[7] = (float)gini;
  f[8] = (float)log2(1.0 + xnnz);
  f[9] = (float)log2(1.0 + bin_len);
  f[10] = (float)log2(1.0 + max_len);
  f[11] = (float)log2(xnnz / (n > 0 ? n : 1.0) + 1e-7);
  f[12] = (float)log2(bin_len / (nnz > 0 ? nnz : 1.0) + 1e-7);
  f[13] = (float)log2(1.0 + bin_len / (xnnz > 1 ? xnnz : 1.0));
  f[14] = (float)log2((gm1 + 1.0) / gm3);
  f[15] = (float)log2(gm2 / gm3);
}

// Hand-written fallback rules (the role of the reference's 3-way rule stubs,
// la/test/new_test_app.cu:121-131): density picks the solution
// (column / row / SpMV), then skew picks naive vs load-balanced.
__host__ __device__ inline int tree_select(double n, double nnz, double avg_row, double std_row,
                                           double xnnz, double bin_len, double max_len) {
  const double density = xnnz / (n > 0 ? n : 1.0);
  const bool skewed = std_row > 2.0 * avg_row + 1.0 || max_len > 8.0 * (avg_row + 1.0);
  if (density >= 0.3) return skewed ? 7 : 6;
  if (density >= 0.05) return skewed ? 5 : 4;
  if (bin_len <= 3072) return 1;
  return skewed ? 2 : 0;
}

// Runs in one workgroup of NT >= 64 threads after ctl has bin_len/max/min.
template <int NT>
__device__ void select_in_block(Ctl* ctl, const SelectorArgs& sel) {
  __shared__ float s_x[kF];
  __shared__ float s_h1[kH1];
  __shared__ float s_h2[kH2];
  __shared__ float s_z[kC];
  const MatFeat* mf = sel.mat;
  if (threadIdx.x == 0) {
    float f[kF];
    build_features(mf->m, mf->n, mf->nnz, mf->max_row, mf->avg_row, mf->std_row, mf->equality,
                   mf->gini, (double)ctl->xnnz, (double)ctl->bin_len, (double)ctl->max_len, f);
    for (int i = 0; i < kF; ++i) {
      ctl->feat[i] = f[i];
      s_x[i] = f[i];
    }
  }
  __syncthreads();
  if (sel.mode == 2) {
    if (threadIdx.x == 0) ctl->kernel_choice = sel.fixed;
    return;
  }
  if (sel.mode == 1 || !sel.weights) {
    if (threadIdx.x == 0)
      ctl->kernel_choice = tree_select(mf->n, mf->nnz, mf->avg_row, mf->std_row, (double)ctl->xnnz,
                                       (double)ctl->bin_len, (double)ctl->max_len);
    return;
  }
  const float* W1 = sel.weights;
  const float* b1 = W1 + kH1 * kF;
  const float* W2 = b1 + kH1;
  const float* b2 = W2 + kH2 * kH1;
  const float* W3 = b2 + kH2;
  const float* b3 = W3 + kC * kH2;
  const float* mean = b3 + kC;
  const float* istd = mean + kF;
  if (threadIdx.x < kF) s_x[threadIdx.x] = (s_x[threadIdx.x] - mean[threadIdx.x]) * istd[threadIdx.x];
  __syncthreads();
  // rows are read at a rotated index (i + j) so that, with the weights in
  // LDS, the 64 threads of a layer hit 64 different banks
  if (threadIdx.x < kH1) {
    const int j = threadIdx.x;
    float a = b1[j];
    for (int q = 0; q < kF; ++q) {
      const int i = (q + j) & (kF - 1);
      a += W1[j * kF + i] * s_x[i];
    }
    s_h1[j] = a > 0.f ? a : 0.f;
  }
  __syncthreads();
  if (threadIdx.x < kH2) {
    const int j = threadIdx.x;
    float a = b2[j];
    for (int q = 0; q < kH1; ++q) {
      const int i = (q + j) & (kH1 - 1);
      a += W2[j * kH1 + i] * s_h1[i];
    }
    s_h2[j] = a > 0.f ? a : 0.f;
  }
  __syncthreads();
  if (threadIdx.x < kC) {
    const int j = threadIdx.x;
    float a = b3[j];
    for (int q = 0; q < kH2; ++q) {
      const int i = (q + j) & (kH2 - 1);
      a += W3[j * kH2 + i] * s_h2[i];
    }
    s_z[j] = a;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int best = 0;
    for (int c = 1; c < kC; ++c)
      if (s_z[c] > s_z[best]) best = c;
    ctl->kernel_choice = best;
  }
}

}  // namespace spmspv
