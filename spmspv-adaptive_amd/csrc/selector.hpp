// Selector features and one-wave inference (fused into the last workgroup
// of k_prep, so one launch does scan + features + MLP).
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

// Feature i of the selector input (0 <= i < kF). One function per feature
// index, so that the device computes the 16 features in 16 lanes at once
// and the host gets the same arithmetic. Every log2 feature is written as
// log2(num / den + add) (den = 1, add = 0 where the formula has no division
// or sum; both are exact then), so that the lanes pick their operands by
// selects and share one double division and one log2 instead of running
// their own branch each.
__host__ __device__ inline float feature_value(int i, double m, double n, double nnz, double max_row,
                                               double avg_row, double std_row, double equality,
                                               double gini, double xnnz, double bin_len,
                                               double max_len) {
  if (i == 5) return (float)(std_row / (avg_row + 1e-9));
  if (i == 6) return (float)equality;
  if (i == 7) return (float)gini;
  const double gm1 = 4.0 * (3.0 * xnnz + 3.0 * bin_len);                   // spmspv.h:2727-2731
  const double gm2 = 4.0 * ((m + 1) + 2.0 * bin_len + nnz + m);            // spmspv.h:2737-2740
  const double gm3 = 4.0 * ((m + 1) + 3.0 * nnz + m);                      // spmspv.h:2746-2749
  double num, den = 1.0, add = 0.0;
  switch (i) {
    case 0: num = m > 1 ? m : 1.0; break;
    case 1: num = n > 1 ? n : 1.0; break;
    case 2: num = 1.0 + nnz; break;
    case 3: num = 1.0 + max_row; break;
    case 4: num = 1.0 + avg_row; break;
    case 8: num = 1.0 + xnnz; break;
    case 9: num = 1.0 + bin_len; break;
    case 10: num = 1.0 + max_len; break;
    case 11: num = xnnz; den = n > 0 ? n : 1.0; add = 1e-7; break;
    case 12: num = bin_len; den = nnz > 0 ? nnz : 1.0; add = 1e-7; break;
    case 13: num = bin_len; den = xnnz > 1 ? xnnz : 1.0; add = 1.0; break;
    case 14: num = gm1 + 1.0; den = gm3; break;
    default: num = gm2; den = gm3; break;
  }
  return (float)log2(num / den + add);
}

__host__ __device__ inline void build_features(double m, double n, double nnz, double max_row,
                                               double avg_row, double std_row, double equality,
                                               double gini, double xnnz, double bin_len,
                                               double max_len, float* f) {
  for (int i = 0; i < kF; ++i)
    f[i] = feature_value(i, m, n, nnz, max_row, avg_row, std_row, equality, gini, xnnz, bin_len,
                         max_len);
}

// Weights as the inference reads them from LDS. The flat order (W1, b1,
// W2, b2, W3, b3, mean, 1/std; W row-major [out][in]) is kept per region,
// but each matrix is permuted so that the one wave running the network
// reads its rows in 16-byte pieces lying side by side across lanes (no
// bank conflicts): W1 and W2 as [in/4][out][4], W3 as [in/8][class][8]
// (lane = part * kC + class, 8 inputs per part).
constexpr int kOB1 = kH1 * kF, kOW2 = kOB1 + kH1, kOB2 = kOW2 + kH2 * kH1, kOW3 = kOB2 + kH2;
constexpr int kOB3 = kOW3 + kC * kH2, kOMean = kOB3 + kC, kOIstd = kOMean + kF;
constexpr int kW3Part = 8;
static_assert(kH1 == kWave && kH2 == kWave && kC * (kH2 / kW3Part) == kWave && kF <= kWave,
              "one lane per hidden unit; the last layer fills the wave with (part, class)");
static_assert(kOW2 % 4 == 0 && kOW3 % 4 == 0, "16-byte aligned matrices");
__host__ __device__ inline int wlds_index(int idx) {
  if (idx < kOB1) {
    const int j = idx / kF, k = idx % kF;
    return ((k >> 2) * kH1 + j) * 4 + (k & 3);
  }
  if (idx >= kOW2 && idx < kOB2) {
    const int r = idx - kOW2, j = r / kH1, k = r % kH1;
    return kOW2 + ((k >> 2) * kH2 + j) * 4 + (k & 3);
  }
  if (idx >= kOW3 && idx < kOB3) {
    const int r = idx - kOW3, c = r / kH2, k = r % kH2;
    return kOW3 + ((k / kW3Part) * kC + c) * kW3Part + (k % kW3Part);
  }
  return idx;
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

__device__ __forceinline__ float lane_bcast(float v, int k) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}

// The tuned size rule behind the selector: a load-balanced column or row
// slot (2, 3, 5, 7) becomes the bucket slot inside [bin_lo, bin_hi].
__device__ inline int bucket_rule(int best, double bin_len, const SelectorArgs& sel) {
  if ((best == 2 || best == 3 || best == 5 || best == 7) && bin_len >= (double)sel.bin_lo &&
      bin_len <= (double)sel.bin_hi)
    return 8;
  return best;
}

// Runs in wave 0 of the calling workgroup (the other waves return at
// once; no barrier inside). sel.weights, when set, points at the weights in
// LDS in the wlds_index layout. Writes ctl->feat and ctl->kernel_choice
// (lane 0 writes the choice).
__device__ void select_in_wave(Ctl* ctl, const SelectorArgs& sel, double xnnz, double bin_len,
                               double max_len) {
  if (threadIdx.x >= kWave) return;
  const int l = threadIdx.x;
  const MatFeat* mf = sel.mat;
  const float f = feature_value(l & (kF - 1), mf->m, mf->n, mf->nnz, mf->max_row, mf->avg_row,
                                mf->std_row, mf->equality, mf->gini, xnnz, bin_len, max_len);
  if (l < kF) ctl->feat[l] = f;
  if (sel.mode == 2) {
    if (l == 0) ctl->kernel_choice = sel.fixed;
    return;
  }
  if (sel.mode == 1 || !sel.weights) {
    if (l == 0)
      ctl->kernel_choice =
          bucket_rule(tree_select(mf->n, mf->nnz, mf->avg_row, mf->std_row, xnnz, bin_len, max_len), bin_len, sel);
    return;
  }
  const float* w = sel.weights;
  const int fi = l & (kF - 1);
  const float xn = (f - w[kOMean + fi]) * w[kOIstd + fi];
  // layer 1: lane j = hidden unit j; inputs broadcast from their lanes
  float a = w[kOB1 + l];
#pragma unroll
  for (int kk = 0; kk < kF / 4; ++kk) {
    const float4 q = *reinterpret_cast<const float4*>(w + ((kk * kH1 + l) << 2));
    a += q.x * lane_bcast(xn, 4 * kk);
    a += q.y * lane_bcast(xn, 4 * kk + 1);
    a += q.z * lane_bcast(xn, 4 * kk + 2);
    a += q.w * lane_bcast(xn, 4 * kk + 3);
  }
  const float h1 = a > 0.f ? a : 0.f;
  // layer 2
  a = w[kOB2 + l];
#pragma unroll
  for (int kk = 0; kk < kH1 / 4; ++kk) {
    const float4 q = *reinterpret_cast<const float4*>(w + kOW2 + ((kk * kH2 + l) << 2));
    a += q.x * lane_bcast(h1, 4 * kk);
    a += q.y * lane_bcast(h1, 4 * kk + 1);
    a += q.z * lane_bcast(h1, 4 * kk + 2);
    a += q.w * lane_bcast(h1, 4 * kk + 3);
  }
  const float h2 = a > 0.f ? a : 0.f;
  // layer 3: lane = part * kC + class; 8 inputs per lane, then the parts
  // are summed across lanes
  const int c = l & (kC - 1), part = l / kC;
  const float4 u0 = *reinterpret_cast<const float4*>(w + kOW3 + l * kW3Part);
  const float4 u1 = *reinterpret_cast<const float4*>(w + kOW3 + l * kW3Part + 4);
  const int k0 = part * kW3Part;
  float z = u0.x * __shfl(h2, k0);
  z += u0.y * __shfl(h2, k0 + 1);
  z += u0.z * __shfl(h2, k0 + 2);
  z += u0.w * __shfl(h2, k0 + 3);
  z += u1.x * __shfl(h2, k0 + 4);
  z += u1.y * __shfl(h2, k0 + 5);
  z += u1.z * __shfl(h2, k0 + 6);
  z += u1.w * __shfl(h2, k0 + 7);
#pragma unroll
  for (int d = kC; d < kWave; d <<= 1) z += __shfl_xor(z, d);
  z += w[kOB3 + c];
  // argmax over the classes in lanes 0..kC-1 (the first maximum wins)
  int best = 0;
  float bz = lane_bcast(z, 0);
#pragma unroll
  for (int k = 1; k < kC; ++k) {
    const float zk = lane_bcast(z, k);
    if (zk > bz) {
      bz = zk;
      best = k;
    }
  }
  // switching-aware: stay in the previous call's family when its best slot
  // is within the penalty (the soft targets make a logit gap a relative
  // time gap over tau, so the penalty is the switching cost in those units)
  if (sel.penalty > 0.f && sel.prev >= 0 && slot_family(best) != slot_family(sel.prev)) {
    int keep = -1;
    float kz = 0.f;
#pragma unroll
    for (int k = 0; k < kC; ++k) {
      const float zk = lane_bcast(z, k);
      if (slot_family(k) == slot_family(sel.prev) && zk >= bz - sel.penalty && (keep < 0 || zk > kz)) {
        keep = k;
        kz = zk;
      }
    }
    if (keep >= 0) best = keep;
  }
  if (l == 0) ctl->kernel_choice = bucket_rule(best, bin_len, sel);
}

}  // namespace spmspv
