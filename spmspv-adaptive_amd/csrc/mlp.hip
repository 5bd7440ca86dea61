// The learned kernel selector: an FP32 MLP 16 -> 64 -> 64 -> 8 (ReLU,
// softmax cross-entropy on soft targets), trained on the GPU by the kernels
// below. The reference trained sklearn trees offline and never wired them
// back (model/1-test_dtree.py:170-196; la/test/new_test_app.cu:121-131 are
// stubs returning 0), so this part has no reference counterpart.
//
// k_mlp_fwd_bwd: one workgroup walks a slice of the minibatch one sample at
// a time; every thread owns a fixed set of parameter gradients and keeps
// them in registers across its slice, then writes one partial row. The
// partial rows are summed in a fixed order by k_mlp_adam (deterministic
// gradients), which also applies the Adam step. Inference for dispatch
// runs inside the last workgroup of k_prep (selector.hpp).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "handle.hpp"
#include "selector.hpp"

using namespace spmspv;

namespace {

constexpr int kNT = 256;
constexpr int kGPT = (kNParam + kNT - 1) / kNT;  // gradient entries per thread
constexpr int oW1 = 0, ob1 = oW1 + kH1 * kF, oW2 = ob1 + kH1, ob2 = oW2 + kH2 * kH1,
              oW3 = ob2 + kH2, ob3 = oW3 + kC * kH2;

// Forward (+ backward when grad_part != nullptr) over samples idx[s0..s1).
__global__ __launch_bounds__(kNT) void k_mlp_fwd_bwd(const float* __restrict__ P,
                                                     const float* __restrict__ X,
                                                     const float* __restrict__ Tg,
                                                     const int* __restrict__ idx, int nsamp,
                                                     int per_block, float* __restrict__ grad_part,
                                                     float* __restrict__ loss_part,
                                                     float* __restrict__ logits_out) {
  __shared__ float s_x[kF], s_h1[kH1], s_h2[kH2], s_z[kC], s_dz[kC], s_d2[kH2], s_d1[kH1];
  __shared__ float s_loss;
  const int tid = threadIdx.x;
  float g[kGPT];
#pragma unroll
  for (int k = 0; k < kGPT; ++k) g[k] = 0.f;
  if (tid == 0) s_loss = 0.f;
  const int s0 = blockIdx.x * per_block;
  const int s1 = min(nsamp, s0 + per_block);
  const bool bwd = grad_part != nullptr;
  for (int si = s0; si < s1; ++si) {
    const int smp = idx ? idx[si] : si;
    if (tid < kF) s_x[tid] = X[(size_t)smp * kF + tid];
    __syncthreads();
    if (tid < kH1) {
      float a = P[ob1 + tid];
      for (int i = 0; i < kF; ++i) a += P[oW1 + tid * kF + i] * s_x[i];
      s_h1[tid] = a > 0.f ? a : 0.f;
    }
    __syncthreads();
    if (tid < kH2) {
      float a = P[ob2 + tid];
      for (int i = 0; i < kH1; ++i) a += P[oW2 + tid * kH1 + i] * s_h1[i];
      s_h2[tid] = a > 0.f ? a : 0.f;
    }
    __syncthreads();
    if (tid < kC) {
      float a = P[ob3 + tid];
      for (int i = 0; i < kH2; ++i) a += P[oW3 + tid * kH2 + i] * s_h2[i];
      s_z[tid] = a;
      if (logits_out) logits_out[(size_t)smp * kC + tid] = a;
    }
    __syncthreads();
    if (tid == 0 && Tg) {
      float zm = s_z[0];
      for (int c = 1; c < kC; ++c) zm = fmaxf(zm, s_z[c]);
      float se = 0.f;
      for (int c = 0; c < kC; ++c) se += expf(s_z[c] - zm);
      const float lse = zm + logf(se);
      float l = 0.f;
      for (int c = 0; c < kC; ++c) {
        const float t = Tg[(size_t)smp * kC + c];
        l += -t * (s_z[c] - lse);
        s_dz[c] = expf(s_z[c] - lse) - t;  // per-sample d loss / d z
      }
      s_loss += l;
    }
    __syncthreads();
    if (!bwd) continue;
    // d2 = W3^T dz masked by relu'(h2); d1 = W2^T d2 masked by relu'(h1)
    if (tid < kH2) {
      float a = 0.f;
      for (int c = 0; c < kC; ++c) a += P[oW3 + c * kH2 + tid] * s_dz[c];
      s_d2[tid] = s_h2[tid] > 0.f ? a : 0.f;
    }
    __syncthreads();
    if (tid < kH1) {
      float a = 0.f;
      for (int k = 0; k < kH2; ++k) a += P[oW2 + k * kH1 + tid] * s_d2[k];
      s_d1[tid] = s_h1[tid] > 0.f ? a : 0.f;
    }
    __syncthreads();
    // each thread accumulates its own parameter entries p = tid + k*NT
#pragma unroll
    for (int k = 0; k < kGPT; ++k) {
      const int p = tid + k * kNT;
      float v = 0.f;
      if (p < ob1) v = s_d1[p / kF] * s_x[p % kF];
      else if (p < oW2) v = s_d1[p - ob1];
      else if (p < ob2) v = s_d2[(p - oW2) / kH1] * s_h1[(p - oW2) % kH1];
      else if (p < oW3) v = s_d2[p - ob2];
      else if (p < ob3) v = s_dz[(p - oW3) / kH2] * s_h2[(p - oW3) % kH2];
      else if (p < kNParam) v = s_dz[p - ob3];
      g[k] += v;
    }
    __syncthreads();
  }
  if (bwd) {
#pragma unroll
    for (int k = 0; k < kGPT; ++k) {
      const int p = tid + k * kNT;
      if (p < kNParam) grad_part[(size_t)blockIdx.x * kNParam + p] = g[k];
    }
  }
  if (tid == 0 && loss_part) loss_part[blockIdx.x] = s_loss;
}

// grad = (sum of partial rows, fixed order) / nsamp; optional Adam step.
__global__ void k_mlp_reduce_adam(float* __restrict__ P, const float* __restrict__ grad_part,
                                  int nparts, int nsamp, float* __restrict__ grad_out,
                                  float* __restrict__ m1, float* __restrict__ m2, float lr,
                                  float b1c, float b2c, int step) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kNParam) return;
  float g = 0.f;
  for (int b = 0; b < nparts; ++b) g += grad_part[(size_t)b * kNParam + p];
  g /= (float)nsamp;
  if (grad_out) grad_out[p] = g;
  if (!m1 || step <= 0) return;
  const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  const float a = beta1 * m1[p] + (1.f - beta1) * g;
  const float v = beta2 * m2[p] + (1.f - beta2) * g * g;
  m1[p] = a;
  m2[p] = v;
  P[p] -= lr * (a / b1c) / (sqrtf(v / b2c) + eps);
}

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

int fail(const char* what, hipError_t e) {
  return set_error(nullptr, SPMSPV_DEVICE_ERROR, std::string(what) + ": " + hipGetErrorString(e));
}

#define MLP_TRY(what, expr)              \
  do {                                   \
    hipError_t e_ = (expr);              \
    if (e_ != hipSuccess) return fail(what, e_); \
  } while (0)

}  // namespace

extern "C" {

int spmspv_mlp_loss_grad(const float* params, const float* X, const float* targets, int B,
                         float* loss, float* grad) {
  if (!params || !X || !targets || B <= 0 || !loss || !grad)
    return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "mlp_loss_grad: bad arguments");
  const int per = 16;
  const int nb = (B + per - 1) / per;
  const size_t bP = sizeof(float) * kNParam, bX = sizeof(float) * (size_t)B * kF,
               bT = sizeof(float) * (size_t)B * kC, bG = sizeof(float) * (size_t)nb * kNParam;
  DevBuf buf;
  MLP_TRY("mlp alloc", hipMalloc(&buf.p, bP * 2 + bX + bT + bG + sizeof(float) * nb));
  char* q = static_cast<char*>(buf.p);
  float *dP = (float*)q, *dX = (float*)(q + bP), *dT = (float*)(q + bP + bX),
        *dG = (float*)(q + bP + bX + bT), *dL = (float*)(q + bP + bX + bT + bG),
        *dg = (float*)(q + bP + bX + bT + bG + sizeof(float) * nb);
  MLP_TRY("mlp h2d", hipMemcpy(dP, params, bP, hipMemcpyHostToDevice));
  MLP_TRY("mlp h2d", hipMemcpy(dX, X, bX, hipMemcpyHostToDevice));
  MLP_TRY("mlp h2d", hipMemcpy(dT, targets, bT, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(nb), dim3(kNT), 0, nullptr, dP, dX, dT, nullptr, B, per, dG,
                     dL, nullptr);
  hipLaunchKernelGGL(k_mlp_reduce_adam, dim3((kNParam + 255) / 256), dim3(256), 0, nullptr, dP, dG, nb,
                     B, dg, nullptr, nullptr, 0.f, 1.f, 1.f, 0);
  MLP_TRY("mlp launch", hipGetLastError());
  std::vector<float> lp(nb);
  MLP_TRY("mlp d2h", hipMemcpy(grad, dg, bP, hipMemcpyDeviceToHost));
  MLP_TRY("mlp d2h", hipMemcpy(lp.data(), dL, sizeof(float) * nb, hipMemcpyDeviceToHost));
  double l = 0;
  for (float v : lp) l += v;
  *loss = (float)(l / B);
  return SPMSPV_SUCCESS;
}

int spmspv_mlp_forward(const float* params, const float* X, int B, float* logits) {
  if (!params || !X || B <= 0 || !logits)
    return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "mlp_forward: bad arguments");
  const int per = 16;
  const int nb = (B + per - 1) / per;
  const size_t bP = sizeof(float) * kNParam, bX = sizeof(float) * (size_t)B * kF,
               bZ = sizeof(float) * (size_t)B * kC;
  DevBuf buf;
  MLP_TRY("mlp alloc", hipMalloc(&buf.p, bP + bX + bZ));
  char* q = static_cast<char*>(buf.p);
  float *dP = (float*)q, *dX = (float*)(q + bP), *dZ = (float*)(q + bP + bX);
  MLP_TRY("mlp h2d", hipMemcpy(dP, params, bP, hipMemcpyHostToDevice));
  MLP_TRY("mlp h2d", hipMemcpy(dX, X, bX, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(nb), dim3(kNT), 0, nullptr, dP, dX, nullptr, nullptr, B, per,
                     nullptr, nullptr, dZ);
  MLP_TRY("mlp launch", hipGetLastError());
  MLP_TRY("mlp d2h", hipMemcpy(logits, dZ, bZ, hipMemcpyDeviceToHost));
  return SPMSPV_SUCCESS;
}

int spmspv_mlp_train(const float* X, const float* targets, int B, int epochs, int batch, float lr,
                     unsigned seed, float* weights_out, float* loss) {
  if (!X || !targets || B <= 0 || epochs <= 0 || batch <= 0 || !weights_out)
    return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "mlp_train: bad arguments");
  // input normaliser from the training set (stored after the parameters)
  std::vector<float> mean(kF, 0.f), istd(kF, 1.f);
  for (int f = 0; f < kF; ++f) {
    double s = 0, s2 = 0;
    for (int i = 0; i < B; ++i) {
      s += X[(size_t)i * kF + f];
      s2 += (double)X[(size_t)i * kF + f] * X[(size_t)i * kF + f];
    }
    const double mu = s / B, var = std::max(0.0, s2 / B - mu * mu);
    mean[f] = (float)mu;
    istd[f] = (float)(1.0 / std::sqrt(var + 1e-6));
  }
  std::vector<float> Xn((size_t)B * kF);
  for (int i = 0; i < B; ++i)
    for (int f = 0; f < kF; ++f) Xn[(size_t)i * kF + f] = (X[(size_t)i * kF + f] - mean[f]) * istd[f];
  // He initialisation, seeded
  std::mt19937 rng(seed);
  std::vector<float> P(kNParam, 0.f);
  auto init = [&](int off, int fan_in, int count) {
    std::normal_distribution<float> nd(0.f, std::sqrt(2.f / fan_in));
    for (int i = 0; i < count; ++i) P[off + i] = nd(rng);
  };
  init(oW1, kF, kH1 * kF);
  init(oW2, kH1, kH2 * kH1);
  init(oW3, kH2, kC * kH2);
  const int per = 8;
  const int nbmax = (batch + per - 1) / per;
  const size_t bP = sizeof(float) * kNParam, bX = sizeof(float) * (size_t)B * kF,
               bT = sizeof(float) * (size_t)B * kC, bG = sizeof(float) * (size_t)nbmax * kNParam,
               bI = sizeof(int) * (size_t)B;
  DevBuf buf;
  MLP_TRY("mlp alloc", hipMalloc(&buf.p, bP * 3 + bX + bT + bG + bI + sizeof(float) * (nbmax + (B + 15) / 16)));
  char* q = static_cast<char*>(buf.p);
  float* dP = (float*)q;
  float* dM1 = (float*)(q + bP);
  float* dM2 = (float*)(q + 2 * bP);
  float* dX = (float*)(q + 3 * bP);
  float* dT = (float*)(q + 3 * bP + bX);
  float* dG = (float*)(q + 3 * bP + bX + bT);
  int* dI = (int*)(q + 3 * bP + bX + bT + bG);
  float* dL = (float*)(q + 3 * bP + bX + bT + bG + bI);
  MLP_TRY("mlp h2d", hipMemcpy(dP, P.data(), bP, hipMemcpyHostToDevice));
  MLP_TRY("mlp memset", hipMemset(dM1, 0, 2 * bP));
  MLP_TRY("mlp h2d", hipMemcpy(dX, Xn.data(), bX, hipMemcpyHostToDevice));
  MLP_TRY("mlp h2d", hipMemcpy(dT, targets, bT, hipMemcpyHostToDevice));
  std::vector<int> order(B);
  for (int i = 0; i < B; ++i) order[i] = i;
  int step = 0;
  double b1p = 1.0, b2p = 1.0;
  for (int ep = 0; ep < epochs; ++ep) {
    std::shuffle(order.begin(), order.end(), rng);
    MLP_TRY("mlp h2d", hipMemcpy(dI, order.data(), bI, hipMemcpyHostToDevice));
    for (int s0 = 0; s0 < B; s0 += batch) {
      const int bs = std::min(batch, B - s0);
      const int nb = (bs + per - 1) / per;
      ++step;
      b1p *= 0.9;
      b2p *= 0.999;
      hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(nb), dim3(kNT), 0, nullptr, dP, dX, dT, dI + s0, bs, per,
                         dG, nullptr, nullptr);
      hipLaunchKernelGGL(k_mlp_reduce_adam, dim3((kNParam + 255) / 256), dim3(256), 0, nullptr, dP, dG,
                         nb, bs, nullptr, dM1, dM2, lr, (float)(1.0 - b1p), (float)(1.0 - b2p), step);
    }
    MLP_TRY("mlp launch", hipGetLastError());
  }
  // final mean loss over the whole set
  const int nbl = (B + 15) / 16;
  hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(nbl), dim3(kNT), 0, nullptr, dP, dX, dT, nullptr, B, 16, nullptr,
                     dL, nullptr);
  MLP_TRY("mlp launch", hipGetLastError());
  std::vector<float> lp(nbl);
  MLP_TRY("mlp d2h", hipMemcpy(lp.data(), dL, sizeof(float) * nbl, hipMemcpyDeviceToHost));
  MLP_TRY("mlp d2h", hipMemcpy(weights_out, dP, bP, hipMemcpyDeviceToHost));
  double l = 0;
  for (float v : lp) l += v;
  if (loss) *loss = (float)(l / B);
  std::memcpy(weights_out + kNParam, mean.data(), sizeof(float) * kF);
  std::memcpy(weights_out + kNParam + kF, istd.data(), sizeof(float) * kF);
  return SPMSPV_SUCCESS;
}

}  // extern "C"
