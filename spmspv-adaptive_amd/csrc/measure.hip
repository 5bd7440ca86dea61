// Measured ceilings of the GPU the handle runs on, for the roofline of the
// kernels (no reference counterpart: the reference reports no bandwidth).
//   read:   streaming loads of a 512 MiB buffer, 16-byte nontemporal loads
//           with 8 in flight per lane (the streams every tile kernel reads);
//   copy:   a 256 MiB device-to-device copy, bytes read + written;
//   gather: 4-byte loads x[col[i]] with col uniformly random over the
//           handle's n columns (a SpMV's x gathers with no locality), per
//           second, 32 M gathers.
// Every timed launch follows a 512 MiB write to another buffer that evicts
// the 256 MiB MALL; the best of `reps` launches is kept.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "handle.hpp"

using namespace spmspv;

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

constexpr long long kBufBytes = 512ll << 20;
constexpr long long kGathers = 32ll << 20;

template <int U>
__global__ void k_read(const f4* __restrict__ a, long long n4, float* out) {
  float s = 0.f;
  const long long stride = (long long)gridDim.x * blockDim.x;
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + (U - 1) * stride < n4; i += U * stride) {
    f4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = __builtin_nontemporal_load(a + i + u * stride);
#pragma unroll
    for (int u = 0; u < U; ++u) s += v[u].x + v[u].y + v[u].z + v[u].w;
  }
  for (; i < n4; i += stride) {
    const f4 v = a[i];
    s += v.x + v.y + v.z + v.w;
  }
  if (s == 1234.5f) out[0] = s;  // keeps the loads; the buffer holds no such sum
}

__global__ void k_copy(const float4* __restrict__ a, float4* __restrict__ b, long long n4) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
       i += (long long)gridDim.x * blockDim.x)
    b[i] = a[i];
}

__device__ __forceinline__ unsigned mix32(unsigned long long z) {
  z = (z ^ (z >> 33)) * 0xff51afd7ed558ccdull;
  z = (z ^ (z >> 33)) * 0xc4ceb9fe1a85ec53ull;
  return (unsigned)(z ^ (z >> 33));
}

__global__ void k_rand_cols(int* col, long long n, int ncols) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    col[i] = (int)(mix32((unsigned long long)i * 0x9E3779B97F4A7C15ull + 1) % (unsigned)ncols);
}

template <int U>
__global__ void k_gather(const int* __restrict__ col, const float* __restrict__ x, long long n, float* out) {
  float s = 0.f;
  const long long stride = (long long)gridDim.x * blockDim.x;
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + (U - 1) * stride < n; i += U * stride) {
    int c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) c[u] = __builtin_nontemporal_load(col + i + u * stride);
#pragma unroll
    for (int u = 0; u < U; ++u) s += x[(unsigned)c[u]];
  }
  for (; i < n; i += stride) s += x[(unsigned)col[i]];
  if (s == 1234.5f) out[0] = s;
}

__global__ void k_fill_byte(unsigned* p, long long n, unsigned v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

}  // namespace

extern "C" {

int spmspv_measure_ceilings(spmspv_handle_t h, int reps, double out[3]) {
  if (!h || !out || reps < 1) return set_error(h, SPMSPV_INVALID_ARGUMENT, "measure: bad arguments");
  hipStream_t s = h->stream;
  int cus = 0;
  HIP_TRY(h, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
  char* buf = nullptr;
  // data (512 MiB) + the eviction buffer (512 MiB); freed on return
  HIP_TRY(h, hipMalloc(&buf, 2 * kBufBytes));
  float* data = reinterpret_cast<float*>(buf);
  unsigned* evict = reinterpret_cast<unsigned*>(buf + kBufBytes);
  const int grid = std::max(1, cus) * 8;
  auto timed = [&](auto launch, float* best) -> hipError_t {
    *best = 1e30f;
    for (int r = 0; r <= reps; ++r) {
      hipLaunchKernelGGL(k_fill_byte, dim3(grid), dim3(256), 0, s, evict, kBufBytes / 4, (unsigned)r);
      hipError_t e = hipEventRecord(h->ev0, s);
      if (e != hipSuccess) return e;
      launch();
      if ((e = hipEventRecord(h->ev1, s)) != hipSuccess) return e;
      if ((e = hipEventSynchronize(h->ev1)) != hipSuccess) return e;
      float ms = 0.f;
      if ((e = hipEventElapsedTime(&ms, h->ev0, h->ev1)) != hipSuccess) return e;
      if (r > 0) *best = std::min(*best, ms);  // the first launch warms up
    }
    return hipGetLastError();
  };
  hipError_t e = hipSuccess;
  float t_read = 0.f, t_copy = 0.f, t_gather = 0.f;
  hipLaunchKernelGGL(k_fill_byte, dim3(grid), dim3(256), 0, s, reinterpret_cast<unsigned*>(data), kBufBytes / 4,
                     0u);
  float* sink = reinterpret_cast<float*>(evict);
  e = timed([&] { hipLaunchKernelGGL((k_read<8>), dim3(grid), dim3(256), 0, s,
                                     reinterpret_cast<const f4*>(data), kBufBytes / 16, sink); },
            &t_read);
  if (e == hipSuccess)
    e = timed([&] { hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, s, reinterpret_cast<const float4*>(data),
                                       reinterpret_cast<float4*>(buf + kBufBytes / 2), (kBufBytes / 2) / 16); },
              &t_copy);
  // gathers: 32 M random columns (128 MiB) then x (n floats) in the data buffer
  int* col = reinterpret_cast<int*>(data);
  float* x = data + kGathers;
  const long long xmax = (kBufBytes / 4) - kGathers;
  const int ncols = (int)std::min<long long>(std::max(h->n, 1), xmax);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_rand_cols, dim3(grid), dim3(256), 0, s, col, kGathers, ncols);
    e = timed([&] { hipLaunchKernelGGL((k_gather<8>), dim3(grid), dim3(256), 0, s, col, x, kGathers, sink); },
              &t_gather);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(buf);
  if (e != hipSuccess) return set_error(h, SPMSPV_DEVICE_ERROR, std::string("measure: ") + hipGetErrorString(e));
  out[0] = (double)kBufBytes / (t_read * 1e-3) / 1e9;
  out[1] = (double)kBufBytes / (t_copy * 1e-3) / 1e9;  // half read, half written
  out[2] = (double)kGathers / (t_gather * 1e-3) / 1e9;
  return SPMSPV_SUCCESS;
}

}  // extern "C"
