// Measures the two ceilings that bound the row kernels on this GPU:
//  * streaming reads of (column, value) pairs (what every SpMV tile reads),
//    with and without several loads in flight per lane;
//  * 4-byte gathers x[col[j]] with uniformly random columns over footprints
//    from 64 KB to 64 MB (a matrix row's x gathers), and sequential ones.
// Each timed launch follows a 512 MB write that evicts the MALL; the best of
// four launches is reported. Usage: gather_ceiling (no arguments).
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include <cstdint>
#define CK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("HIP error %s at %d\n", hipGetErrorString(e), __LINE__); exit(1);} } while (0)

__global__ void k_stream(const int* __restrict__ c, const float* __restrict__ v, long long n, float* out) {
  float s = 0.f;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    s += v[i] * (float)c[i];
  if (s == 12345.f) out[0] = s;
}
template <int U>
__global__ void k_stream_u(const int* __restrict__ c, const float* __restrict__ v, long long n, float* out) {
  float s = 0.f;
  const long long stride = (long long)gridDim.x * blockDim.x;
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + (U - 1) * stride < n; i += U * stride) {
    int cc[U]; float vv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) { cc[u] = __builtin_nontemporal_load(c + i + u * stride); vv[u] = __builtin_nontemporal_load(v + i + u * stride); }
#pragma unroll
    for (int u = 0; u < U; ++u) s += vv[u] * (float)cc[u];
  }
  if (s == 12345.f) out[0] = s;
}
template <int U>
__global__ void k_gather(const int* __restrict__ c, const float* __restrict__ v, const float* __restrict__ x, long long n, float* out) {
  float s = 0.f;
  const long long stride = (long long)gridDim.x * blockDim.x;
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + (U - 1) * stride < n; i += U * stride) {
    int cc[U]; float vv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) { cc[u] = __builtin_nontemporal_load(c + i + u * stride); vv[u] = __builtin_nontemporal_load(v + i + u * stride); }
#pragma unroll
    for (int u = 0; u < U; ++u) s += vv[u] * x[(unsigned)cc[u]];
  }
  if (s == 12345.f) out[0] = s;
}
__global__ void k_copy(const float4* __restrict__ a, float4* __restrict__ b, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) b[i] = a[i];
}

static uint64_t sm(uint64_t& s) { uint64_t z = (s += 0x9e3779b97f4a7c15ull); z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull; z = (z ^ (z >> 27)) * 0x94d049bb133111ebull; return z ^ (z >> 31); }

int main() {
  const long long n = 32ll << 20;  // 32M entries: 128 MB col + 128 MB val
  int* c; float* v; float* x; float* out; float* flush;
  CK(hipMalloc(&c, n * 4)); CK(hipMalloc(&v, n * 4)); CK(hipMalloc(&x, 64ll << 20)); CK(hipMalloc(&out, 64)); CK(hipMalloc(&flush, 512ll << 20));
  std::vector<int> hc(n); std::vector<float> hv(n, 1.0f);
  CK(hipMemcpy(v, hv.data(), n * 4, hipMemcpyHostToDevice));
  CK(hipMemset(x, 0, 64ll << 20));
  hipEvent_t e0, e1; CK(hipEventCreate(&e0)); CK(hipEventCreate(&e1));
  int cus = 0; CK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
  auto timeit = [&](auto launch, const char* name, double bytes, double gathers) {
    float best = 1e30f;
    for (int r = 0; r < 5; ++r) {
      CK(hipMemsetAsync(flush, r, 512ll << 20, 0));
      CK(hipEventRecord(e0, 0)); launch(); CK(hipEventRecord(e1, 0)); CK(hipEventSynchronize(e1));
      float ms; CK(hipEventElapsedTime(&ms, e0, e1)); if (r && ms < best) best = ms;
    }
    printf("%-40s %9.1f us  %7.2f TB/s  %7.1f Ggath/s\n", name, best * 1e3, bytes / best / 1e9, gathers / best / 1e6);
    fflush(stdout);
  };
  // sequential col for stream
  for (long long i = 0; i < n; ++i) hc[i] = (int)(i & 1023);
  CK(hipMemcpy(c, hc.data(), n * 4, hipMemcpyHostToDevice));
  timeit([&] { k_copy<<<cus * 8, 256>>>((const float4*)flush, (float4*)(flush + (64ll << 20)), (256ll << 20) / 16); }, "copy 256MB->256MB", 512e6 * 1.048576, 0);
  for (int g : {1, 2, 4, 8, 16}) {
    char nm[64]; snprintf(nm, 64, "stream grid %d/CU", g);
    timeit([&] { k_stream<<<cus * g, 256>>>(c, v, n, out); }, nm, n * 8.0, 0);
  }
  for (int g : {2, 4, 8}) {
    char nm[64]; snprintf(nm, 64, "stream_u8 nt grid %d/CU", g);
    timeit([&] { k_stream_u<8><<<cus * g, 256>>>(c, v, n, out); }, nm, n * 8.0, 0);
  }
  const long long foots[] = {1 << 14, 1 << 16, 1 << 18, 1 << 20, 1 << 22, 1 << 24};
  for (long long f : foots) {
    uint64_t s = 42 + f;
    for (long long i = 0; i < n; ++i) hc[i] = (int)(sm(s) % (uint64_t)f);
    CK(hipMemcpy(c, hc.data(), n * 4, hipMemcpyHostToDevice));
    for (int g : {4, 8}) {
      char nm[64]; snprintf(nm, 64, "gather uniform %lld KB grid %d/CU", f * 4 / 1024, g);
      timeit([&] { k_gather<8><<<cus * g, 256>>>(c, v, x, n, out); }, nm, n * 8.0, (double)n);
    }
  }
  // sorted-within-64 (banded-like): col = i/16 (sequential)
  for (long long i = 0; i < n; ++i) hc[i] = (int)((i / 32) & ((1 << 20) - 1));
  CK(hipMemcpy(c, hc.data(), n * 4, hipMemcpyHostToDevice));
  timeit([&] { k_gather<8><<<cus * 8, 256>>>(c, v, x, n, out); }, "gather sequential (banded-like)", n * 8.0, (double)n);
  return 0;
}
