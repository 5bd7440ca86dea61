// Handle-free column SpMSpV that allocates nothing: the load-balanced
// atomic column slot (the kernels of spmspv_run with
// SPMSPV_K_COL_LB_ATOMIC) with every scratch buffer carved from one
// file-scope device array instead of a handle's arena. For callers that
// cannot call hipMalloc (code under graph capture, serving loops).
//
// Per call, on the caller's stream:
//   1. a fill kernel zeroes the part of the scratch that must start at zero
//      (control block, look-back words, accumulator, row marks), whatever
//      an earlier call left there;
//   2. k_prep: column-length scan of x's support and merge-path splits;
//   3. k_lb_gather: products added into the accumulator, rows marked;
//   4. k_compact: marks -> sorted (key, value), accumulator cleared;
//   5. the result count copied to the caller's device word.
// Calls that overlap on one GPU share its scratch array and race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "handle.hpp"

using namespace spmspv;

namespace {

constexpr size_t kScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

size_t up256(size_t v) { return (v + 255) & ~size_t(255); }

// Byte offsets of every buffer inside the scratch; the ones that must start
// at zero come first, so one fill of [0, zero_bytes) clears them.
struct Layout {
  size_t ctl, state, tile_mm, acc, marks, zero_bytes, ptr, xbeg, split, total;
};

Layout layout(int m, int xnnz, long long nnz) {
  Layout L{};
  const size_t prep_tiles = ((size_t)xnnz + kPrepTile - 1) / kPrepTile + 2;
  const size_t mwords = ((size_t)m + 31) / 32;
  const size_t cmp_tiles = (mwords + kCmpNT - 1) / kCmpNT + 2;
  const size_t lb_tiles = ((size_t)xnnz + (size_t)nnz + kLbTile - 1) / kLbTile + 2;
  size_t off = 0;
  L.ctl = off;
  off = up256(off + sizeof(Ctl));
  L.state = off;
  off = up256(off + sizeof(unsigned long long) * std::max(prep_tiles, cmp_tiles));
  L.tile_mm = off;
  off = up256(off + sizeof(int) * 2 * prep_tiles);
  L.acc = off;
  off = up256(off + sizeof(float) * (size_t)m);
  L.marks = off;
  off = up256(off + mwords * 32 + 64);
  L.zero_bytes = off;
  L.ptr = off;
  off = up256(off + sizeof(int) * ((size_t)xnnz + 1));
  L.xbeg = off;
  off = up256(off + sizeof(int) * std::max<size_t>(xnnz, 1));
  L.split = off;
  off = up256(off + sizeof(int) * lb_tiles);
  L.total = off;
  return L;
}

// Device address of this GPU's scratch array, looked up once per device
// with that device current.
int scratch_of_current_device(char** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess)
    return set_error(nullptr, SPMSPV_DEVICE_ERROR, std::string("noalloc: hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices)
    return set_error(nullptr, SPMSPV_DEVICE_ERROR, "noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(nullptr, SPMSPV_DEVICE_ERROR,
                       std::string("noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[dev] = p;
  }
  *out = static_cast<char*>(g_scratch[dev]);
  return SPMSPV_SUCCESS;
}

__global__ void k_copy_count(const Ctl* ctl, int* dst) { *dst = ctl->ynnz; }

}  // namespace

extern "C" {

int spmspv_csc_spmspv_noalloc(int m, int n, long long nnz, const int* d_col_ptr, const int* d_row,
                              const float* d_val, const int* d_x_key, const float* d_x_val, int xnnz,
                              int* d_y_key, float* d_y_val, int* d_y_nnz, void* stream) {
  if (m <= 0 || n <= 0 || nnz < 0 || nnz > 0x7fffffffLL || xnnz < 0 || xnnz > n)
    return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "noalloc: bad sizes");
  if (!d_col_ptr || !d_y_key || !d_y_val || !d_y_nnz || (nnz > 0 && (!d_row || !d_val)) ||
      (xnnz > 0 && (!d_x_key || !d_x_val)))
    return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "noalloc: null argument");
  const Layout L = layout(m, xnnz, nnz);
  if (L.total > kScratchBytes)
    return set_error(nullptr, SPMSPV_INVALID_ARGUMENT,
                     "noalloc: needs " + std::to_string(L.total) + " scratch bytes, the maximum is " +
                         std::to_string(kScratchBytes) + " (m, xnnz or nnz too large)");
  char* base = nullptr;
  int rc = scratch_of_current_device(&base);
  if (rc) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  Ctl* ctl = reinterpret_cast<Ctl*>(base + L.ctl);
  auto* state = reinterpret_cast<unsigned long long*>(base + L.state);
  int* tile_mm = reinterpret_cast<int*>(base + L.tile_mm);
  float* acc = reinterpret_cast<float*>(base + L.acc);
  auto* marks = reinterpret_cast<unsigned char*>(base + L.marks);
  int* ptr = reinterpret_cast<int*>(base + L.ptr);
  int* xbeg = reinterpret_cast<int*>(base + L.xbeg);
  int* split = reinterpret_cast<int*>(base + L.split);
  // a fill kernel, not hipMemsetAsync: the call may be captured in a graph
  launch_fill_i32(s, reinterpret_cast<int*>(base), (long long)(L.zero_bytes / sizeof(int)), 0);
  // look-back words were just cleared, so fixed epochs never meet a stale word
  launch_prep(s, d_col_ptr, d_x_key, n, xnnz, nullptr, ptr, xbeg, split, state, tile_mm, ctl, nullptr,
              Ep{1u, -1});
  launch_lb_gather(s, 0, ptr, xbeg, d_x_val, split, d_row, d_val, ctl, xnnz, nullptr, acc, marks, nullptr,
                   nullptr, -1);
  launch_compact(s, nullptr, marks, m, acc, 1, d_y_key, d_y_val, state, ctl, Ep{2u, -1});
  hipLaunchKernelGGL(k_copy_count, dim3(1), dim3(1), 0, s, ctl, d_y_nnz);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return set_error(nullptr, SPMSPV_DEVICE_ERROR, std::string("noalloc: launch: ") + hipGetErrorString(e));
  return SPMSPV_SUCCESS;
}

}  // extern "C"
