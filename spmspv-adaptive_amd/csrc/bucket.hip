// Bucket-based scatter/accumulate SpMSpV (slot 8). The reference's version
// (CscBasedBucketSpmspvDriver, la/include/spmspv/csc-spmspv/detail/device/
// csc_bucket_spmspv.cuh:766-1086) counts products per row bucket, scans,
// extracts into buckets and reduces each bucket; its reduction is compiled
// out (#if 0, :999-1071) and its atomic variant truncates values to int
// (:684). Here:
//   1. merge-path load-balanced gather writes every product (row, a*x);
//   2. one stable counted partition pass (per-tile digit counts, their
//      scan, scatter) on the top bucket_bits bits of the row groups the
//      products into 2^bucket_bits row buckets, 256 unless the caller or the
//      autotuner chose fewer (the count/scan/extract);
//   3. the bucketed products are cut into fixed chunks of 4,096 (so a
//      bucket holding most products -- R-MAT hub rows -- is spread over many
//      workgroups): a chunk sums its products per row in an LDS window over
//      its first bucket, then adds each row's sum to the accumulator once
//      (products outside the window go to it directly);
//   4. the compaction turns the accumulator's hit marks into the sorted
//      result and clears both.
#include "handle.hpp"

namespace spmspv {

void launch_partition_pass(hipStream_t s, const SortBufs& b, int shift, const long long* n_dev, Ctl* ctl,
                           long long n_hint, long long n_cap, Ep ep);

constexpr int kWinBits = 13;  // LDS window of at most 8192 rows (32 KB of floats)
constexpr int kChunkNT = 256, kChunkIPT = 16, kChunk = kChunkNT * kChunkIPT;

template <int NT, int IPT>
__global__ __launch_bounds__(NT) void k_bucket_chunk(const unsigned* __restrict__ keys,
                                                     const float* __restrict__ vals,
                                                     const long long* __restrict__ n_dev, int wbits,
                                                     float* acc, unsigned char* marks) {
  constexpr int CH = NT * IPT;
  constexpr int W = 1 << kWinBits;
  __shared__ float s_acc[W];
  __shared__ unsigned s_hit[W / 32];
  __shared__ unsigned s_base;
  const long long n = *n_dev;
  const unsigned win = 1u << wbits;
  const int nwords = (int)((win + 31) / 32);
  for (int k = threadIdx.x; k < (int)win; k += NT) s_acc[k] = 0.f;
  for (int k = threadIdx.x; k < nwords; k += NT) s_hit[k] = 0u;
  const int w = threadIdx.x / kWave;
  const long long nchunks = (n + CH - 1) / CH;
  for (long long c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const long long base = c * CH;
    const long long b0 = base + (long long)w * kWave * IPT + lane();
    unsigned key[IPT];
    float v[IPT];
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const long long i = b0 + (long long)j * kWave;
      const long long ic = i < n ? i : n - 1;
      key[j] = keys[ic];
      v[j] = vals[ic];
    }
    // the chunk's first item (thread 0, item 0) lies in its lowest bucket
    if (threadIdx.x == 0) s_base = (key[0] >> wbits) << wbits;
    __syncthreads();
    const unsigned wb = s_base;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      if (b0 + (long long)j * kWave < n) {
        const unsigned r = key[j];
        const unsigned lr = r - wb;
        if (lr < win) {
          atomicAdd(&s_acc[lr], v[j]);
          atomicOr(&s_hit[lr >> 5], 1u << (lr & 31));
        } else {
          atomic_add_f32(acc + r, v[j]);
          marks[r] = 1;
        }
      }
    }
    __syncthreads();
    // one accumulator update per row of the window; the window is left
    // zeroed for the next chunk
    for (int k = threadIdx.x; k < nwords; k += NT) {
      unsigned bits = s_hit[k];
      if (!bits) continue;
      s_hit[k] = 0u;
      while (bits) {
        const int bit = __ffs(bits) - 1;
        bits &= bits - 1;
        const unsigned lr = ((unsigned)k << 5) + bit;
        atomic_add_f32(acc + wb + lr, s_acc[lr]);
        marks[wb + lr] = 1;
        s_acc[lr] = 0.f;
      }
    }
    __syncthreads();
  }
}

// Bucket geometry for m rows: 2^bucket_bits buckets (1..8 bits, the
// partition pass's one 8-bit digit) over the row range, windows of at most
// 8192 rows inside each bucket.
static void bucket_geometry(int m, int bucket_bits, int* bshift, int* wbits) {
  int bits = 1;
  while (bits < 31 && (1u << bits) < (unsigned)m) ++bits;
  *bshift = bits > bucket_bits ? bits - bucket_bits : 0;
  *wbits = std::min(*bshift, kWinBits);
}

void launch_bucket(hipStream_t s, const SortBufs& b, int m, int bucket_bits, Ctl* ctl, float* acc,
                   unsigned char* marks, long long binlen_hint, long long n_cap, Ep ep) {
  int bshift, wbits;
  bucket_geometry(m, bucket_bits, &bshift, &wbits);
  const long long* n_dev = &ctl->bin_len;
  launch_partition_pass(s, b, bshift, n_dev, ctl, binlen_hint, n_cap, ep);
  const long long n = binlen_hint >= 0 ? binlen_hint : n_cap;
  const int grid = (int)std::max<long long>(1, std::min<long long>(2048, (n + kChunk - 1) / kChunk));
  hipLaunchKernelGGL((k_bucket_chunk<kChunkNT, kChunkIPT>), dim3(grid), dim3(kChunkNT), 0, s, b.keyB,
                     b.valB, n_dev, wbits, acc, marks);
}

}  // namespace spmspv
