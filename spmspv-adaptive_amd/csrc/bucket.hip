// Bucket-based scatter/accumulate SpMSpV (slot 8). The reference's version
// (CscBasedBucketSpmspvDriver, la/include/spmspv/csc-spmspv/detail/device/
// csc_bucket_spmspv.cuh:766-1086) counts products per row bucket, scans,
// extracts into buckets and reduces each bucket; its reduction is compiled
// out (#if 0, :999-1071) and its atomic variant truncates values to int
// (:684). Here:
//   1. merge-path load-balanced gather writes every product (row, a*x);
//   2. one stable counted partition pass (per-tile digit counts, their
//      scan, scatter) on the top 8 bits of the row groups the products into
//      256 row buckets (the count/scan/extract);
//   3. one workgroup per (bucket, row window of 8192 rows) accumulates its
//      products into an LDS row window with LDS float atomics and a hit
//      bit array, then writes the window's structural support in row order
//      at the position found by decoupled look-back over windows -- so the
//      output is sorted without a full sort.
#include "handle.hpp"

namespace spmspv {

void launch_partition_pass(hipStream_t s, const SortBufs& b, int shift, const long long* n_dev, Ctl* ctl,
                           long long n_hint, long long n_cap, Ep ep);

constexpr int kWinBits = 13;  // 8192 rows per LDS window (32 KB of floats)

template <int NT>
__global__ __launch_bounds__(NT) void k_bucket_acc(const unsigned* __restrict__ keys,
                                                   const float* __restrict__ vals,
                                                   const unsigned* __restrict__ dstart,
                                                   const long long* __restrict__ n_dev, int m,
                                                   int bshift, int wbits, Ctl* ctl,
                                                   unsigned long long* state, Ep epa,
                                                   int* __restrict__ ykey, float* __restrict__ yval) {
  constexpr int W = 1 << kWinBits;
  constexpr int WW = W / 32;
  __shared__ float s_acc[W];
  __shared__ unsigned s_hit[WW];
  __shared__ unsigned long long s_pre;
  __shared__ int s_i[NT / kWave];
  const int nbuckets = ((m - 1) >> bshift) + 1;
  const int wpb = 1 << (bshift - wbits);  // windows per bucket
  const unsigned ntiles = (unsigned)(nbuckets * wpb);
  const int win = 1 << wbits;
  const unsigned t = blockIdx.x;
  if (t >= ntiles) return;
  const unsigned epoch = ep_of(ctl, epa);
  {
    const int b = (int)(t / wpb), w = (int)(t % wpb);
    const int r0 = (b << bshift) + (w << wbits);
    for (int k = threadIdx.x; k < win; k += NT) s_acc[k] = 0.f;
    for (int k = threadIdx.x; k < (win + 31) / 32; k += NT) s_hit[k] = 0u;
    __syncthreads();
    // bucket b's items: the partition pass left the start of digit d at
    // dstart[d * stiles] (stiles = its tiles)
    const long long n = *n_dev;
    const long long stiles = (n + kSortTile - 1) / kSortTile;
    const long long i0 = n > 0 ? (long long)dstart[(size_t)b * stiles] : 0;
    const long long i1 = n > 0 && b + 1 < 256 ? (long long)dstart[(size_t)(b + 1) * stiles] : n;
    for (long long i = i0 + threadIdx.x; i < i1; i += NT) {
      const int r = (int)keys[i] - r0;
      if ((unsigned)r < (unsigned)win) {
        atomicAdd(&s_acc[r], vals[i]);
        atomicOr(&s_hit[r >> 5], 1u << (r & 31));
      }
    }
    __syncthreads();
    // compaction of the window: words blocked per thread
    constexpr int WPT = (WW + NT - 1) / NT;
    const int nw = (win + 31) / 32;
    int cnt = 0;
    unsigned wd[WPT];
#pragma unroll
    for (int j = 0; j < WPT; ++j) {
      const int k = threadIdx.x * WPT + j;
      wd[j] = k < nw ? s_hit[k] : 0u;
      cnt += __popc(wd[j]);
    }
    int btot;
    const int bex = block_excl_sum<NT, int>(cnt, s_i, &btot);
    const unsigned long long pre = tile_prefix(state, t, epoch, (unsigned long long)btot, &s_pre);
    int pos = (int)pre + bex;
#pragma unroll
    for (int j = 0; j < WPT; ++j) {
      unsigned bits = wd[j];
      const int k = threadIdx.x * WPT + j;
      while (bits) {
        const int bit = __ffs(bits) - 1;
        bits &= bits - 1;
        const int lr = (k << 5) + bit;
        ykey[pos] = r0 + lr;
        yval[pos] = s_acc[lr];
        ++pos;
      }
    }
    if (t == ntiles - 1 && threadIdx.x == 0) ctl->ynnz = (int)(pre + btot);
  }
}

// Bucket geometry for m rows: 256 buckets over the row range, windows of
// at most 8192 rows inside each bucket.
static void bucket_geometry(int m, int* bshift, int* wbits) {
  int bits = 1;
  while (bits < 31 && (1u << bits) < (unsigned)m) ++bits;
  *bshift = bits > 8 ? bits - 8 : 0;
  *wbits = std::min(*bshift, kWinBits);
}

void launch_bucket(hipStream_t s, const SortBufs& b, int m, Ctl* ctl, unsigned long long* state,
                   int* ykey, float* yval, long long binlen_hint, long long n_cap, Ep ep) {
  int bshift, wbits;
  bucket_geometry(m, &bshift, &wbits);
  const long long* n_dev = &ctl->bin_len;
  launch_partition_pass(s, b, bshift, n_dev, ctl, binlen_hint, n_cap, ep);
  const int ntiles = (((m - 1) >> bshift) + 1) << (bshift - wbits);
  const Ep e2{ep.v + 1u, ep.dev < 0 ? -1 : ep.dev + 1};
  hipLaunchKernelGGL(k_bucket_acc<256>, dim3(ntiles), dim3(256), 0, s, b.keyB, b.valB, b.dcount, n_dev, m,
                     bshift, wbits, ctl, state, e2, ykey, yval);
}

}  // namespace spmspv
