// Row-bin scatter/accumulate SpMSpV (slot 8). The reference's bucket method
// (CscBasedBucketSpmspvDriver, la/include/spmspv/csc-spmspv/detail/device/
// csc_bucket_spmspv.cuh:766-1086) counts the products per row bucket, scans
// the counts, extracts the products into their buckets and reduces each
// bucket; its reduction is compiled out (#if 0, :999-1071) and its atomic
// variant truncates values to int (:684). Here the buckets are row bins of
// at most 2^15 rows, so that one bin's FP32 sums fit a 128 KB LDS window:
//   1. k_bin_count: the load-balanced gather's merge-path tiles (one at a
//      time per workgroup, grid-stride over at most 512 workgroups) count
//      their products per bin in LDS, then one global atomic per (workgroup,
//      bin) adds the counts;
//   2. k_bin_plan (one workgroup): each bin's start (exclusive scan of the
//      counts) and write cursor, and its accumulate chunks: a bin is cut
//      into equal chunks of about kBinChunk products, at most kBinChunksMax
//      (so a hub bin's partial sums meet in at most that many atomics per
//      row); the counts are cleared for the next call;
//   3. k_bin_scatter: the same tiles, two per workgroup round, form every
//      product (row, a*x) again and write it into its bin, at positions the
//      round reserves per bin with one global atomic;
//   4. k_bin_accum: each chunk is summed per row in an LDS window over its
//      bin's rows (LDS FP32 atomics, the window sized to the bin); a bin
//      that is a single chunk stores its row sums into the accumulator, a
//      longer bin adds them (one FP32 atomic per row and chunk); hit rows
//      are marked;
//   5. the caller's compaction turns the marks into the sorted result and
//      clears the accumulator and the marks.
// HBM bytes per product: 4 (count) + 8 read and 8 written (scatter) + 8
// (accumulate), against 44 and up for gather, radix passes and reduce.
#include <mutex>

#include "handle.hpp"

namespace spmspv {

namespace {

constexpr int kBinTPW = 2;          // merge-path tiles per scatter workgroup round
constexpr int kBinChunksMax = 16;   // accumulate chunks per bin at most
constexpr int kAccNT = 1024;        // accumulate workgroup
constexpr int kAccPPT = 16;         // products per thread per sub-round

// The products of merge-path tile t of the load-balanced gather (k_prep's
// split; kLbTile merged positions: x-entry markers and products). For the
// thread's k-th item q = threadIdx.x + k * NT < np: idx[k] = its CSC index,
// xs[k] = its x value (0 without xval). Returns np; b0 gets the index of the
// tile's first product. The LDS tables are the caller's; the function's two
// barriers order their use, and the caller syncs before the next tile.
template <int NT, int VT>
__device__ __forceinline__ int lb_tile(long long t, const int* __restrict__ ptr, const int* __restrict__ xbeg,
                                       const float* __restrict__ xval, const int* __restrict__ split, int xnnz,
                                       long long W, int* s_ptr, int* s_beg, float* s_xv, int* s_seg,
                                       int (&idx)[VT], float (&xs)[VT], long long* b0out) {
  constexpr int T = NT * VT;
  const long long d0 = t * T;
  const long long d1 = min(d0 + T, W);
  const int a0 = split[t];
  const int a1 = split[t + 1];
  const long long b0 = d0 - a0;
  const long long b1 = d1 - a1;
  const int sa = a0 > 0 ? a0 - 1 : 0;
  const int nseg = a1 - sa + 1;  // <= T + 2
  {
    constexpr int SEGPT = (T + 2 + NT - 1) / NT;
    int pv[SEGPT], bv[SEGPT];
    float xv[SEGPT];
#pragma unroll
    for (int u = 0; u < SEGPT; ++u) {
      const int k = threadIdx.x + u * NT;
      const int i = sa + (k < nseg ? k : 0);
      const int ic = i < xnnz ? i : (xnnz > 0 ? xnnz - 1 : 0);
      pv[u] = ptr[i];
      bv[u] = xbeg[ic];
      xv[u] = xval ? xval[ic] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < SEGPT; ++u) {
      const int k = threadIdx.x + u * NT;
      if (k < nseg) {
        s_ptr[k] = pv[u];
        s_beg[k] = bv[u];
        s_xv[k] = xv[u];
      }
    }
  }
  __syncthreads();
  long long D = d0 + (long long)threadIdx.x * VT;
  if (D < d1) {
    int lo = a0, hi = a1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if ((long long)mid + s_ptr[mid - sa] < D) lo = mid + 1;
      else hi = mid;
    }
    int i = lo;
    long long p = D - i;
#pragma unroll
    for (int k = 0; k < VT; ++k) {
      if (D >= d1) break;
      if (i < a1 && (long long)i + s_ptr[i - sa] == D) {
        ++i;
      } else {
        s_seg[p - b0] = i - 1 - sa;
        ++p;
      }
      ++D;
    }
  }
  __syncthreads();
  const int np = (int)(b1 - b0);
#pragma unroll
  for (int k = 0; k < VT; ++k) {
    const int q0 = threadIdx.x + k * NT;
    const bool ok = q0 < np;
    const int sg = ok ? s_seg[q0] : 0;
    idx[k] = ok ? s_beg[sg] + (int)(b0 + q0 - s_ptr[sg]) : 0;
    xs[k] = ok ? s_xv[sg] : 0.f;
  }
  *b0out = b0;
  return np;
}

template <int NT, int VT>
__global__ __launch_bounds__(NT) void k_bin_count(const int* __restrict__ ptr, const int* __restrict__ xbeg,
                                                  const int* __restrict__ split, const int* __restrict__ crow,
                                                  const Ctl* __restrict__ ctl, int xnnz, int rbits, int nbins,
                                                  unsigned* __restrict__ hist) {
  constexpr int T = NT * VT;
  __shared__ int s_ptr[T + 2];
  __shared__ int s_beg[T + 2];
  __shared__ float s_xv[T + 2];
  __shared__ int s_seg[T];
  __shared__ unsigned s_h[kBinMax];
  for (int b = threadIdx.x; b < nbins; b += NT) s_h[b] = 0u;
  const long long ntiles = ctl->n_tiles;
  const long long W = (long long)xnnz + ctl->bin_len;
  for (long long t = blockIdx.x; t < ntiles; t += gridDim.x) {
    int idx[VT];
    float xs[VT];
    long long b0;
    const int np = lb_tile<NT, VT>(t, ptr, xbeg, nullptr, split, xnnz, W, s_ptr, s_beg, s_xv, s_seg, idx, xs, &b0);
    if (np > 0) {
      int r[VT];
#pragma unroll
      for (int k = 0; k < VT; ++k) r[k] = crow[idx[k]];
#pragma unroll
      for (int k = 0; k < VT; ++k)
        if ((int)threadIdx.x + k * NT < np) atomicAdd(&s_h[(unsigned)r[k] >> rbits], 1u);
    }
    __syncthreads();  // the tables are reused by the next tile
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nbins; b += NT) {
    const unsigned c = s_h[b];
    if (c) atomicAdd(hist + b, c);
  }
}

// One workgroup: starts, cursors and accumulate chunks of every bin from
// the counts, which it clears.
template <int NT>
__global__ __launch_bounds__(NT) void k_bin_plan(unsigned* __restrict__ hist, int nbins, BinWs ws) {
  constexpr int BPT = kBinMax / NT;
  __shared__ long long s_l[NT / kWave];
  __shared__ int s_i[NT / kWave];
  unsigned c[BPT];
  int nch[BPT];
  long long csum = 0;
  int nsum = 0;
#pragma unroll
  for (int j = 0; j < BPT; ++j) {
    const int b = threadIdx.x * BPT + j;
    c[j] = b < nbins ? hist[b] : 0u;
    nch[j] = (int)std::min<unsigned>((c[j] + kBinChunk - 1) / kBinChunk, (unsigned)kBinChunksMax);
    csum += c[j];
    nsum += nch[j];
  }
  long long ctot;
  int ntot;
  long long cpre = block_excl_sum<NT, long long>(csum, s_l, &ctot);
  int npre = block_excl_sum<NT, int>(nsum, s_i, &ntot);
#pragma unroll
  for (int j = 0; j < BPT; ++j) {
    const int b = threadIdx.x * BPT + j;
    if (b < nbins) {
      ws.bstart[b] = cpre;
      ws.cursor[b] = (unsigned long long)cpre;
      ws.cstart[b] = npre;
      hist[b] = 0u;
    }
    cpre += c[j];
    npre += nch[j];
  }
  if (threadIdx.x == 0) {
    ws.bstart[nbins] = ctot;
    ws.cstart[nbins] = ntot;
  }
}

template <int NT, int VT>
__global__ __launch_bounds__(NT) void k_bin_scatter(const int* __restrict__ ptr, const int* __restrict__ xbeg,
                                                    const float* __restrict__ xval, const int* __restrict__ split,
                                                    const int* __restrict__ crow, const float* __restrict__ cval,
                                                    const Ctl* __restrict__ ctl, int xnnz, int rbits, int nbins,
                                                    unsigned long long* __restrict__ cursor,
                                                    unsigned* __restrict__ okey, float* __restrict__ oval) {
  constexpr int T = NT * VT;
  __shared__ int s_ptr[T + 2];
  __shared__ int s_beg[T + 2];
  __shared__ float s_xv[T + 2];
  __shared__ int s_seg[T];
  __shared__ unsigned s_c[kBinMax];
  __shared__ unsigned long long s_base[kBinMax];
  const long long ntiles = ctl->n_tiles;
  const long long W = (long long)xnnz + ctl->bin_len;
  for (long long g = blockIdx.x; g * kBinTPW < ntiles; g += gridDim.x) {
    for (int b = threadIdx.x; b < nbins; b += NT) s_c[b] = 0u;
    __syncthreads();
    // every product of the round in registers, with its rank in its bin
    unsigned row[kBinTPW][VT];
    float val[kBinTPW][VT];
    unsigned lo[kBinTPW][VT];  // ~0u: no product
#pragma unroll
    for (int u = 0; u < kBinTPW; ++u) {
      const long long t = g * kBinTPW + u;
#pragma unroll
      for (int k = 0; k < VT; ++k) lo[u][k] = ~0u;
      if (t < ntiles) {
        int idx[VT];
        float xs[VT];
        long long b0;
        const int np = lb_tile<NT, VT>(t, ptr, xbeg, xval, split, xnnz, W, s_ptr, s_beg, s_xv, s_seg, idx, xs, &b0);
        if (np > 0) {
          float a[VT];
#pragma unroll
          for (int k = 0; k < VT; ++k) {
            row[u][k] = (unsigned)crow[idx[k]];
            a[k] = cval[idx[k]];
          }
#pragma unroll
          for (int k = 0; k < VT; ++k) {
            val[u][k] = a[k] * xs[k];
            if ((int)threadIdx.x + k * NT < np) lo[u][k] = atomicAdd(&s_c[row[u][k] >> rbits], 1u);
          }
        }
      }
      __syncthreads();  // the tables are reused by the next tile
    }
    for (int b = threadIdx.x; b < nbins; b += NT) {
      const unsigned c = s_c[b];
      if (c) s_base[b] = atomicAdd(cursor + b, (unsigned long long)c);
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kBinTPW; ++u) {
#pragma unroll
      for (int k = 0; k < VT; ++k) {
        if (lo[u][k] != ~0u) {
          const unsigned long long p = s_base[row[u][k] >> rbits] + lo[u][k];
          okey[p] = row[u][k];
          oval[p] = val[u][k];
        }
      }
    }
    __syncthreads();  // s_c and s_base are reused by the next round
  }
}

template <int NT, int PPT>
__global__ __launch_bounds__(NT) void k_bin_accum(const unsigned* __restrict__ key, const float* __restrict__ val,
                                                  BinWs ws, int nbins, int rbits, float* acc,
                                                  unsigned char* marks) {
  constexpr int SUB = NT * PPT;
  constexpr int BPT = kBinMax / NT;
  // the window: 2^rbits FP32 sums, then their hit bits
  extern __shared__ float s_dyn[];
  float* s_acc = s_dyn;
  const int R = 1 << rbits;
  const int RW = (R + 31) / 32;
  unsigned* s_hit = reinterpret_cast<unsigned*>(s_dyn + R);
  __shared__ int s_b;
  for (int r = threadIdx.x; r < R; r += NT) s_acc[r] = 0.f;
  for (int w = threadIdx.x; w < RW; w += NT) s_hit[w] = 0u;
  const int nch = ws.cstart[nbins];
  for (int c = blockIdx.x; c < nch; c += gridDim.x) {
    // the bin holding chunk c: the b with cstart[b] <= c < cstart[b + 1]
#pragma unroll
    for (int j = 0; j < BPT; ++j) {
      const int b = threadIdx.x * BPT + j;
      if (b < nbins && ws.cstart[b] <= c && c < ws.cstart[b + 1]) s_b = b;
    }
    __syncthreads();
    const int b = s_b;
    const int c0 = ws.cstart[b], nb = ws.cstart[b + 1] - c0;
    const bool owned = nb == 1;
    // chunk c - c0 of nb equal parts of the bin
    const long long bs = ws.bstart[b], size = ws.bstart[b + 1] - bs;
    const long long p0 = bs + size * (c - c0) / nb;
    const long long pe = bs + size * (c - c0 + 1) / nb;
    const unsigned rb = (unsigned)b << rbits;
    for (long long q0 = p0; q0 < pe; q0 += SUB) {
      unsigned kk[PPT];
      float vv[PPT];
#pragma unroll
      for (int j = 0; j < PPT; ++j) {
        const long long i = q0 + threadIdx.x + (long long)j * NT;
        const long long ic = i < pe ? i : p0;
        kk[j] = __builtin_nontemporal_load(key + ic);
        vv[j] = __builtin_nontemporal_load(val + ic);
      }
#pragma unroll
      for (int j = 0; j < PPT; ++j) {
        if (q0 + threadIdx.x + (long long)j * NT < pe) {
          const unsigned r = kk[j] - rb;
          atomicAdd(&s_acc[r], vv[j]);
          atomicOr(&s_hit[r >> 5], 1u << (r & 31));
        }
      }
    }
    __syncthreads();
    // the window's rows into the accumulator; the window is left zeroed
    for (int w = threadIdx.x; w < RW; w += NT) {
      unsigned bits = s_hit[w];
      if (!bits) continue;
      s_hit[w] = 0u;
      while (bits) {
        const int bit = __ffs(bits) - 1;
        bits &= bits - 1;
        const unsigned lr = ((unsigned)w << 5) + bit;
        const float v = s_acc[lr];
        s_acc[lr] = 0.f;
        if (owned) acc[rb + lr] = v;  // the accumulator holds 0 there
        else atomic_add_f32(acc + rb + lr, v);
        marks[rb + lr] = 1;
      }
    }
    __syncthreads();
  }
}

}  // namespace

// Bin geometry for m rows: at least 2^bucket_bits bins and at most 2^15 rows
// per bin; *rbits = log2(rows per bin). Returns the number of bins, or -1
// when m needs more than kBinMax bins.
int bin_geometry(int m, int bucket_bits, int* rbits) {
  int bits = 0;
  while (bits < 31 && (1ll << bits) < (long long)m) ++bits;
  int rb = bits > bucket_bits ? bits - bucket_bits : 0;
  if (rb > kBinRowBits) rb = kBinRowBits;
  const long long nb = m > 0 ? (((long long)m - 1) >> rb) + 1 : 1;
  *rbits = rb;
  return nb > kBinMax ? -1 : (int)nb;
}

bool launch_bucket(hipStream_t s, const int* ptr, const int* xbeg, const float* xval, const int* split,
                   const int* crow, const float* cval, const Ctl* ctl, int xnnz, int m, int bucket_bits,
                   const BinWs& ws, unsigned* bkey, float* bval, float* acc, unsigned char* marks,
                   long long tiles_hint, long long binlen_hint, long long binlen_cap) {
  int rbits = 0;
  const int nbins = bin_geometry(m, bucket_bits, &rbits);
  if (nbins < 0) return false;
  const long long tiles = tiles_hint > 0 ? tiles_hint : ((long long)xnnz + binlen_cap) / kLbTile + 2;
  // a few tiles per workgroup: each workgroup ends with one atomic per bin
  // it counted, on addresses every workgroup shares
  const int cgrid = (int)std::max<long long>(1, std::min<long long>(512, tiles));
  const int rgrid = (int)std::max<long long>(1, std::min<long long>(2048, (tiles + kBinTPW - 1) / kBinTPW));
  hipLaunchKernelGGL((k_bin_count<kLbNT, kLbVT>), dim3(cgrid), dim3(kLbNT), 0, s, ptr, xbeg, split, crow, ctl, xnnz,
                     rbits, nbins, ws.hist);
  hipLaunchKernelGGL((k_bin_plan<1024>), dim3(1), dim3(1024), 0, s, ws.hist, nbins, ws);
  hipLaunchKernelGGL((k_bin_scatter<kLbNT, kLbVT>), dim3(rgrid), dim3(kLbNT), 0, s, ptr, xbeg, xval, split, crow,
                     cval, ctl, xnnz, rbits, nbins, ws.cursor, bkey, bval);
  const long long bl = binlen_hint >= 0 ? binlen_hint : binlen_cap;
  const int agrid = (int)std::max<long long>(1, std::min<long long>(4096, bl / kBinChunk + nbins));
  const size_t lds = sizeof(float) * ((size_t)(1 << rbits) + ((1 << rbits) + 31) / 32);
  // a window of 2^15 rows takes 132 KB of the 160 KB of LDS
  static std::once_flag attr_once;
  std::call_once(attr_once, [] {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_bin_accum<kAccNT, kAccPPT>),
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)(sizeof(float) * ((1 << kBinRowBits) + (1 << kBinRowBits) / 32)));
  });
  hipLaunchKernelGGL((k_bin_accum<kAccNT, kAccPPT>), dim3(agrid), dim3(kAccNT), lds, s, bkey, bval, ws, nbins,
                     rbits, acc, marks);
  return true;
}

}  // namespace spmspv
