// Row-bin scatter/accumulate SpMSpV (slot 8). The reference's bucket method
// (CscBasedBucketSpmspvDriver, la/include/spmspv/csc-spmspv/detail/device/
// csc_bucket_spmspv.cuh:766-1086) counts the products per row bucket, scans
// the counts, extracts the products into their buckets and reduces each
// bucket; its reduction is compiled out (#if 0, :999-1071) and its atomic
// variant truncates values to int (:684). Here the buckets are row bins,
// ranges of consecutive rows holding about equal numbers of the matrix's
// nonzeros and at most 2^12 rows (one bin's sums fit a 48 KB LDS window,
// two accumulate workgroups per CU). The layout is fixed per matrix and bin count (BinLayout, built
// on the host from the row pointer): a bin never receives more products
// than its rows have nonzeros, so bin b owns the product slots [base[b],
// base[b+1]) with base[b] = nonzeros in the rows before it, and no per-call
// count is needed; equal nonzero counts keep the bins' product counts
// level on a power-law matrix, where equal row ranges put a tenth of them
// into the first bin. Per call:
//   1. k_bin_scatter: the load-balanced gather's merge-path tiles (k_prep's
//      split, kBinS of its tiles at a time), one per workgroup round, form
//      every product (row, a*x), find its bin (a table of row blocks, then a
//      search among the bins that start inside the row's block) and rank it
//      there in LDS; the round reserves each bin's slots with one global
//      atomic on the bin's cursor (one cursor per 128-byte line), places its
//      products bin by bin in LDS and writes them out in that order;
//   2. k_bin_accum: every workgroup reads the cursors (products per bin) and
//      plans the chunks itself: a bin is cut into chunks of kBinChunk
//      products (one round of registers), so a bin that an uneven x fills
//      far beyond its share still spreads over workgroups; one chunk per
//      workgroup. Each chunk is summed per row in an LDS window over its
//      bin's rows, in 64-bit fixed point scaled to the row's largest
//      product (integer LDS atomics: FP32 LDS atomic adds are about 15
//      times slower on gfx950), or in registers when the bin has at most
//      kFewRows rows (a hub row's products would all queue on one LDS
//      word); a bin that is a single chunk stores its row sums into the
//      accumulator, a longer one adds them (one FP32 atomic per row and
//      chunk); hit rows are marked;
//   3. the caller's compaction turns the marks into the sorted result,
//      clears the accumulator and the marks, and clears the cursors.
// HBM bytes per product: 8 read and 8 written (scatter) + 8 (accumulate),
// against 44 and up for gather, radix passes and reduce.
#include <algorithm>
#include <mutex>
#include <vector>

#include "handle.hpp"

namespace spmspv {

namespace {

// The scatter: 1,024-thread workgroups whose merge-path tile is kBinS of
// the prep's tiles (one latency chain for 7,168 merged positions), one
// such tile per round. Each round reserves its slots with one device-scope
// atomic per bin it fills (about 17 G/s on the chip, one address taking
// them in turn at about 16 ns).
constexpr int kScatNT = 1024, kBinS = 4, kScatGrid = 2048;
static_assert(kScatNT * kLbVT == kBinS * kLbTile, "the scatter's tile is kBinS prep tiles");
static_assert(kBinS * kLbTile < 65536 && kBinMax < 65535, "bin and rank share a word");
constexpr int kAccNT = 1024;        // accumulate workgroup
constexpr int kAccPPT = 16;         // products per thread per sub-round
constexpr int kAccGrid = 2048;      // accumulate workgroups at most
constexpr int kFewRows = 16;        // bins of at most this many rows sum in registers
constexpr int kMaxDevices = 64;     // devices whose kernel attributes are tracked

// LDS of the accumulate kernel for a window of rows rows and nbins bins
constexpr size_t accum_lds_bytes(int rows, int nbins) {
  return (sizeof(unsigned long long) + sizeof(unsigned)) * (size_t)rows + sizeof(int) * (3 * (size_t)nbins + 1);
}

// The products of merge-path tile t of the load-balanced gather, a tile of
// S of k_prep's tiles (T = NT * VT = S * kLbTile merged positions: x-entry
// markers and products; split[t * S] and split[min(t * S + S, ntiles)] are
// its first and end markers). For the thread's k-th item q = threadIdx.x +
// k * NT < np: idx[k] = its CSC index, xi[k] = its x entry. Returns np. The
// LDS tables are the caller's; the function's two barriers order their
// use, and the caller syncs before the next tile.
template <int NT, int VT, int S>
__device__ __forceinline__ int lb_tile(long long t, const int* __restrict__ ptr, const int* __restrict__ xbeg,
                                       const int* __restrict__ split, long long ntiles, int xnnz, long long W,
                                       int* s_ptr, int* s_beg, unsigned short* s_seg, int (&idx)[VT],
                                       int (&xi)[VT]) {
  constexpr int T = NT * VT;
  static_assert(T == S * kLbTile, "a tile is S of the prep's tiles");
  static_assert(T + 2 < 65536, "segment indices fit 16 bits");
  const long long d0 = t * T;
  const long long d1 = min(d0 + T, W);
  const int a0 = split[t * S];
  const int a1 = split[std::min<long long>(t * S + S, ntiles)];
  const long long b0 = d0 - a0;
  const long long b1 = d1 - a1;
  const int sa = a0 > 0 ? a0 - 1 : 0;
  const int nseg = a1 - sa + 1;  // <= T + 2
  {
    constexpr int SEGPT = (T + 2 + NT - 1) / NT;
    int pv[SEGPT], bv[SEGPT];
#pragma unroll
    for (int u = 0; u < SEGPT; ++u) {
      const int k = threadIdx.x + u * NT;
      const int i = sa + (k < nseg ? k : 0);
      const int ic = i < xnnz ? i : (xnnz > 0 ? xnnz - 1 : 0);
      pv[u] = ptr[i];
      bv[u] = xbeg[ic];
    }
#pragma unroll
    for (int u = 0; u < SEGPT; ++u) {
      const int k = threadIdx.x + u * NT;
      if (k < nseg) {
        s_ptr[k] = pv[u];
        s_beg[k] = bv[u];
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
        s_seg[p - b0] = (unsigned short)(i - 1 - sa);
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
    xi[k] = ok ? sa + sg : 0;
  }
  return np;
}

// The bin of row r: the last b with start[b] <= r. tbl[r >> shift] is the
// bin holding the first row of r's block, and tbl of the next block bounds
// the search, so a row in a block that lies inside one bin needs no step.
__device__ __forceinline__ int bin_of(const int* s_start, const int* s_tbl, int ntbl, int shift, int nbins,
                                      unsigned r) {
  const int blk = (int)(r >> shift);
  int lo = s_tbl[blk];
  int hi = blk + 1 < ntbl ? s_tbl[blk + 1] : nbins - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((unsigned)s_start[mid] <= r) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Dynamic LDS: the bin starts (nbins + 1), the row-block table (ntbl), the
// round's per-bin counts, turned into offsets (nbins + 1), and reserved
// bases (nbins). A round is one merge-path tile of kBinS prep tiles: its
// products are ranked per bin in LDS, the round reserves each bin's slots
// with one device-scope atomic, places the products bin by bin in LDS (the
// merge tables are free by then) and writes them out in that order, so
// consecutive lanes store consecutive slots of a bin (the texture
// addresser's cost follows the lines a store touches, and products stored
// straight from the lanes touched one line each).
template <int NT, int VT>
__global__ __launch_bounds__(NT) void k_bin_scatter(const int* __restrict__ ptr, const int* __restrict__ xbeg,
                                                    const float* __restrict__ xval, const int* __restrict__ split,
                                                    const int* __restrict__ crow, const float* __restrict__ cval,
                                                    const Ctl* __restrict__ ctl, int xnnz, int nbins, BinWs ws,
                                                    unsigned* __restrict__ okey, float* __restrict__ oval, Gate g) {
  if (gate_shut(g)) return;
  constexpr int T = NT * VT;
  constexpr int BPT = (kBinMax + 1 + NT - 1) / NT;
  __shared__ int s_ptr[T + 2];
  __shared__ int s_beg[T + 2];
  __shared__ unsigned short s_seg[T];
  __shared__ int s_w[NT / kWave];
  extern __shared__ int s_bins[];
  const int ntbl = ws.ntbl, shift = ws.tbl_shift;
  int* s_start = s_bins;
  int* s_tbl = s_start + nbins + 1;
  unsigned* s_c = reinterpret_cast<unsigned*>(s_tbl + ntbl);
  unsigned* s_base = s_c + nbins + 1;
  for (int b = threadIdx.x; b <= nbins; b += NT) s_start[b] = ws.start[b];
  for (int b = threadIdx.x; b < ntbl; b += NT) s_tbl[b] = ws.tbl[b];
  const long long ntiles = ctl->n_tiles;
  const long long nbig = (ntiles + kBinS - 1) / kBinS;
  const long long W = (long long)xnnz + ctl->bin_len;
  for (long long t = blockIdx.x; t < nbig; t += gridDim.x) {
    for (int b = threadIdx.x; b <= nbins; b += NT) s_c[b] = 0u;
    int idx[VT], xi[VT];
    const int np = lb_tile<NT, VT, kBinS>(t, ptr, xbeg, split, ntiles, xnnz, W, s_ptr, s_beg, s_seg, idx, xi);
    unsigned row[VT];
    float val[VT];
    unsigned br[VT];  // bin << 16 | rank in the bin (~0u: no product)
    {
      float a[VT], xs[VT];
#pragma unroll
      for (int k = 0; k < VT; ++k) {
        const bool ok = (int)threadIdx.x + k * NT < np;
        row[k] = ok ? (unsigned)crow[idx[k]] : 0u;
        a[k] = ok ? cval[idx[k]] : 0.f;
        xs[k] = ok ? xval[xi[k]] : 0.f;
      }
#pragma unroll
      for (int k = 0; k < VT; ++k) {
        val[k] = a[k] * xs[k];
        br[k] = ~0u;
        if ((int)threadIdx.x + k * NT < np) {
          const int b = bin_of(s_start, s_tbl, ntbl, shift, nbins, row[k]);
          br[k] = ((unsigned)b << 16) | atomicAdd(&s_c[b], 1u);
        }
      }
    }
    __syncthreads();
    // reserve each bin's slots, and the bins' offsets in the round
    {
      unsigned c[BPT];
      int sum = 0;
#pragma unroll
      for (int j = 0; j < BPT; ++j) {
        const int b = threadIdx.x * BPT + j;
        c[j] = b < nbins ? s_c[b] : 0u;
        sum += (int)c[j];
        if (c[j]) s_base[b] = (unsigned)ws.base[b] + atomicAdd(ws.cursor + (long long)b * kBinCurStride, c[j]);
      }
      int tot;
      unsigned pre = (unsigned)block_excl_sum<NT, int>(sum, s_w, &tot);
#pragma unroll
      for (int j = 0; j < BPT; ++j) {
        const int b = threadIdx.x * BPT + j;
        if (b <= nbins) s_c[b] = pre;  // every thread read its counts before the scan's barriers
        pre += c[j];
      }
    }
    __syncthreads();
    unsigned* s_row = reinterpret_cast<unsigned*>(s_ptr);
    float* s_val = reinterpret_cast<float*>(s_beg);
#pragma unroll
    for (int k = 0; k < VT; ++k) {
      if (br[k] != ~0u) {
        const unsigned b = br[k] >> 16;
        const unsigned q = s_c[b] + (br[k] & 0xFFFFu);
        s_row[q] = row[k];
        s_val[q] = val[k];
        s_seg[q] = (unsigned short)b;
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < np; i += NT) {
      const unsigned b = s_seg[i];
      const unsigned p = s_base[b] + ((unsigned)i - s_c[b]);
      okey[p] = s_row[i];
      oval[p] = s_val[i];
    }
    __syncthreads();  // the tables, counts and bases are reused by the next round
  }
}

// Fixed-point row sums. A product v = m * 2^(ef - 150), with m its 24-bit
// significand and ef = max(exponent field, 1), enters row r's sum as the
// integer q = m * 2^(ef - E + 22) (rounded to nearest when the shift is to
// the right), where E is the largest ef among the products of r in the
// chunk: |q| < 2^46, so a chunk of at most 2^14 products sums exactly in 64
// bits, in any order, and the row's value is the sum times 2^(E - 172),
// rounded once to FP32. Each product's error is at most 2^-46 of the row's
// largest |product|. Integer LDS atomics (max for E, add for the sum) run
// about 15 times faster on gfx950 than FP32 LDS atomic adds, which were
// the kernel's limit.
__device__ __forceinline__ unsigned prod_ef(float v) {
  const unsigned e = (__float_as_uint(v) >> 23) & 0xFFu;
  return e ? e : 1u;
}

__device__ __forceinline__ long long prod_fixed(float v, unsigned E) {
  const unsigned b = __float_as_uint(v);
  const unsigned ef = (b >> 23) & 0xFFu;
  const unsigned long long m = (b & 0x7FFFFFu) | (ef ? 0x800000u : 0u);
  const int d = (int)E - (int)(ef ? ef : 1u);
  unsigned long long q;
  if (d <= 22) q = m << (22 - d);
  else if (d - 22 <= 24) q = (m + (1ull << (d - 23))) >> (d - 22);
  else q = 0;
  return (b >> 31) ? -(long long)q : (long long)q;
}

__device__ __forceinline__ float fixed_value(long long sum, unsigned E) {
  return (float)ldexp((double)sum, (int)E - 172);
}

// Dynamic LDS: the window (per row of the largest bin: the 64-bit fixed-
// point sum, then the exponent E, 0 while the row has no product), then
// the plan: products, first slot and first chunk of every bin.
template <int NT, int PPT>
__global__ __launch_bounds__(NT) void k_bin_accum(const unsigned* __restrict__ key, const float* __restrict__ val,
                                                  BinWs ws, int nbins, float* acc, unsigned char* marks, Gate g) {
  if (gate_shut(g)) return;
  constexpr int SUB = NT * PPT;
  static_assert(SUB == kBinChunk, "a chunk is one round of registers");
  constexpr int BPT = kBinMax / NT;
  constexpr int NW = NT / kWave;
  extern __shared__ unsigned long long s_q[];
  const int R = ws.max_rows;
  unsigned* s_e = reinterpret_cast<unsigned*>(s_q + R);
  int* s_cnt = reinterpret_cast<int*>(s_e + R);
  int* s_base = s_cnt + nbins;
  int* s_cst = s_base + nbins;
  __shared__ int s_i[NW];
  __shared__ float s_few[NW][kFewRows];  // per-wave sums of a bin of few rows
  __shared__ unsigned s_fewhit[NW];
  // the plan, the same in every workgroup: a bin is cut into chunks of
  // kBinChunk products (the last one shorter)
  {
    unsigned c[BPT];
    int bb[BPT];
#pragma unroll
    for (int j = 0; j < BPT; ++j) {
      const int b = threadIdx.x * BPT + j;
      c[j] = b < nbins ? ws.cursor[(long long)b * kBinCurStride] : 0u;
      bb[j] = b < nbins ? ws.base[b] : 0;
    }
    int nsum = 0;
    int nch[BPT];
#pragma unroll
    for (int j = 0; j < BPT; ++j) {
      nch[j] = (int)((c[j] + kBinChunk - 1) / kBinChunk);
      nsum += nch[j];
    }
    int ntot;
    int npre = block_excl_sum<NT, int>(nsum, s_i, &ntot);
#pragma unroll
    for (int j = 0; j < BPT; ++j) {
      const int b = threadIdx.x * BPT + j;
      if (b < nbins) {
        s_cnt[b] = (int)c[j];
        s_base[b] = bb[j];
        s_cst[b] = npre;
      }
      npre += nch[j];
    }
    if (threadIdx.x == 0) s_cst[nbins] = ntot;
  }
  for (int r = threadIdx.x; r < R; r += NT) {
    s_q[r] = 0ull;
    s_e[r] = 0u;
  }
  __syncthreads();
  const int nch = s_cst[nbins];
  for (int c = blockIdx.x; c < nch; c += gridDim.x) {
    // the bin holding chunk c: the last b with s_cst[b] <= c
    int lo = 0, hi = nbins - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (s_cst[mid] <= c) lo = mid;
      else hi = mid - 1;
    }
    const int b = lo;
    const int c0 = s_cst[b], nb = s_cst[b + 1] - c0;
    const bool owned = nb == 1;
    // chunk c - c0 of nb parts of at most kBinChunk products; rows [rs, rs + nr)
    const long long p0 = (long long)s_base[b] + (long long)(c - c0) * kBinChunk;
    const long long pe = std::min<long long>(p0 + kBinChunk, (long long)s_base[b] + s_cnt[b]);
    const unsigned rs = (unsigned)ws.start[b];
    const int nr = ws.start[b + 1] - (int)rs;
    unsigned kk[PPT];
    float vv[PPT];
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
      const long long i = p0 + threadIdx.x + (long long)j * NT;
      const long long ic = i < pe ? i : p0;
      kk[j] = __builtin_nontemporal_load(key + ic);
      vv[j] = __builtin_nontemporal_load(val + ic);
    }
    if (nr <= kFewRows) {
      // a bin of few (hub) rows: the lanes of a wave would all add into the
      // same LDS words, so each thread sums per row in registers and the
      // workgroup reduces the kFewRows sums
      float part[kFewRows];
      unsigned hit = 0u;
#pragma unroll
      for (int q = 0; q < kFewRows; ++q) part[q] = 0.f;
#pragma unroll
      for (int j = 0; j < PPT; ++j) {
        if (p0 + threadIdx.x + (long long)j * NT < pe) {
          const unsigned r = kk[j] - rs;
          hit |= 1u << r;
#pragma unroll
          for (int q = 0; q < kFewRows; ++q) part[q] += r == (unsigned)q ? vv[j] : 0.f;
        }
      }
#pragma unroll
      for (int q = 0; q < kFewRows; ++q) part[q] = wave_sum(part[q]);
#pragma unroll
      for (int o = kWave / 2; o > 0; o >>= 1) hit |= __shfl_xor(hit, o, kWave);
      const int w = threadIdx.x / kWave;
      if (lane() == 0) {
#pragma unroll
        for (int q = 0; q < kFewRows; ++q) s_few[w][q] = part[q];
        s_fewhit[w] = hit;
      }
      __syncthreads();
      if ((int)threadIdx.x < nr) {
        const int q = threadIdx.x;
        float v = 0.f;
        unsigned h = 0u;
#pragma unroll
        for (int k = 0; k < NW; ++k) {
          v += s_few[k][q];
          h |= s_fewhit[k];
        }
        if ((h >> q) & 1u) {
          if (owned) acc[rs + q] = v;  // the accumulator holds 0 there
          else atomic_add_f32(acc + rs + q, v);
          marks[rs + q] = 1;
        }
      }
      __syncthreads();  // s_few is reused by the next chunk
      continue;
    }
    // every row's largest exponent first; a chunk with an Inf or NaN
    // product sums in FP32 instead (the window's sums read as floats)
    int special = 0;
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
      if (p0 + threadIdx.x + (long long)j * NT < pe) {
        const unsigned ef = prod_ef(vv[j]);
        special |= ef == 0xFFu;
        atomicMax(&s_e[kk[j] - rs], ef);
      }
    }
    special = __syncthreads_or(special);
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
      if (p0 + threadIdx.x + (long long)j * NT < pe) {
        const unsigned r = kk[j] - rs;
        if (special) atomicAdd(reinterpret_cast<float*>(s_q + r), vv[j]);
        else atomicAdd(s_q + r, (unsigned long long)prod_fixed(vv[j], s_e[r]));
      }
    }
    __syncthreads();
    // the window's hit rows into the accumulator, consecutive rows on
    // consecutive threads; the window returns to zero
    for (int r = threadIdx.x; r < nr; r += NT) {
      const unsigned E = s_e[r];
      if (E) {
        const float v = special ? *reinterpret_cast<const float*>(s_q + r) : fixed_value((long long)s_q[r], E);
        s_q[r] = 0ull;
        s_e[r] = 0u;
        if (owned) acc[rs + r] = v;  // the accumulator holds 0 there
        else atomic_add_f32(acc + rs + r, v);
        marks[rs + r] = 1;
      }
    }
    __syncthreads();
  }
}

// Nonzeros per row from the CSC row indices (a CSC-only matrix), added into
// cnt (m ints, zero before).
__global__ void k_row_counts(const int* __restrict__ crow, long long nnz, int* __restrict__ cnt) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += stride)
    atomicAdd(cnt + crow[i], 1);
}

}  // namespace

int bin_layout(const std::vector<int>& row_ptr, int m, int bucket_bits, BinLayout* out) {
  out->start.clear();
  out->base.clear();
  out->tbl.clear();
  const long long nnz = m > 0 ? row_ptr[m] : 0;
  const long long want = std::max<long long>(1, (nnz + (1ll << bucket_bits) - 1) >> bucket_bits);
  const int cap = 1 << kBinRowBits;
  int mr = 0;
  int r = 0;
  out->start.push_back(0);
  out->base.push_back(0);
  while (r < m) {
    // the bin grows row by row while it stays within its share of nonzeros
    // (a row beyond the share on its own still makes a bin) and the cap
    int e = r + 1;
    const long long b0 = row_ptr[r];
    while (e < m && e - r < cap && row_ptr[e + 1] - b0 <= want) ++e;
    mr = std::max(mr, e - r);
    out->start.push_back(e);
    out->base.push_back(row_ptr[e]);
    r = e;
    if ((int)out->start.size() - 1 > kBinMax) return -1;
  }
  if (m == 0) {  // one empty bin
    out->start.push_back(0);
    out->base.push_back(0);
  }
  const int nb = (int)out->start.size() - 1;
  out->max_rows = std::max(mr, 1);
  // the row-block table: blocks of 2^shift rows, at most kBinTbl of them
  int shift = 0;
  while (((long long)std::max(m, 1) + (1ll << shift) - 1) >> shift > kBinTbl) ++shift;
  out->tbl_shift = shift;
  const int ntbl = (int)(((long long)std::max(m, 1) + (1ll << shift) - 1) >> shift);
  int b = 0;
  for (int k = 0; k < ntbl; ++k) {
    const long long first = (long long)k << shift;
    while (b + 1 < nb && out->start[b + 1] <= first) ++b;
    out->tbl.push_back(b);
  }
  return nb;
}

void launch_row_counts(hipStream_t s, const int* crow, long long nnz, int* cnt) {
  const int grid = (int)std::max<long long>(1, std::min<long long>(4096, (nnz + 255) / 256));
  hipLaunchKernelGGL(k_row_counts, dim3(grid), dim3(256), 0, s, crow, nnz, cnt);
}

void launch_bucket(hipStream_t s, const int* ptr, const int* xbeg, const float* xval, const int* split,
                   const int* crow, const float* cval, const Ctl* ctl, int xnnz, const BinWs& ws, int nbins,
                   unsigned* bkey, float* bval, float* acc, unsigned char* marks, long long tiles_hint,
                   long long binlen_hint, long long binlen_cap, Gate g) {
  // dynamic LDS beyond 64 KB (many bins, wide windows) needs the attribute
  // once per device (the attribute belongs to the device's copy of the kernels)
  static std::once_flag attr_once[kMaxDevices];
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::call_once(attr_once[dev % kMaxDevices], [] {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_bin_scatter<kScatNT, kLbVT>),
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)(sizeof(int) * (3 * (size_t)kBinMax + 2 + kBinTbl)));
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_bin_accum<kAccNT, kAccPPT>),
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)accum_lds_bytes(1 << kBinRowBits, kBinMax));
  });
  // one round per big tile, the workgroups looping when there are more
  const long long tiles = tiles_hint > 0 ? tiles_hint : ((long long)xnnz + binlen_cap) / kLbTile + 2;
  // no more workgroups than fit at once: each loads the bin tables once and
  // loops over its rounds
  const int sres = resident_blocks(reinterpret_cast<const void*>(k_bin_scatter<kScatNT, kLbVT>), kScatNT);
  const int rgrid = (int)std::max<long long>(
      1, std::min<long long>(std::min(kScatGrid, sres), (tiles + kBinS - 1) / kBinS));
  const size_t slds = sizeof(int) * (3 * (size_t)nbins + 2 + ws.ntbl);
  hipLaunchKernelGGL((k_bin_scatter<kScatNT, kLbVT>), dim3(rgrid), dim3(kScatNT), slds, s, ptr, xbeg, xval, split, crow,
                     cval, ctl, xnnz, nbins, ws, bkey, bval, g);
  // one chunk per workgroup where the grid allows; with bin_len unknown on
  // the host, two per bin on average (the workgroups loop over the rest)
  const long long extra = binlen_hint >= 0 ? std::min(binlen_hint, binlen_cap) / kBinChunk : nbins;
  // no more workgroups than fit at once: one that loops over further chunks
  // keeps its plan, where a fresh one would build it again
  const int ares = resident_blocks(reinterpret_cast<const void*>(k_bin_accum<kAccNT, kAccPPT>), kAccNT);
  const int agrid = (int)std::max<long long>(1, std::min<long long>(std::min(kAccGrid, ares), nbins + extra));
  hipLaunchKernelGGL((k_bin_accum<kAccNT, kAccPPT>), dim3(agrid), dim3(kAccNT), accum_lds_bytes(ws.max_rows, nbins),
                     s, bkey, bval, ws, nbins, acc, marks, g);
}

}  // namespace spmspv
