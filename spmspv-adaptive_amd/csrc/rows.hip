// Row-driven kernels for gfx950 (CSR, one output row per matrix row):
//  * k_row_merge: nnz-and-row balanced SpMV / masked SpMSpV. Each tile owns
//    a fixed share of the merged sequence (row ends + nonzeros), so skewed
//    rows (R-MAT hubs) and runs of empty rows cost the same as any other
//    tile. Products are formed with coalesced loads into LDS, each thread
//    walks its items, and partial rows are joined by a fixed-order
//    segmented scan inside the tile and a fixed-order carry fix-up across
//    tiles (deterministic; y needs no pre-zeroing).
//    Role of HolaSpMV / HolaSpMSpV (la/.../detail/device/holaspmv.h:2069-2366,
//    MultiplyAlong :1551, MultiplyAlongSpMspV :1613) and of the CSR5 kernels
//    (csr5_spmv_cuda.h:440-582, csr5_spmspv.h:449-610): K5/K7.
//  * k_row_naive: row-granular CSR-vector with a power-of-two team of lanes
//    per row and no nnz balancing (naive_spmv_ker / naive_spmspv_ker,
//    naivespmv.h:12-50): K4/K6.
// Masked variants skip a nonzero unless its column is set in x's bit array
// (the bitvector test of naivespmv.h:12-31, holaspmv.h:747-830). With
// `hits`, each row also reports whether any masked-in column touched it, so
// the structural support can be compacted afterwards.
#include "kernels.hpp"

namespace spmspv {

__device__ __forceinline__ bool test_bit(const unsigned* __restrict__ bits, int c) {
  return (bits[c >> 5] >> (c & 31)) & 1u;
}

// split[t] = rows whose end marker lies before merged position t*T
// (MergePath search over row_ptr[1..m] and 0..nnz-1).
__global__ void k_row_split(const int* __restrict__ row_ptr, int m, int nnz, int T, int ntiles,
                            int* __restrict__ split) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > ntiles) return;
  const long long d = std::min((long long)t * T, (long long)m + nnz);
  long long lo = std::max(0ll, d - nnz), hi = std::min(d, (long long)m);
  while (lo < hi) {
    const long long mid = (lo + hi) >> 1;
    if (row_ptr[mid + 1] <= d - mid - 1) lo = mid + 1;
    else hi = mid;
  }
  split[t] = (int)lo;
}

void launch_row_split(hipStream_t s, const int* row_ptr, int m, int nnz, int* split, int ntiles,
                      int tile) {
  const int grid = (ntiles + 1 + 255) / 256;
  hipLaunchKernelGGL(k_row_split, dim3(grid), dim3(256), 0, s, row_ptr, m, nnz, tile, ntiles, split);
}

template <int NT>
__device__ __forceinline__ void seg_scan_row(bool f, float v, int h, float* cin, int* hin,
                                             float* tail_v, int* tail_h, int* s_f, float* s_v,
                                             int* s_h) {
  constexpr int NW = NT / kWave;
  const int w = threadIdx.x / kWave;
  int fi = f ? 1 : 0;
  float vi = v;
  int hi = h;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const int fo = __shfl_up(fi, o, kWave);
    const float vo = __shfl_up(vi, o, kWave);
    const int ho = __shfl_up(hi, o, kWave);
    if (lane() >= o) {
      if (!fi) {
        vi = vo + vi;
        hi = ho | hi;
      }
      fi |= fo;
    }
  }
  if (lane() == kWave - 1) {
    s_f[w] = fi;
    s_v[w] = vi;
    s_h[w] = hi;
  }
  __syncthreads();
  float pv = 0.f;
  int ph = 0;
  for (int k = 0; k < w; ++k) {
    if (s_f[k]) {
      pv = s_v[k];
      ph = s_h[k];
    } else {
      pv += s_v[k];
      ph |= s_h[k];
    }
  }
  float tv = 0.f;
  int th = 0;
  for (int k = 0; k < NW; ++k) {
    if (s_f[k]) {
      tv = s_v[k];
      th = s_h[k];
    } else {
      tv += s_v[k];
      th |= s_h[k];
    }
  }
  const float incl_v = fi ? vi : pv + vi;
  const int incl_h = fi ? hi : (ph | hi);
  float ev = __shfl_up(incl_v, 1, kWave);
  int eh = __shfl_up(incl_h, 1, kWave);
  if (lane() == 0) {
    ev = pv;
    eh = ph;
  }
  __syncthreads();
  *cin = ev;
  *hin = eh;
  *tail_v = tv;
  *tail_h = th;
}

// Walks tiles t0, t0+tstep, ... < ntiles of one CSR (or one column block
// of it). Completed rows are stored to y; hits go to whole bit words
// (MARKS false) or to one mark byte per row (MARKS true, used when several
// column blocks report hits for the same rows).
template <int NT, int VT, bool MASKED, bool HITS, bool MARKS>
__device__ __forceinline__ void merge_tiles(const int* __restrict__ row_ptr,
                                            const int* __restrict__ col,
                                            const float* __restrict__ val,
                                            const float* __restrict__ x,
                                            const unsigned* __restrict__ xbits, int m, int nnz,
                                            const int* __restrict__ split, int ntiles, int t0,
                                            int tstep, float* __restrict__ y, unsigned* ybits,
                                            unsigned char* marks, float* __restrict__ carry_val,
                                            int* __restrict__ carry_row,
                                            int* __restrict__ carry_hit) {
  constexpr int T = NT * VT;
  constexpr int NW = NT / kWave;
  // ~16 KB of LDS per workgroup (rows are written straight to y), so that
  // more workgroups per CU keep loads in flight
  __shared__ int s_re[T + 1];
  __shared__ float s_p[T];
  __shared__ unsigned char s_ph[HITS ? T : 1];
  __shared__ unsigned char s_yh[HITS ? T : 1];
  __shared__ int s_f[NW];
  __shared__ float s_sv[NW];
  __shared__ int s_sh[NW];
  const long long W = (long long)m + nnz;
  // the tile's nonzeros are loaded one tile ahead: the loads of tile t+grid
  // are in flight while tile t is walked, scanned and written
  int cj[VT];
  float vj[VT];
  auto load_tile = [&](int t) {
    const long long d0 = (long long)t * T;
    const long long d1 = std::min(d0 + T, W);
    const long long j0 = d0 - split[t];
    const int nj = (int)(d1 - split[t + 1] - j0);
#pragma unroll
    for (int i = 0; i < VT; ++i) {
      const int k = threadIdx.x + i * NT;
      cj[i] = 0;
      vj[i] = 0.f;
      if (k < nj) {
        cj[i] = __builtin_nontemporal_load(col + j0 + k);
        vj[i] = __builtin_nontemporal_load(val + j0 + k);
      }
    }
  };
  int t = t0;
  if (t < ntiles) load_tile(t);
  for (; t < ntiles; t += tstep) {
    const long long d0 = (long long)t * T;
    const long long d1 = std::min(d0 + T, W);
    const int i0 = split[t];
    const int i1 = split[t + 1];
    const long long j0 = d0 - i0;
    const long long j1 = d1 - i1;
    const int nr = i1 - i0;
    const int nj = (int)(j1 - j0);
    for (int k = threadIdx.x; k <= nr; k += NT) {
      const int r = i0 + k;
      s_re[k] = r < m ? row_ptr[r + 1] : nnz;
    }
#pragma unroll
    for (int i = 0; i < VT; ++i) {
      const int k = threadIdx.x + i * NT;
      if (k < nj) {
        const int c = cj[i];
        float p;
        int h = 0;
        if (MASKED) {
          h = test_bit(xbits, c) ? 1 : 0;
          p = h ? vj[i] * x[c] : 0.f;
        } else {
          p = vj[i] * x[c];
          if (HITS) h = test_bit(xbits, c) ? 1 : 0;
        }
        s_p[k] = p;
        if (HITS) s_ph[k] = (unsigned char)h;
      }
    }
    __syncthreads();
    if (t + tstep < ntiles) load_tile(t + tstep);
    long long D = d0 + (long long)threadIdx.x * VT;
    bool had_end = false;
    float first_sum = 0.f, sum = 0.f;
    int first_hit = 0, hit = 0, first_row = 0;
    if (D < d1) {
      long long lo = std::max((long long)i0, D - nnz), hi = std::min(D, (long long)i1);
      while (lo < hi) {
        const long long mid = (lo + hi) >> 1;
        if (s_re[mid - i0] <= D - mid - 1) lo = mid + 1;
        else hi = mid;
      }
      int i = (int)lo;
      long long j = D - i;
#pragma unroll
      for (int k = 0; k < VT; ++k) {
        if (D >= d1) break;
        if (j < s_re[i - i0]) {
          sum += s_p[j - j0];
          if (HITS) hit |= s_ph[j - j0];
          ++j;
        } else {
          if (!had_end) {
            first_row = i;
            first_sum = sum;
            first_hit = hit;
            had_end = true;
          } else {
            y[i] = sum;
            if (HITS) s_yh[i - i0] = (unsigned char)hit;
          }
          sum = 0.f;
          hit = 0;
          ++i;
        }
        ++D;
      }
    }
    float cin, tail_v;
    int hin, tail_h;
    seg_scan_row<NT>(had_end, sum, hit, &cin, &hin, &tail_v, &tail_h, s_f, s_sv, s_sh);
    if (had_end) {
      y[first_row] = cin + first_sum;
      if (HITS) s_yh[first_row - i0] = (unsigned char)(hin | first_hit);
    }
    if (threadIdx.x == 0) {
      carry_row[t] = (i1 < m) ? i1 : -1;
      carry_val[t] = tail_v;
      carry_hit[t] = tail_h;
    }
    if (HITS && MARKS) {
      __syncthreads();
      for (int k = threadIdx.x; k < nr; k += NT)
        if (s_yh[k]) marks[i0 + k] = 1;
    } else if (HITS) {
      __syncthreads();
      if (nr > 0) {
        // hit bits of rows [i0, i1): a word wholly inside the range belongs
        // to this tile alone and is stored; the two edge words are OR-ed in
        const int w0 = i0 >> 5, w1 = (i1 - 1) >> 5;
        for (int w = w0 + threadIdx.x; w <= w1; w += NT) {
          const int rlo = max(i0, w << 5), rhi = min(i1, (w << 5) + 32);
          unsigned bits = 0u;
          for (int r = rlo; r < rhi; ++r) bits |= (unsigned)s_yh[r - i0] << (r & 31);
          if (rlo == (w << 5) && rhi == (w << 5) + 32) ybits[w] = bits;
          else if (bits) atomicOr(ybits + w, bits);
        }
      }
    }
    __syncthreads();
  }
}

template <int NT, int VT, bool MASKED, bool HITS>
__global__ __launch_bounds__(NT) void k_row_merge(const int* __restrict__ row_ptr,
                                                  const int* __restrict__ col,
                                                  const float* __restrict__ val,
                                                  const float* __restrict__ x,
                                                  const unsigned* __restrict__ xbits, int m, int nnz,
                                                  const int* __restrict__ split, int ntiles,
                                                  float* __restrict__ y, unsigned* ybits,
                                                  float* __restrict__ carry_val,
                                                  int* __restrict__ carry_row,
                                                  int* __restrict__ carry_hit) {
  merge_tiles<NT, VT, MASKED, HITS, false>(row_ptr, col, val, x, xbits, m, nnz, split, ntiles,
                                           blockIdx.x, gridDim.x, y, ybits, nullptr, carry_val,
                                           carry_row, carry_hit);
}

// Column-blocked variant: the matrix is split into nb <= 8 column blocks of
// about equal nnz, each a CSR over all m rows. Workgroups are dispatched to
// the 8 XCDs round-robin (XCD = blockIdx % 8), so workgroup g serves block
// (g % 8) % nb: every XCD gathers x only from its block's column range,
// which stays resident in that XCD's L2. Block b writes its partial rows to
// y + b*m (every row, so no clearing is needed); k_xsum adds the blocks.
template <int NT, int VT, bool MASKED, bool HITS>
__global__ __launch_bounds__(NT) void k_row_merge_x(XBlocks xb, const float* __restrict__ x,
                                                    const unsigned* __restrict__ xbits, int m,
                                                    float* __restrict__ ypart,
                                                    unsigned char* marks,
                                                    float* __restrict__ carry_val,
                                                    int* __restrict__ carry_row,
                                                    int* __restrict__ carry_hit) {
  const int xcd = blockIdx.x & 7;
  const int slot = blockIdx.x >> 3;
  const int b = xcd % xb.nb;
  const int reps = 8 / xb.nb;
  const int idx = slot * reps + xcd / xb.nb;
  const int step = (gridDim.x >> 3) * reps;
  const int base = xb.tile_base[b];
  merge_tiles<NT, VT, MASKED, HITS, true>(
      xb.row_ptr + (size_t)b * (m + 1), xb.col + xb.col_base[b], xb.val + xb.col_base[b], x, xbits, m,
      xb.nnz[b], xb.split + xb.split_base[b], xb.ntiles[b], idx, step, ypart + (size_t)b * m, nullptr,
      marks, carry_val + base, carry_row + base, carry_hit + base);
}

// Carry fix-up of the column-blocked variant: tile t belongs to the block
// whose tile range holds it; runs of equal carry rows never cross blocks.
__global__ void k_row_fixup_x(XBlocks xb, const float* __restrict__ carry_val,
                              const int* __restrict__ carry_row, const int* __restrict__ carry_hit,
                              int m, float* ypart, unsigned char* marks) {
  const int total = xb.tile_base[xb.nb - 1] + xb.ntiles[xb.nb - 1];
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    int b = 0;
    while (b + 1 < xb.nb && t >= xb.tile_base[b + 1]) ++b;
    const int lo = xb.tile_base[b], hi = lo + xb.ntiles[b];
    const int r = carry_row[t];
    if (r < 0) continue;
    if (t > lo && carry_row[t - 1] == r) continue;
    float s = 0.f;
    int h = 0;
    for (int u = t; u < hi && carry_row[u] == r; ++u) {
      s += carry_val[u];
      h |= carry_hit[u];
    }
    ypart[(size_t)b * m + r] += s;
    if (marks && h) marks[r] = 1;
  }
}

// y[r] = sum over blocks of the partial rows, in block order.
__global__ void k_xsum(const float* __restrict__ ypart, int nb, int m, float* __restrict__ y) {
  for (long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x; r < m;
       r += (long long)gridDim.x * blockDim.x) {
    float s = ypart[r];
    for (int b = 1; b < nb; ++b) s += ypart[(size_t)b * m + r];
    y[r] = s;
  }
}

void launch_row_merge_x(hipStream_t s, bool masked, bool hits, const XBlocks& xb, const float* x,
                        const unsigned* xbits, int m, int vt, float* ypart, unsigned char* marks,
                        float* carry_val, int* carry_row, int* carry_hit, float* y) {
  const int total = xb.tile_base[xb.nb - 1] + xb.ntiles[xb.nb - 1];
  if (total <= 0) return;
  const int grid = 8 * std::max(1, std::min(256, (total + 7) / 8));
#define SPMSPV_ROW_MERGE_X(V, MK, HT)                                                           \
  hipLaunchKernelGGL((k_row_merge_x<kRowNT, V, MK, HT>), dim3(grid), dim3(kRowNT), 0, s, xb, x, \
                     xbits, m, ypart, marks, carry_val, carry_row, carry_hit)
#define SPMSPV_ROW_MERGE_X_VT(V)                           \
  if (masked && hits) SPMSPV_ROW_MERGE_X(V, true, true);   \
  else if (masked) SPMSPV_ROW_MERGE_X(V, true, false);     \
  else if (hits) SPMSPV_ROW_MERGE_X(V, false, true);       \
  else SPMSPV_ROW_MERGE_X(V, false, false);
  if (vt == 5) {
    SPMSPV_ROW_MERGE_X_VT(5)
  } else if (vt == 11) {
    SPMSPV_ROW_MERGE_X_VT(11)
  } else {
    SPMSPV_ROW_MERGE_X_VT(7)
  }
#undef SPMSPV_ROW_MERGE_X_VT
#undef SPMSPV_ROW_MERGE_X
  const int fgrid = std::min((total + 255) / 256, 1024);
  hipLaunchKernelGGL(k_row_fixup_x, dim3(fgrid), dim3(256), 0, s, xb, carry_val, carry_row, carry_hit,
                     m, ypart, hits ? marks : nullptr);
  hipLaunchKernelGGL(k_xsum, dim3(std::min(8192, (m + 255) / 256)), dim3(256), 0, s, ypart, xb.nb, m, y);
}

// ---- preprocessing of the column-blocked copy
// bounds[0..nb]: column boundaries with ~nnz/nb stored entries per block
// (from the CSC column pointer), so dense low-index columns get narrow blocks.
__global__ void k_xbounds(const int* __restrict__ colptr, int n, int nnz, int nb, int* bounds) {
  const int k = threadIdx.x;
  if (k > nb) return;
  if (k == 0) {
    bounds[0] = 0;
    return;
  }
  if (k == nb) {
    bounds[nb] = n;
    return;
  }
  const long long target = (long long)nnz * k / nb;
  int lo = 0, hi = n;
  while (lo < hi) {  // first column whose start >= target
    const int mid = (lo + hi) >> 1;
    if (colptr[mid] < target) lo = mid + 1;
    else hi = mid;
  }
  bounds[k] = lo;
}

__device__ __forceinline__ int col_block(const int* __restrict__ bounds, int nb, int c) {
  int b = 0;
  while (b + 1 < nb && c >= bounds[b + 1]) ++b;
  return b;
}

// cnt[b*(m+1) + r] = entries of row r in block b. One wave per row; the
// columns of a row need not be sorted.
__global__ void k_xcount(const int* __restrict__ row_ptr, const int* __restrict__ col, int m,
                         const int* __restrict__ bounds, int nb, int* cnt) {
  const long long waves = (long long)gridDim.x * (blockDim.x / kWave);
  for (long long r = (long long)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; r < m;
       r += waves) {
    const int b0 = row_ptr[r], e0 = row_ptr[r + 1];
    int mine = 0;  // lane q < nb counts block q
    for (int j0 = b0; j0 < e0; j0 += kWave) {
      const int j = j0 + lane();
      const int b = j < e0 ? col_block(bounds, nb, col[j]) : -1;
      for (int q = 0; q < nb; ++q) {
        const int k = __popcll(__ballot(b == q));
        if (lane() == q) mine += k;
      }
    }
    if (lane() < nb) cnt[(size_t)lane() * (m + 1) + r] = mine;
  }
}

// In-place exclusive scan of n ints in three passes (2048 items per part).
constexpr int kScanPart = 2048;
__global__ __launch_bounds__(256) void k_scan_part(const int* __restrict__ a, long long n,
                                                   int* __restrict__ sums) {
  __shared__ int lds[4];
  const long long base = (long long)blockIdx.x * kScanPart;
  int v = 0;
#pragma unroll
  for (int k = 0; k < kScanPart / 256; ++k) {
    const long long i = base + k * 256 + threadIdx.x;
    if (i < n) v += a[i];
  }
  v = block_sum<256>(v, lds);
  if (threadIdx.x == 0) sums[blockIdx.x] = v;
}

__global__ __launch_bounds__(256) void k_scan_sums(int* sums, int nparts) {
  __shared__ int lds[4];
  int run = 0;
  for (int c = 0; c < nparts; c += 256) {
    const int i = c + threadIdx.x;
    const int v = i < nparts ? sums[i] : 0;
    int tot;
    const int ex = block_excl_sum<256>(v, lds, &tot);
    if (i < nparts) sums[i] = run + ex;
    run += tot;
  }
}

__global__ __launch_bounds__(256) void k_scan_apply(int* a, long long n, const int* __restrict__ sums) {
  __shared__ int lds[4];
  constexpr int PT = kScanPart / 256;
  const long long base = (long long)blockIdx.x * kScanPart + (long long)threadIdx.x * PT;
  int v[PT];
  int loc = 0;
#pragma unroll
  for (int k = 0; k < PT; ++k) {
    v[k] = base + k < n ? a[base + k] : 0;
    loc += v[k];
  }
  int tot;
  int run = sums[blockIdx.x] + block_excl_sum<256>(loc, lds, &tot);
#pragma unroll
  for (int k = 0; k < PT; ++k) {
    if (base + k < n) a[base + k] = run;
    run += v[k];
  }
}

// base[b] = global offset of block b (b <= nb) in the scanned counts
__global__ void k_xbase(const int* __restrict__ bptr, int m, int nb, int* base) {
  const int b = threadIdx.x;
  if (b < nb) base[b] = bptr[(size_t)b * (m + 1)];
  if (b == nb) base[nb] = bptr[(size_t)nb * (m + 1) - 1];
}

// block-local row pointers: subtract each block's global offset
__global__ void k_xrebase(int* bptr, int m, int nb, const int* __restrict__ base) {
  const long long total = (long long)nb * (m + 1);
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x)
    bptr[i] -= base[i / (m + 1)];
}

// One wave per row: every entry goes to its block at bptr[b][r] plus its
// rank among the row's earlier entries of that block (row order is kept).
__global__ void k_xcopy(const int* __restrict__ row_ptr, const int* __restrict__ col,
                        const float* __restrict__ val, int m, const int* __restrict__ bounds, int nb,
                        const int* __restrict__ bptr, const int* __restrict__ col_base,
                        int* __restrict__ bcol, float* __restrict__ bval) {
  const long long waves = (long long)gridDim.x * (blockDim.x / kWave);
  const unsigned long long below = (1ull << lane()) - 1ull;
  for (long long r = (long long)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; r < m;
       r += waves) {
    const int b0 = row_ptr[r], e0 = row_ptr[r + 1];
    int run[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int j0 = b0; j0 < e0; j0 += kWave) {
      const int j = j0 + lane();
      const int c = j < e0 ? col[j] : 0;
      const int b = j < e0 ? col_block(bounds, nb, c) : -1;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        if (q >= nb) break;
        const unsigned long long mask = __ballot(b == q);
        if (b == q) {
          const long long dst = (long long)col_base[q] + bptr[(size_t)q * (m + 1) + r] + run[q] +
                                __popcll(mask & below);
          bcol[dst] = c;
          bval[dst] = val[j];
        }
        run[q] += __popcll(mask);
      }
    }
  }
}

XStoreLayout xstore_layout(int m, int nnz, int nb, int tile) {
  XStoreLayout L{};
  auto up = [](size_t v) { return (v + 255) & ~size_t(255); };
  const long long cells = (long long)nb * (m + 1);
  L.max_tiles = (int)(((long long)nb * m + nnz) / tile + 2 * nb + 2);
  size_t off = 0;
  L.bounds = off; off += up(sizeof(int) * (nb + 1));
  L.base = off; off += up(sizeof(int) * (nb + 1));
  L.bptr = off; off += up(sizeof(int) * cells);
  L.sums = off; off += up(sizeof(int) * (cells / kScanPart + 2));
  L.col = off; off += up(sizeof(int) * (size_t)std::max(nnz, 1));
  L.val = off; off += up(sizeof(float) * (size_t)std::max(nnz, 1));
  L.split = off; off += up(sizeof(int) * (L.max_tiles + nb));
  L.carry_val = off; off += up(sizeof(float) * L.max_tiles);
  L.carry_row = off; off += up(sizeof(int) * L.max_tiles);
  L.carry_hit = off; off += up(sizeof(int) * L.max_tiles);
  L.ypart = off; off += up(sizeof(float) * (size_t)nb * std::max(m, 1));
  L.bytes = off;
  return L;
}

hipError_t build_xblocks(hipStream_t s, const int* row_ptr, const int* col, const float* val,
                         const int* colptr, int m, int n, int nnz, int nb, int tile, char* mem,
                         int* h_pin, XStore* out) {
  const XStoreLayout L = xstore_layout(m, nnz, nb, tile);
  int* bounds = reinterpret_cast<int*>(mem + L.bounds);
  int* base = reinterpret_cast<int*>(mem + L.base);
  int* bptr = reinterpret_cast<int*>(mem + L.bptr);
  int* sums = reinterpret_cast<int*>(mem + L.sums);
  int* bcol = reinterpret_cast<int*>(mem + L.col);
  float* bval = reinterpret_cast<float*>(mem + L.val);
  int* split = reinterpret_cast<int*>(mem + L.split);
  const long long cells = (long long)nb * (m + 1);
  hipLaunchKernelGGL(k_xbounds, dim3(1), dim3(64), 0, s, colptr, n, nnz, nb, bounds);
  launch_fill_i32(s, bptr, cells, 0);
  const int gw = (int)std::min<long long>((m + 3) / 4 + 1, 16384);
  hipLaunchKernelGGL(k_xcount, dim3(gw), dim3(256), 0, s, row_ptr, col, m, bounds, nb, bptr);
  const long long parts = (cells + kScanPart - 1) / kScanPart;
  hipLaunchKernelGGL(k_scan_part, dim3((unsigned)parts), dim3(256), 0, s, bptr, cells, sums);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(256), 0, s, sums, (int)parts);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)parts), dim3(256), 0, s, bptr, cells, sums);
  hipLaunchKernelGGL(k_xbase, dim3(1), dim3(64), 0, s, bptr, m, nb, base);
  hipLaunchKernelGGL(k_xrebase, dim3(std::min<long long>((cells + 255) / 256, 8192)), dim3(256), 0, s,
                     bptr, m, nb, base);
  hipLaunchKernelGGL(k_xcopy, dim3(gw), dim3(256), 0, s, row_ptr, col, val, m, bounds, nb, bptr, base,
                     bcol, bval);
  hipError_t e = hipMemcpyAsync(h_pin, base, sizeof(int) * (nb + 1), hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  XBlocks& xb = out->xb;
  xb = XBlocks{};
  xb.nb = nb;
  xb.row_ptr = bptr;
  xb.col = bcol;
  xb.val = bval;
  xb.split = split;
  int tiles = 0, splits = 0;
  for (int b = 0; b < nb; ++b) {
    xb.col_base[b] = h_pin[b];
    xb.nnz[b] = h_pin[b + 1] - h_pin[b];
    xb.ntiles[b] = (int)(((long long)m + xb.nnz[b] + tile - 1) / tile);
    xb.tile_base[b] = tiles;
    xb.split_base[b] = splits;
    launch_row_split(s, bptr + (size_t)b * (m + 1), m, xb.nnz[b], split + splits, xb.ntiles[b], tile);
    tiles += xb.ntiles[b];
    splits += xb.ntiles[b] + 1;
  }
  if (tiles > L.max_tiles) return hipErrorInvalidValue;
  out->ypart = reinterpret_cast<float*>(mem + L.ypart);
  out->carry_val = reinterpret_cast<float*>(mem + L.carry_val);
  out->carry_row = reinterpret_cast<int*>(mem + L.carry_row);
  out->carry_hit = reinterpret_cast<int*>(mem + L.carry_hit);
  out->tile = tile;
  return hipGetLastError();
}


template <bool HITS>
__global__ void k_row_fixup(const float* __restrict__ carry_val, const int* __restrict__ carry_row,
                            const int* __restrict__ carry_hit, int ntiles, float* y, unsigned* ybits) {
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < ntiles; t += gridDim.x * blockDim.x) {
    const int r = carry_row[t];
    if (r < 0) continue;
    if (t > 0 && carry_row[t - 1] == r) continue;
    float s = 0.f;
    int h = 0;
    for (int u = t; u < ntiles && carry_row[u] == r; ++u) {
      s += carry_val[u];
      h |= carry_hit[u];
    }
    y[r] += s;
    if (HITS && h) atomicOr(ybits + (r >> 5), 1u << (r & 31));
  }
}

void launch_row_merge(hipStream_t s, bool masked, bool hits, const int* row_ptr, const int* col,
                      const float* val, const float* x, const unsigned* xbits, int m, int nnz,
                      const int* split, int ntiles, int vt, float* y, unsigned* ybits,
                      float* carry_val, int* carry_row, int* carry_hit) {
  if (ntiles <= 0) return;
  // about eight workgroups per CU, each walking several tiles with the next
  // tile's loads in flight
  const int grid = std::min(ntiles, 2048);
#define SPMSPV_ROW_MERGE(V, MK, HT)                                                             \
  hipLaunchKernelGGL((k_row_merge<kRowNT, V, MK, HT>), dim3(grid), dim3(kRowNT), 0, s, row_ptr,  \
                     col, val, x, xbits, m, nnz, split, ntiles, y, ybits, carry_val, carry_row,  \
                     carry_hit)
#define SPMSPV_ROW_MERGE_VT(V)                          \
  if (masked && hits) SPMSPV_ROW_MERGE(V, true, true);  \
  else if (masked) SPMSPV_ROW_MERGE(V, true, false);    \
  else if (hits) SPMSPV_ROW_MERGE(V, false, true);      \
  else SPMSPV_ROW_MERGE(V, false, false);
  if (vt == 5) {
    SPMSPV_ROW_MERGE_VT(5)
  } else if (vt == 11) {
    SPMSPV_ROW_MERGE_VT(11)
  } else {
    SPMSPV_ROW_MERGE_VT(7)
  }
#undef SPMSPV_ROW_MERGE_VT
#undef SPMSPV_ROW_MERGE
  const int fgrid = std::min((ntiles + 255) / 256, 1024);
  if (hits)
    hipLaunchKernelGGL(k_row_fixup<true>, dim3(fgrid), dim3(256), 0, s, carry_val, carry_row,
                       carry_hit, ntiles, y, ybits);
  else
    hipLaunchKernelGGL(k_row_fixup<false>, dim3(fgrid), dim3(256), 0, s, carry_val, carry_row,
                       carry_hit, ntiles, y, ybits);
}

template <int NT, bool MASKED, bool HITS>
__global__ __launch_bounds__(NT) void k_row_naive(const int* __restrict__ row_ptr,
                                                  const int* __restrict__ col,
                                                  const float* __restrict__ val,
                                                  const float* __restrict__ x,
                                                  const unsigned* __restrict__ xbits, int m,
                                                  int lg_team, float* __restrict__ y,
                                                  unsigned* ybits) {
  const int team = 1 << lg_team;
  const int tpb = NT >> lg_team;
  const int tl = threadIdx.x & (team - 1);
  const int tm = threadIdx.x >> lg_team;
  for (long long r = (long long)blockIdx.x * tpb + tm; r < m; r += (long long)gridDim.x * tpb) {
    const int b = row_ptr[r];
    const int e = row_ptr[r + 1];
    float s = 0.f;
    int h = 0;
    // four independent loads in flight per lane (long rows are latency bound)
    int j = b + tl;
    for (; j + 3 * team < e; j += 4 * team) {
      int c[4];
      float v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        c[u] = __builtin_nontemporal_load(col + j + u * team);
        v[u] = __builtin_nontemporal_load(val + j + u * team);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (MASKED) {
          if (test_bit(xbits, c[u])) {
            s += v[u] * x[c[u]];
            h = 1;
          }
        } else {
          s += v[u] * x[c[u]];
          if (HITS) h |= test_bit(xbits, c[u]) ? 1 : 0;
        }
      }
    }
    for (; j < e; j += team) {
      const int c = col[j];
      if (MASKED) {
        if (test_bit(xbits, c)) {
          s += val[j] * x[c];
          h = 1;
        }
      } else {
        s += val[j] * x[c];
        if (HITS) h |= test_bit(xbits, c) ? 1 : 0;
      }
    }
    for (int o = team >> 1; o > 0; o >>= 1) {
      s += __shfl_xor(s, o, kWave);
      if (HITS) h |= __shfl_xor(h, o, kWave);
    }
    if (tl == 0) {
      y[r] = s;
      if (HITS && h) atomicOr(ybits + (r >> 5), 1u << (r & 31));
    }
  }
}

void launch_row_naive(hipStream_t s, bool masked, bool hits, const int* row_ptr, const int* col,
                      const float* val, const float* x, const unsigned* xbits, int m, int team,
                      float* y, unsigned* ybits) {
  if (m <= 0) return;
  int lg = 0;
  while ((1 << lg) < team && lg < 6) ++lg;
  const int tpb = 256 >> lg;
  const int grid = (int)std::min<long long>(16384, ((long long)m + tpb - 1) / tpb);
#define SPMSPV_ROW_NAIVE(MK, HT)                                                                   \
  hipLaunchKernelGGL((k_row_naive<256, MK, HT>), dim3(grid), dim3(256), 0, s, row_ptr, col, val, x, \
                     xbits, m, lg, y, ybits)
  if (masked && hits) SPMSPV_ROW_NAIVE(true, true);
  else if (masked) SPMSPV_ROW_NAIVE(true, false);
  else if (hits) SPMSPV_ROW_NAIVE(false, true);
  else SPMSPV_ROW_NAIVE(false, false);
#undef SPMSPV_ROW_NAIVE
}

}  // namespace spmspv
