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
  int t = blockIdx.x;
  if (t < ntiles) load_tile(t);
  for (; t < ntiles; t += gridDim.x) {
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
    if (t + (int)gridDim.x < ntiles) load_tile(t + gridDim.x);
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
    if (HITS) {
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
