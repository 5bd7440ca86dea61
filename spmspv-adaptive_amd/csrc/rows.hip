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
#include <map>
#include <mutex>

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

// Block-wide segmented scan of one (flag, value, hit) per thread in thread
// order. Returns the exclusive carry into this thread (cin, hin) and the
// tile's trailing segment (tail_v, tail_h: everything after the last flag).
template <int NT>
__device__ __forceinline__ void seg_scan_row(bool f, float v, int h, float* cin, int* hin,
                                             float* tail_v, int* tail_h, int* s_f, float* s_v,
                                             int* s_h) {
  constexpr int NW = NT / kWave;
  const int w = threadIdx.x / kWave;
  int fi = f ? 1 : 0;
  float vi = v;
  int hi = h;
  wave_seg_scan(fi, vi, hi);
  if (lane() == kWave - 1) {
    s_f[w] = fi;
    s_v[w] = vi;
    s_h[w] = hi;
  }
  __syncthreads();
  float pv = 0.f;  // carry into this wave
  int ph = 0;
  float tv = 0.f;  // carry out of the block
  int th = 0;
#pragma unroll
  for (int k = 0; k < NW; ++k) {
    const int fk = s_f[k];
    const float vk = s_v[k];
    const int hk = s_h[k];
    if (k < w) {
      pv = fk ? vk : pv + vk;
      ph = fk ? hk : (ph | hk);
    }
    tv = fk ? vk : tv + vk;
    th = fk ? hk : (th | hk);
  }
  const float incl_v = fi ? vi : pv + vi;
  const int incl_h = fi ? hi : (ph | hi);
  *cin = dpp_f32<kDppWaveShr1, 0xf>(pv, incl_v);
  *hin = dpp_i32<kDppWaveShr1, 0xf>(ph, incl_h);
  *tail_v = tv;
  *tail_h = th;
}

constexpr int kXPlain = 0, kXEnc = 1, kXBits = 2;

// One merge-path tile: rows i0..i1-1 end in it, nonzeros j0..j0+nj-1.
struct RowTile {
  int i0, i1, nr, nj, j0;
};

__device__ __forceinline__ RowTile row_tile(const int* __restrict__ split, int t, int T, long long W) {
  RowTile g;
  const long long d0 = (long long)t * T;
  const long long d1 = d0 + T < W ? d0 + T : W;
  g.i0 = split[t];
  g.i1 = split[t + 1];
  g.nr = g.i1 - g.i0;
  g.j0 = (int)(d0 - g.i0);
  g.nj = (int)(d1 - g.i1) - g.j0;
  return g;
}

// Nonzero-and-row balanced SpMV / masked SpMSpV over merge-path tiles of
// T = NT*VT merged items (row ends + nonzeros), software-pipelined three
// tiles deep so that no memory latency sits between two tiles of a
// workgroup:
//   tile t+2: column indices and values are loading,
//   tile t+1: x is gathered and the tile's row pointers are loading,
//   tile t  : products go to LDS, each row's end is scattered to the
//             position of its last nonzero, every thread sums its VT
//             consecutive products, and a block-wide segmented scan joins
//             rows that cross threads.
// XIN selects what is gathered per nonzero:
//   kXPlain: x values (support not needed: SpMV with dense output),
//   kXEnc  : the support-encoded x (+0.0 outside the support, a nonzero bit
//            pattern inside it, an explicit zero stored as -0.0), so one
//            gather yields both the value and the support bit,
//   kXBits : only x's bit array (boolean semiring of the BFS levels): no
//            values are read (val may be null), a row's result is its hit
//            bit alone and y is not written (may be null).
// Loads use a per-tile base pointer with 32-bit offsets and are
// unconditional (clamped) so that a stage's loads issue back to back.
// Rows that cross tiles leave their tail in carry_*[t] (row i1), which
// k_row_fixup adds in tile order.
template <int NT, int VT, int XIN, bool MASKED, bool HITS>
__global__ __launch_bounds__(NT) void k_row_merge(const int* __restrict__ row_ptr,
                                                  const int* __restrict__ col,
                                                  const float* __restrict__ val,
                                                  const float* __restrict__ x,
                                                  const unsigned* __restrict__ xbits, int m, int nnz,
                                                  const int* __restrict__ split, int ntiles,
                                                  float* __restrict__ y, unsigned* ybits,
                                                  float* __restrict__ carry_val,
                                                  int* __restrict__ carry_row,
                                                  int* __restrict__ carry_hit, Gate g) {
  if (gate_shut(g)) return;
  constexpr int T = NT * VT;
  constexpr int NW = NT / kWave;
  __shared__ float s_p[T];
  __shared__ int s_end[T];  // row whose last nonzero sits here, or -1
  __shared__ unsigned char s_ph[HITS ? T : 1];
  __shared__ unsigned char s_yh[HITS ? T : 1];
  __shared__ int s_f[NW];
  __shared__ float s_sv[NW];
  __shared__ int s_sh[NW];
  const unsigned tid = threadIdx.x;
  const int tstep = gridDim.x;
  const long long W = (long long)m + nnz;
  int t = blockIdx.x;
  if (t >= ntiles) return;

#pragma unroll
  for (int i = 0; i < VT; ++i) s_end[tid + i * NT] = -1;

  int cN[VT];  // stage 2: columns / values loading
  float vN[VT];
  float vG[VT], xG[VT];  // stage 1: values, gathered x (or bit words), rows
  int cG[VT];
  int rs[VT], re[VT];

  auto load_cv = [&](int tt) {
    const RowTile g = row_tile(split, tt, T, W);
    // past the data (a tile of row ends only) the last nonzero is reloaded
    const int jb = g.j0 < nnz ? g.j0 : nnz - 1;
    const unsigned lim = g.nj > 0 ? (unsigned)(g.nj - 1) : 0u;
    const int* __restrict__ ct = col + jb;
#pragma unroll
    for (int i = 0; i < VT; ++i) {
      unsigned k = tid + i * NT;
      k = k < lim ? k : lim;
      cN[i] = __builtin_nontemporal_load(ct + k);
      if (XIN != kXBits) vN[i] = __builtin_nontemporal_load(val + jb + k);
    }
  };
  auto gather = [&](int tt) {
    const RowTile g = row_tile(split, tt, T, W);
#pragma unroll
    for (int i = 0; i < VT; ++i) {
      if (XIN == kXBits) xG[i] = __builtin_bit_cast(float, xbits[(unsigned)cN[i] >> 5]);
      else xG[i] = x[(unsigned)cN[i]];
    }
    const unsigned lim = (unsigned)(m - 1 - g.i0);
    const int* __restrict__ rt = row_ptr + g.i0;
#pragma unroll
    for (int i = 0; i < VT; ++i) {
      unsigned k = tid + i * NT;
      k = k < lim ? k : lim;
      rs[i] = rt[k];
      re[i] = rt[k + 1];
    }
    // a real move rather than a rename, so that the stage-2 and stage-1
    // registers keep their roles from tile to tile and the loop needs no
    // copy of a register whose load is in flight (a copy waits for it)
#pragma unroll
    for (int i = 0; i < VT; ++i) {
      if (XIN != kXBits) asm volatile("v_mov_b32 %0, %1" : "=v"(vG[i]) : "v"(vN[i]));
      else asm volatile("v_mov_b32 %0, %1" : "=v"(cG[i]) : "v"(cN[i]));
    }
  };

  // Stages run unconditionally (past the last tile they reload it, from
  // cache): a conditional stage would also make the compiler copy
  // registers whose loads are still in flight.
  load_cv(t);
  gather(t);
  load_cv(min(t + tstep, ntiles - 1));
  __syncthreads();

  for (; t < ntiles; t += tstep) {
    const RowTile g = row_tile(split, t, T, W);
    // products of tile t
#pragma unroll
    for (int i = 0; i < VT; ++i) {
      const unsigned k = tid + i * NT;
      bool in;
      float p;
      if (XIN == kXBits) {
        in = (__builtin_bit_cast(unsigned, xG[i]) >> (cG[i] & 31)) & 1u;
        p = 0.f;
      } else {
        in = __builtin_bit_cast(unsigned, xG[i]) != 0u;
        p = vG[i] * xG[i];
        if (MASKED) p = in ? p : 0.f;
      }
      if (k < (unsigned)g.nj) {
        s_p[k] = p;
        if (HITS) s_ph[k] = (unsigned char)in;
      }
    }
    // row ends of tile t; a row with no nonzero in the tile (empty, or row
    // i0 whose last nonzero lies in an earlier tile) is stored here as 0
    {
      float* __restrict__ yt = y + g.i0;
#pragma unroll
      for (int i = 0; i < VT; ++i) {
        const unsigned k = tid + i * NT;
        if (k < (unsigned)g.nr) {
          if (re[i] > rs[i] && re[i] > g.j0) {
            s_end[re[i] - 1 - g.j0] = g.i0 + (int)k;
          } else {
            if (XIN != kXBits) yt[k] = 0.f;
            if (HITS) s_yh[k] = 0;
          }
        }
      }
    }
    // start the next two stages; the scheduling fences keep each stage's
    // registers from overlapping the next stage's
    __builtin_amdgcn_sched_barrier(0);
    gather(min(t + tstep, ntiles - 1));
    __builtin_amdgcn_sched_barrier(0);
    load_cv(min(t + 2 * tstep, ntiles - 1));
    __builtin_amdgcn_sched_barrier(0);
    __syncthreads();

    // each thread sums its VT consecutive products
    const unsigned b = tid * VT;
    float pv[VT];
    int ev[VT];
    unsigned char hv[VT];
#pragma unroll
    for (int q = 0; q < VT; ++q) {
      pv[q] = s_p[b + q];
      ev[q] = s_end[b + q];
      if (HITS) hv[q] = s_ph[b + q];
    }
#pragma unroll
    for (int q = 0; q < VT; ++q) s_end[b + q] = -1;  // ready for the next tile
    bool had_end = false;
    float first_sum = 0.f, sum = 0.f;
    int first_hit = 0, hit = 0, first_row = 0;
#pragma unroll
    for (int q = 0; q < VT; ++q) {
      if (b + q < (unsigned)g.nj) {
        sum += pv[q];
        if (HITS) hit |= hv[q];
        const int r = ev[q];
        if (r >= 0) {
          if (!had_end) {
            first_row = r;
            first_sum = sum;
            first_hit = hit;
            had_end = true;
          } else {
            if (XIN != kXBits) y[(unsigned)r] = sum;
            if (HITS) s_yh[r - g.i0] = (unsigned char)hit;
          }
          sum = 0.f;
          hit = 0;
        }
      }
    }
    float cin, tail_v;
    int hin, tail_h;
    seg_scan_row<NT>(had_end, sum, hit, &cin, &hin, &tail_v, &tail_h, s_f, s_sv, s_sh);
    if (had_end) {
      if (XIN != kXBits) y[(unsigned)first_row] = cin + first_sum;
      if (HITS) s_yh[first_row - g.i0] = (unsigned char)(hin | first_hit);
    }
    if (tid == 0) {
      carry_row[t] = (g.i1 < m) ? g.i1 : -1;
      carry_val[t] = tail_v;
      carry_hit[t] = tail_h;
    }
    if (HITS) {
      __syncthreads();
      // hit bits of rows [i0, i1), one ballot per 64 rows from a 64-aligned
      // base: a word wholly inside the range belongs to this tile alone and
      // is stored; the two edge words are OR-ed in
      const int rb = g.i0 & ~63;
      for (int base = rb + (int)(tid & ~63u); base < g.i1; base += NT) {
        const int r = base + lane();
        const bool f = r >= g.i0 && r < g.i1 && s_yh[r - g.i0];
        const unsigned long long bal = __ballot(f);
        if (lane() == 0 || lane() == 32) {
          const int wr = base + lane();  // first row of this word
          const unsigned bits = lane() == 0 ? (unsigned)bal : (unsigned)(bal >> 32);
          if (wr >= g.i0 && wr + 32 <= g.i1) ybits[wr >> 5] = bits;
          else if (bits) atomicOr(ybits + (wr >> 5), bits);
        }
      }
      __syncthreads();
    }
  }
}

template <bool HITS>
__global__ void k_row_fixup(const float* __restrict__ carry_val, const int* __restrict__ carry_row,
                            const int* __restrict__ carry_hit, int ntiles, float* y, unsigned* ybits,
                            Gate g) {
  if (gate_shut(g)) return;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < ntiles; t += gridDim.x * blockDim.x) {
    int r, h;
    float s;
    if (!carry_run<int>(carry_val, carry_row, HITS ? carry_hit : nullptr, t, ntiles, r, s, h))
      continue;
    if (y) y[r] += s;  // null: hit bits only (boolean rows)
    if (HITS && h) atomicOr(ybits + (r >> 5), 1u << (r & 31));
  }
}

int resident_blocks(const void* kernel, int block) {
  static std::mutex mu;
  static std::map<std::pair<const void*, int>, int> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 1024;
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find({kernel, dev});
  if (it != cache.end()) return it->second;
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      per_cu <= 0 || cus <= 0)
    return 1024;
  const int n = per_cu * cus;
  cache[{kernel, dev}] = n;
  return n;
}

void launch_row_merge(hipStream_t s, bool masked, bool hits, const int* row_ptr, const int* col,
                      const float* val, const float* x, const unsigned* xbits, int m, int nnz,
                      const int* split, int ntiles, int vt, float* y, unsigned* ybits,
                      float* carry_val, int* carry_row, int* carry_hit, hipEvent_t ev_begin,
                      hipEvent_t ev_end, Gate g) {
  if (m <= 0 || ntiles <= 0) return;
  if (nnz == 0) {  // every row is empty; no hits
    if (y) launch_fill_f32(s, y, m, 0.f, g);
    return;
  }
  // one resident wave of workgroups, each walking many tiles
#define SPMSPV_ROW_MERGE(V, XI, MK, HT)                                                          \
  {                                                                                             \
    const void* fn = (const void*)k_row_merge<kRowNT, V, XI, MK, HT>;                           \
    hipLaunchKernelGGL((k_row_merge<kRowNT, V, XI, MK, HT>),                                    \
                       dim3(std::min(ntiles, resident_blocks(fn, kRowNT))), dim3(kRowNT), 0, s, \
                       row_ptr, col, val, x, xbits, m, nnz, split, ntiles, y, ybits, carry_val, \
                       carry_row, carry_hit, g);                                                \
  }
#define SPMSPV_ROW_MERGE_VT(V)                                   \
  if (!x) SPMSPV_ROW_MERGE(V, kXBits, true, true)                \
  else if (masked && hits) SPMSPV_ROW_MERGE(V, kXEnc, true, true) \
  else if (masked) SPMSPV_ROW_MERGE(V, kXEnc, true, false)       \
  else if (hits) SPMSPV_ROW_MERGE(V, kXEnc, false, true)         \
  else SPMSPV_ROW_MERGE(V, kXPlain, false, false)
  if (ev_begin) (void)hipEventRecord(ev_begin, s);
  if (vt == 5) {
    SPMSPV_ROW_MERGE_VT(5)
  } else if (vt == 11) {
    SPMSPV_ROW_MERGE_VT(11)
  } else {
    SPMSPV_ROW_MERGE_VT(7)
  }
#undef SPMSPV_ROW_MERGE_VT
#undef SPMSPV_ROW_MERGE
  if (ev_end) (void)hipEventRecord(ev_end, s);
  const int fgrid = std::min((ntiles + 255) / 256, 1024);
  if (hits)
    hipLaunchKernelGGL(k_row_fixup<true>, dim3(fgrid), dim3(256), 0, s, carry_val, carry_row,
                       carry_hit, ntiles, y, ybits, g);
  else
    hipLaunchKernelGGL(k_row_fixup<false>, dim3(fgrid), dim3(256), 0, s, carry_val, carry_row,
                       carry_hit, ntiles, y, ybits, g);
}

// BOOL (val and x null): the boolean semiring of the BFS levels; a row's
// result is its hit bit alone (no values read, y not written).
template <int NT, bool MASKED, bool HITS, bool BOOL>
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
        if (!BOOL) v[u] = __builtin_nontemporal_load(val + j + u * team);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (BOOL) {
          h |= test_bit(xbits, c[u]) ? 1 : 0;
        } else if (MASKED) {
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
      if (BOOL) {
        h |= test_bit(xbits, c) ? 1 : 0;
      } else if (MASKED) {
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
      if (!BOOL) s += __shfl_xor(s, o, kWave);
      if (HITS || BOOL) h |= __shfl_xor(h, o, kWave);
    }
    if (tl == 0) {
      if (!BOOL) y[r] = s;
      if ((HITS || BOOL) && h) atomicOr(ybits + (r >> 5), 1u << (r & 31));
    }
  }
}

// val == nullptr and x == nullptr: the boolean form (hit bits of the rows
// that touch xbits' support, into ybits; y unused).
void launch_row_naive(hipStream_t s, bool masked, bool hits, const int* row_ptr, const int* col,
                      const float* val, const float* x, const unsigned* xbits, int m, int team,
                      float* y, unsigned* ybits) {
  if (m <= 0) return;
  int lg = 0;
  while ((1 << lg) < team && lg < 6) ++lg;
  const int tpb = 256 >> lg;
  const int grid = (int)std::min<long long>(16384, ((long long)m + tpb - 1) / tpb);
#define SPMSPV_ROW_NAIVE(MK, HT, BO)                                                                   \
  hipLaunchKernelGGL((k_row_naive<256, MK, HT, BO>), dim3(grid), dim3(256), 0, s, row_ptr, col, val, x, \
                     xbits, m, lg, y, ybits)
  if (!val && !x) SPMSPV_ROW_NAIVE(true, true, true);
  else if (masked && hits) SPMSPV_ROW_NAIVE(true, true, false);
  else if (masked) SPMSPV_ROW_NAIVE(true, false, false);
  else if (hits) SPMSPV_ROW_NAIVE(false, true, false);
  else SPMSPV_ROW_NAIVE(false, false, false);
#undef SPMSPV_ROW_NAIVE
}

}  // namespace spmspv
