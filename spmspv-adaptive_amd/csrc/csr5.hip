// CSR5-style SpMV / masked SpMSpV for 64-lane wavefronts (the spmv_type 1
// path of K5/K7; the reference's csr5_spmv_cuda.h:440-582 and
// csr5_spmspv.h:449-610 with CSR5Preprocess, spmspv.h:1425-1582, built for
// 32-lane warps). The format is re-derived for wave64:
//   * a tile is 64 lanes x S items (S = sigma in {8, 16, 24, 32}); lane l
//     owns the tile's items l*S .. l*S+S-1 in CSR order, stored in groups
//     of four (item i of lane l at tile*64*S + (i/4)*256 + l*4 + i%4) so
//     that every four item steps are one 16-byte load per lane, 1 KB per
//     wave side by side;
//   * the descriptor is one 32-bit word per lane: bit i marks a row start
//     at item i (bit 0 of lane 0 always marks the tile start). Marker
//     offsets (the reference's y_offset) come from a DPP prefix sum of the
//     popcounts and the cross-lane segmented sum (its seg_offset) from a
//     DPP segmented scan, so neither is stored;
//   * tile_info[t] = index, in the list of non-empty rows, of the row that
//     holds the tile's first item, with bit 31 set when that item continues
//     a row from the previous tile. The k-th marker of tile t is non-empty
//     row rows_ne[base + k], which replaces the reference's per-tile
//     empty_offset arrays; empty rows are written from their own list;
//   * the copy's columns are renumbered by nonzero count and every row's
//     items stored in that order, so that the densest columns come first.
// One wave per tile, no barrier inside the tile loop; each wave walks tiles
// w, w + waves, ... with the next tile's loads in flight (column indices
// and values two tiles ahead, x gathered one tile ahead). The only LDS use
// is a copy of the first x entries (the densest columns), made by each block
// at its start: one block per CU then serves most gathers from LDS.
#include <algorithm>

#include "handle.hpp"

namespace spmspv {

namespace {

constexpr int kXPlain5 = 0, kXEnc5 = 1, kXBits5 = 2;
// columns (after renumbering, the ones with the most nonzeros) whose x
// entries the tile kernel keeps in LDS: all 160 KB of a CU's LDS, shared by
// one block per CU that holds as many waves as the registers allow. At R-MAT
// s20 they take 72 % of the nonzeros, and the distinct cache lines per
// wave-wide gather fall to under a third.
constexpr int kCsr5Hot = 40960;
// threads per block of the tile kernel: 256 without the LDS copy; with it,
// all the waves a CU can hold at the register use of each width (168
// registers at S = 24: 3 waves per SIMD; 241 at S = 32: 2; <= 128 below: 4)
constexpr int csr5_nt(int S, int HOT) { return HOT == 0 ? 256 : S >= 32 ? 512 : S >= 24 ? 768 : 1024; }
static_assert(kCsr5Hot * 4 <= 160 * 1024, "LDS of one CU");

// ---- preprocessing --------------------------------------------------------
constexpr int kScanPart = 2048;

__global__ __launch_bounds__(256) void k_nonempty_part(const int* __restrict__ row_ptr, int m,
                                                       int* __restrict__ sums) {
  __shared__ int lds[4];
  const long long base = (long long)blockIdx.x * kScanPart;
  int v = 0;
#pragma unroll
  for (int k = 0; k < kScanPart / 256; ++k) {
    const long long r = base + k * 256 + threadIdx.x;
    if (r < m) v += row_ptr[r + 1] > row_ptr[r];
  }
  v = block_sum<256>(v, lds);
  if (threadIdx.x == 0) sums[blockIdx.x] = v;
}

// exclusive scan of the part sums in one workgroup; sums[nparts] = total
__global__ __launch_bounds__(256) void k_scan_parts(int* sums, int nparts) {
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
  if (threadIdx.x == 0) sums[nparts] = run;
}

// rank[r] = non-empty rows before r; rows_ne / rows_empty lists (ascending)
__global__ __launch_bounds__(256) void k_rank_rows(const int* __restrict__ row_ptr, int m,
                                                   const int* __restrict__ sums, int* __restrict__ rank,
                                                   int* __restrict__ rows_ne, int* __restrict__ rows_empty) {
  __shared__ int lds[4];
  constexpr int PT = kScanPart / 256;
  const long long base = (long long)blockIdx.x * kScanPart + (long long)threadIdx.x * PT;
  int f[PT];
  int loc = 0;
#pragma unroll
  for (int k = 0; k < PT; ++k) {
    const long long r = base + k;
    f[k] = r < m ? (row_ptr[r + 1] > row_ptr[r]) : 0;
    loc += f[k];
  }
  int tot;
  int run = sums[blockIdx.x] + block_excl_sum<256>(loc, lds, &tot);
#pragma unroll
  for (int k = 0; k < PT; ++k) {
    const long long r = base + k;
    if (r < m) {
      rank[r] = run;
      if (f[k]) rows_ne[run] = (int)r;
      else rows_empty[r - run] = (int)r;
    }
    run += f[k];
  }
}

// tile_info and the tile-start marker of every tile
__global__ void k_tile_info(const int* __restrict__ row_ptr, const int* __restrict__ rank, int m,
                            int nnz, int T, int ntiles, int n_ne, unsigned* __restrict__ tinfo,
                            unsigned* __restrict__ flags) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > ntiles) return;
  const long long j0 = (long long)t * T;
  if (j0 >= nnz) {
    tinfo[t] = (unsigned)n_ne;  // past the data: a "real start" sentinel
    return;
  }
  // the row holding item j0: first r with row_ptr[r + 1] > j0
  int lo = 0, hi = m - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (row_ptr[mid + 1] > j0) hi = mid;
    else lo = mid + 1;
  }
  const bool cont = row_ptr[lo] < j0;
  tinfo[t] = (unsigned)rank[lo] | (cont ? 0x80000000u : 0u);
  flags[(size_t)t * kWave] |= 1u;  // only this thread touches lane 0 here
}

// row-start markers (one-time preprocessing, atomics are fine)
__global__ void k_row_markers(const int* __restrict__ row_ptr, int m, int T, int S,
                              unsigned* __restrict__ flags) {
  for (long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x; r < m;
       r += (long long)gridDim.x * blockDim.x) {
    const int j = row_ptr[r];
    if (row_ptr[r + 1] <= j) continue;
    const int t = j / T, o = j % T;
    atomicOr(flags + (size_t)t * kWave + o / S, 1u << (o % S));
  }
}

__global__ void k_csr5_transpose(const int* __restrict__ col, const float* __restrict__ val, int nnz,
                                 int T, int S, long long total, const int* __restrict__ cperm,
                                 const int* __restrict__ item, int* __restrict__ ccol,
                                 float* __restrict__ cval) {
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (long long)gridDim.x * blockDim.x) {
    // position o of a tile holds item i of lane l with o = (i / 4) * 256 +
    // l * 4 + i % 4: each lane's items come in 16-byte groups of four, one
    // vector load per group, a wave's load 1 KB side by side
    const long long t = q / T;
    const int o = (int)(q - t * T);
    const int g = o >> 8, r = o & 255;
    const int l = r >> 2, i = g * 4 + (r & 3);
    const long long src = t * T + (long long)l * S + i;
    const bool ok = src < nnz;
    // with an item order, position src of the copy holds CSR item item[src]
    const long long j = ok ? (item ? (long long)item[src] : src) : 0;
    const int c = ok ? col[j] : 0;
    ccol[q] = ok ? (cperm ? cperm[c] : c) : 0;
    cval[q] = ok ? val[j] : 0.f;
  }
}

// ---- SpMV / masked SpMSpV ---------------------------------------------------
// HOT > 0: the first `hot` (<= HOT) entries of x, the columns with the most
// nonzeros once the columns are renumbered, are copied into LDS at kernel
// start and gathered from there; the other lanes of a gather read global x.
template <int S, int XIN, bool MASKED, bool HITS, int HOT>
__global__ __launch_bounds__(csr5_nt(S, HOT)) void k_csr5(const int* __restrict__ ccol,
                                              const float* __restrict__ cval,
                                              const unsigned* __restrict__ flags,
                                              const unsigned* __restrict__ tinfo,
                                              const float* __restrict__ x,
                                              const unsigned* __restrict__ xbits, int nnz, int ntiles,
                                              float* __restrict__ y, unsigned char* __restrict__ marks,
                                              float* __restrict__ carry_val,
                                              int* __restrict__ carry_row,
                                              int* __restrict__ carry_hit, int hot, int xn,
                                              float* __restrict__ clr, Gate g) {
  if (gate_shut(g)) return;
  if (clr) {
    // a buffer the caller wants all +0.0 (xn floats, 16-byte aligned)
    const int n4 = xn >> 2, gt = (int)(gridDim.x * blockDim.x);
    for (int i = (int)(blockIdx.x * blockDim.x + threadIdx.x); i < n4; i += gt)
      reinterpret_cast<float4*>(clr)[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    const int i = (n4 << 2) + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < xn) clr[i] = 0.f;
  }
  constexpr int T = kWave * S;
  const int lanes = lane();
  const int waves = gridDim.x * (blockDim.x / kWave);
  __shared__ float sx[HOT > 0 ? HOT : 1];
  // x through a buffer descriptor: one 32-bit offset per gather, and an
  // offset past x reads 0 instead of faulting
  const __amdgpu_buffer_rsrc_t xr =
      __builtin_amdgcn_make_buffer_rsrc((void*)x, (short)0, xn * 4, 0x00020000);
  if (HOT > 0) {
    // every wave of the block takes part (before any of them may leave)
    constexpr int NT = csr5_nt(S, HOT);
    constexpr int PT = HOT > 0 ? (HOT + NT - 1) / NT : 1;
    float v[PT];
#pragma unroll
    for (int k = 0; k < PT; ++k) {
      const int c = k * NT + (int)threadIdx.x;
      v[k] = x[c < hot ? c : 0];
    }
#pragma unroll
    for (int k = 0; k < PT; ++k) {
      const int c = k * NT + (int)threadIdx.x;
      if (HOT % NT == 0 || c < HOT) sx[c] = v[k];
    }
    __syncthreads();
  }
  // wave-uniform, so that per-tile words come through scalar loads
  int t = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave);
  if (t >= ntiles) return;

  int cN[S];  // stage 2: columns / values loading
  float vN[S];
  float vG[S], xG[S];  // stage 1: values and gathered x (or bit words)
  int cG[S];
  // hot items: their product (masked) is formed as soon as LDS answers and
  // kept in vG; bit i of hG marks item i hot, bit i of bG its support bit
  unsigned hG = 0u, bG = 0u;
  unsigned fG;  // the tile's marker word

  static_assert(S % 4 == 0, "items come in groups of four per lane");
  typedef int i4v __attribute__((ext_vector_type(4)));
  typedef float f4v __attribute__((ext_vector_type(4)));
  auto load_cv = [&](int tt) {
    // group g of this lane: items 4g .. 4g+3, one 16-byte load each
    const i4v* __restrict__ ct = reinterpret_cast<const i4v*>(ccol + (size_t)tt * T) + lanes;
    const f4v* __restrict__ vt = reinterpret_cast<const f4v*>(cval + (size_t)tt * T) + lanes;
#pragma unroll
    for (int g4 = 0; g4 < S / 4; ++g4) {
      const i4v c = __builtin_nontemporal_load(ct + g4 * kWave);
      const f4v v = __builtin_nontemporal_load(vt + g4 * kWave);
      cN[4 * g4] = c.x;
      cN[4 * g4 + 1] = c.y;
      cN[4 * g4 + 2] = c.z;
      cN[4 * g4 + 3] = c.w;
      vN[4 * g4] = v.x;
      vN[4 * g4 + 1] = v.y;
      vN[4 * g4 + 2] = v.z;
      vN[4 * g4 + 3] = v.w;
    }
  };
  auto gather = [&](int tt) {
    // real moves rather than renames (see the row-merge kernel): the stage
    // registers keep their roles and no in-flight load is copied. They come
    // first so that vN is free before the gathers' registers are taken (its
    // loads are a tile old; the column loads issued beside them are waited
    // for by the gathers anyway)
#pragma unroll
    for (int i = 0; i < S; ++i) {
      asm volatile("v_mov_b32 %0, %1" : "=v"(vG[i]) : "v"(vN[i]));
      if (XIN == kXBits5) asm volatile("v_mov_b32 %0, %1" : "=v"(cG[i]) : "v"(cN[i]));
    }
    unsigned hm = 0u;
    if (HOT > 0) {
#pragma unroll
      for (int i = 0; i < S; ++i) hm |= ((unsigned)cN[i] < (unsigned)hot ? 1u : 0u) << i;
    }
    float xL[HOT > 0 ? S : 1];
#pragma unroll
    for (int i = 0; i < S; ++i) {
      if (XIN == kXBits5) {
        xG[i] = __builtin_bit_cast(float, xbits[(unsigned)cN[i] >> 5]);
      } else if (HOT > 0) {
        // both reads unconditional: the hot lanes give the global read an
        // offset past the end of x (the buffer read returns 0 without a
        // memory access), the others read LDS word 0 (a broadcast)
        const unsigned c = (unsigned)cN[i];
        const bool hc = (hm >> i) & 1u;
        xG[i] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(xr, hc ? 0xFFFFFFF0u : c * 4u, 0, 0));
        xL[i] = sx[hc ? c : 0u];
      } else {
        xG[i] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(xr, (unsigned)cN[i] * 4u, 0, 0));
      }
    }
    fG = flags[(size_t)tt * kWave + lanes];
    if (HOT > 0) {
      // LDS answers within a few hundred cycles: the hot products are formed
      // here, so that xL dies at once and only two mask words live on
      unsigned bm = 0u;
#pragma unroll
      for (int i = 0; i < S; ++i) {
        const bool hc = (hm >> i) & 1u;
        const bool xb = __builtin_bit_cast(unsigned, xL[i]) != 0u;
        float p = vG[i] * xL[i];
        if (MASKED) p = xb ? p : 0.f;
        vG[i] = hc ? p : vG[i];
        bm |= (hc && xb ? 1u : 0u) << i;
      }
      hG = hm;
      bG = bm;
    }
  };

  load_cv(t);
  gather(t);
  load_cv(min(t + waves, ntiles - 1));

  for (; t < ntiles; t += waves) {
    const unsigned f = fG;
    const unsigned hh = hG, hb = bG;
    const unsigned ti = tinfo[t];
    const unsigned tn = tinfo[t + 1];
    const int base = (int)(ti & 0x7fffffffu);
    const int lim = nnz - t * T;  // items of this tile that exist (>= T except the last)

    // markers before this lane (DPP inclusive prefix of the popcounts)
    const int cnt = __popc(f);
    int incl = cnt;
    incl += dpp_i32<kDppRowShr1, 0xf>(0, incl);
    incl += dpp_i32<kDppRowShr2, 0xf>(0, incl);
    incl += dpp_i32<kDppRowShr4, 0xf>(0, incl);
    incl += dpp_i32<kDppRowShr8, 0xf>(0, incl);
    incl += dpp_i32<kDppRowBcast15, 0xa>(0, incl);
    incl += dpp_i32<kDppRowBcast31, 0xc>(0, incl);
    const int excl = incl - cnt;
    const int total = __shfl(incl, kWave - 1, kWave);

    // walk this lane's items; a segment that ends at a marker inside the
    // lane is a finished row. Products are formed on the way, so that the
    // stage-1 registers are free before the next tile is gathered into them.
    float s = 0.f, head_s = 0.f;
    int h = 0, head_h = 0, j = 0;
    bool head = true;
#pragma unroll
    for (int i = 0; i < S; ++i) {
      if ((f >> i) & 1u) {
        if (head) {
          head_s = s;
          head_h = h;
          head = false;
        } else {
          const unsigned k = (unsigned)(base + excl + j - 1);
          y[k] = s;
          if (HITS && h) marks[k] = 1;
        }
        s = 0.f;
        h = 0;
        ++j;
      }
      const bool ok = lanes * S + i < lim;
      bool b;
      float q;
      if (XIN == kXBits5) {
        b = (__builtin_bit_cast(unsigned, xG[i]) >> (cG[i] & 31)) & 1u;
        q = b ? vG[i] : 0.f;
      } else {
        b = __builtin_bit_cast(unsigned, xG[i]) != 0u;
        q = vG[i] * xG[i];
        if (MASKED) q = b ? q : 0.f;
        if (HOT > 0 && ((hh >> i) & 1u)) {  // a select, not a branch
          b = (hb >> i) & 1u;
          q = vG[i];
        }
      }
      s += ok ? q : 0.f;
      if (HITS) h |= (ok && b) ? 1 : 0;
    }
    __builtin_amdgcn_sched_barrier(0);
    gather(min(t + waves, ntiles - 1));
    __builtin_amdgcn_sched_barrier(0);
    load_cv(min(t + 2 * waves, ntiles - 1));
    __builtin_amdgcn_sched_barrier(0);

    // the segment open at the lane's end runs into the next lanes
    int fl = cnt > 0 ? 1 : 0;
    float sv = s;
    int hv = h;
    wave_seg_scan(fl, sv, hv);  // inclusive: segment open after this lane
    const float cin = dpp_f32<kDppWaveShr1, 0xf>(0.f, sv);
    const int hin = dpp_i32<kDppWaveShr1, 0xf>(0, hv);
    if (cnt > 0 && excl > 0) {
      // the head segment closes the row of the marker before this lane
      const unsigned k = (unsigned)(base + excl - 1);
      y[k] = cin + head_s;
      if (HITS && (hin | head_h)) marks[k] = 1;
    }
    if (lanes == kWave - 1) {
      // the tile's open segment: complete when the next tile starts a row
      const int k = base + total - 1;
      if (!(tn & 0x80000000u)) {
        y[(unsigned)k] = sv;
        if (HITS && hv) marks[k] = 1;
        carry_row[t] = -1;
      } else {
        carry_row[t] = k;
        carry_val[t] = sv;
        carry_hit[t] = hv;
      }
    }
  }
}

// rows that run through several tiles: their tail parts, added in tile order
__global__ void k_csr5_fixup(const float* __restrict__ carry_val, const int* __restrict__ carry_row,
                             const int* __restrict__ carry_hit, int ntiles, float* y,
                             unsigned char* marks, Gate g) {
  if (gate_shut(g)) return;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < ntiles; t += gridDim.x * blockDim.x) {
    int r, h;
    float s;
    if (!carry_run<int>(carry_val, carry_row, carry_hit, t, ntiles, r, s, h)) continue;
    y[r] += s;
    if (marks && h) marks[r] = 1;
  }
}

// dense y from the non-empty-row results: y[r] = y_ne[rank[r]] or 0
__global__ void k_csr5_expand(const int* __restrict__ row_ptr, const int* __restrict__ rank, int m,
                              const float* __restrict__ y_ne, float* __restrict__ y, Gate g) {
  if (gate_shut(g)) return;
  for (long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x; r < m;
       r += (long long)gridDim.x * blockDim.x)
    y[r] = row_ptr[r + 1] > row_ptr[r] ? y_ne[rank[r]] : 0.f;
}

}  // namespace

namespace {

// ---- column order (most nonzeros first) -------------------------------------
// Sort key of column c: nnz - count(c), so an ascending stable sort puts the
// densest columns first and keeps equal counts in index order; the column
// index rides along as the value's bits.
__global__ void k_col_keys_from_ptr(const int* __restrict__ colptr, int n, long long nnz,
                                    unsigned* __restrict__ keys, float* __restrict__ vals) {
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (long long)gridDim.x * blockDim.x) {
    keys[c] = (unsigned)(nnz - (colptr[c + 1] - colptr[c]));
    vals[c] = __int_as_float((int)c);
  }
}
__global__ void k_col_count(const int* __restrict__ col, int nnz, int n, unsigned* cnt) {
  for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < nnz;
       j += (long long)gridDim.x * blockDim.x) {
    const int c = col[j];
    if ((unsigned)c < (unsigned)n) atomicAdd(cnt + c, 1u);
  }
}
__global__ void k_col_keys_from_count(const unsigned* __restrict__ cnt, int n, long long nnz,
                                      unsigned* __restrict__ keys, float* __restrict__ vals) {
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (long long)gridDim.x * blockDim.x) {
    keys[c] = (unsigned)(nnz - (long long)cnt[c]);
    vals[c] = __int_as_float((int)c);
  }
}
__global__ void k_set_count(long long* p, long long v) { *p = v; }
// sort 1 of the item order: key = new column of item j, value = j
__global__ void k_item_keys_col(const int* __restrict__ col, int nnz, const int* __restrict__ cperm,
                                unsigned* __restrict__ keys, float* __restrict__ vals) {
  for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < nnz;
       j += (long long)gridDim.x * blockDim.x) {
    keys[j] = (unsigned)cperm[col[j]];
    vals[j] = __int_as_float((int)j);
  }
}
// sort 2: key = row of the item (binary search in row_ptr), value = item,
// written into (keys, vals_out) from the order of sort 1 (in place allowed)
__global__ void k_item_keys_row(const float* vals_in, int nnz, const int* __restrict__ row_ptr, int m,
                                unsigned* keys, float* vals_out) {
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < nnz;
       q += (long long)gridDim.x * blockDim.x) {
    const float v = vals_in[q];
    const int j = __float_as_int(v);
    int lo = 0, hi = m;  // last row r with row_ptr[r] <= j
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (row_ptr[mid] <= j) lo = mid;
      else hi = mid;
    }
    keys[q] = (unsigned)lo;
    vals_out[q] = v;
  }
}
__global__ void k_invert_order(const float* __restrict__ vals, int n, int* __restrict__ cperm) {
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x)
    cperm[__float_as_int(vals[q])] = (int)q;
}

}  // namespace

bool launch_column_order(hipStream_t s, const ColumnOrderArgs& a, const int* col, int nnz, int n,
                         int* cperm) {
  if (n <= 0 || n > a.sb_cap || nnz >= 0x7ffffffe) return false;
  const int grid = std::max(1, std::min(8192, (n + 255) / 256));
  if (a.csc_colptr) {
    hipLaunchKernelGGL(k_col_keys_from_ptr, dim3(grid), dim3(256), 0, s, a.csc_colptr, n,
                       (long long)nnz, a.sb.keyA, a.sb.valA);
  } else {
    launch_fill_i32(s, reinterpret_cast<int*>(a.sb.keyB), n, 0);
    hipLaunchKernelGGL(k_col_count, dim3(std::max(1, std::min(8192, (nnz + 255) / 256))), dim3(256), 0, s,
                       col, nnz, n, a.sb.keyB);
    hipLaunchKernelGGL(k_col_keys_from_count, dim3(grid), dim3(256), 0, s, a.sb.keyB, n, (long long)nnz,
                       a.sb.keyA, a.sb.valA);
  }
  hipLaunchKernelGGL(k_set_count, dim3(1), dim3(1), 0, s, a.n_dev, (long long)n);
  const int fb = launch_radix_sort(s, a.sb, nnz + 1, a.n_dev, a.ctl, n, n, false, a.ep);
  hipLaunchKernelGGL(k_invert_order, dim3(grid), dim3(256), 0, s, fb ? a.sb.valB : a.sb.valA, n, cperm);
  return hipGetLastError() == hipSuccess;
}

// The items of every row ordered by their new column (two stable sorts: by
// new column, then by row); returns the order (CSR item of every position)
// in the sort buffers, valid until they are reused, or null.
static const int* launch_item_order(hipStream_t s, const ColumnOrderArgs& a, const int* row_ptr,
                                    const int* col, int nnz, int m, int n, const int* cperm) {
  if (nnz <= 0 || nnz > a.sb_cap || m <= 0) return nullptr;
  const int grid = std::max(1, std::min(8192, (nnz + 255) / 256));
  hipLaunchKernelGGL(k_item_keys_col, dim3(grid), dim3(256), 0, s, col, nnz, cperm, a.sb.keyA, a.sb.valA);
  hipLaunchKernelGGL(k_set_count, dim3(1), dim3(1), 0, s, a.n_dev, (long long)nnz);
  const Ep e1{a.ep.v + 4u, -1};
  const int f1 = launch_radix_sort(s, a.sb, n, a.n_dev, a.ctl, nnz, nnz, false, e1);
  // sort 2 reads its keys from wherever sort 1 left the items, into A
  hipLaunchKernelGGL(k_item_keys_row, dim3(grid), dim3(256), 0, s, f1 ? a.sb.valB : a.sb.valA, nnz, row_ptr,
                     m, a.sb.keyA, a.sb.valA);
  const Ep e2{a.ep.v + 8u, -1};
  const int f2 = launch_radix_sort(s, a.sb, m, a.n_dev, a.ctl, nnz, nnz, false, e2);
  if (hipGetLastError() != hipSuccess) return nullptr;
  return reinterpret_cast<const int*>(f2 ? a.sb.valB : a.sb.valA);
}

int csr5_sigma_of(int sigma, int m, long long nnz) {
  int s = sigma;
  if (s <= 0) {
    // nonzeros per row -> sigma, the reference's auto rule (spmspv.h:1397-1411)
    const long long per = m > 0 ? nnz / m : 0;
    if (per <= 4) s = 4;
    else if (per <= 32) s = (int)per;
    else if (per <= 256) s = 32;
    else s = 6;
  }
  // the instantiated widths
  if (s <= 8) return 8;
  if (s <= 16) return 16;
  if (s <= 24) return 24;
  return 32;
}

size_t csr5_bytes(int m, int n, long long nnz, int S) {
  const long long T = (long long)kWave * S;
  const long long ntiles = (nnz + T - 1) / T;
  auto up = [](size_t v) { return (v + 255) & ~size_t(255); };
  size_t b = 0;
  b += up(sizeof(int) * (size_t)std::max<long long>(ntiles * T, 1)) * 2;  // ccol, cval
  b += up(sizeof(unsigned) * (size_t)(ntiles + 1) * kWave);              // flags
  b += up(sizeof(unsigned) * (size_t)(ntiles + 2));                      // tinfo
  b += up(sizeof(int) * (size_t)m) * 4;                                  // rank, rows_ne, rows_empty, y_ne
  b += up(sizeof(int) * (size_t)(m / kScanPart + 2));                    // part sums
  b += up(sizeof(float) * (size_t)(ntiles + 1)) * 3;                     // carries
  b += up(sizeof(int) * (size_t)std::max(n, 1));                         // column order
  return b;
}

hipError_t csr5_build(hipStream_t s, const int* row_ptr, const int* col, const float* val, int m,
                      int n, int nnz, int S, void* mem, int* h_pin, const ColumnOrderArgs* order,
                      Csr5* out) {
  const int T = kWave * S;
  const int ntiles = (int)(((long long)nnz + T - 1) / T);
  auto up = [](size_t v) { return (v + 255) & ~size_t(255); };
  char* p = static_cast<char*>(mem);
  Csr5 c{};
  c.S = S;
  c.row_ptr = row_ptr;
  c.ntiles = ntiles;
  c.ccol = reinterpret_cast<int*>(p);
  p += up(sizeof(int) * (size_t)std::max<long long>((long long)ntiles * T, 1));
  c.cval = reinterpret_cast<float*>(p);
  p += up(sizeof(int) * (size_t)std::max<long long>((long long)ntiles * T, 1));
  c.flags = reinterpret_cast<unsigned*>(p);
  p += up(sizeof(unsigned) * (size_t)(ntiles + 1) * kWave);
  c.tinfo = reinterpret_cast<unsigned*>(p);
  p += up(sizeof(unsigned) * (size_t)(ntiles + 2));
  int* rank = reinterpret_cast<int*>(p);
  c.rank = rank;
  p += up(sizeof(int) * (size_t)m);
  c.y_ne = reinterpret_cast<float*>(p);
  p += up(sizeof(float) * (size_t)m);
  c.rows_ne = reinterpret_cast<int*>(p);
  p += up(sizeof(int) * (size_t)m);
  c.rows_empty = reinterpret_cast<int*>(p);
  p += up(sizeof(int) * (size_t)m);
  int* sums = reinterpret_cast<int*>(p);
  p += up(sizeof(int) * (size_t)(m / kScanPart + 2));
  c.carry_val = reinterpret_cast<float*>(p);
  p += up(sizeof(float) * (size_t)(ntiles + 1));
  c.carry_row = reinterpret_cast<int*>(p);
  p += up(sizeof(float) * (size_t)(ntiles + 1));
  c.carry_hit = reinterpret_cast<int*>(p);
  p += up(sizeof(float) * (size_t)(ntiles + 1));
  int* cperm = reinterpret_cast<int*>(p);
  p += up(sizeof(int) * (size_t)std::max(n, 1));
  c.n = n;
  const int* item = nullptr;
  if (order && launch_column_order(s, *order, col, nnz, n, cperm)) {
    c.cperm = cperm;
    item = launch_item_order(s, *order, row_ptr, col, nnz, m, n, cperm);
  }

  const int parts = (m + kScanPart - 1) / kScanPart;
  if (parts > 0) {
    hipLaunchKernelGGL(k_nonempty_part, dim3(parts), dim3(256), 0, s, row_ptr, m, sums);
    hipLaunchKernelGGL(k_scan_parts, dim3(1), dim3(256), 0, s, sums, parts);
    hipLaunchKernelGGL(k_rank_rows, dim3(parts), dim3(256), 0, s, row_ptr, m, sums, rank, c.rows_ne,
                       c.rows_empty);
  }
  launch_fill_i32(s, reinterpret_cast<int*>(c.flags), (long long)(ntiles + 1) * kWave, 0);
  hipError_t e = hipMemcpyAsync(h_pin, sums + parts, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  c.n_ne = parts > 0 ? h_pin[0] : 0;
  c.n_empty = m - c.n_ne;
  if (ntiles > 0) {
    hipLaunchKernelGGL(k_tile_info, dim3((ntiles + 1 + 255) / 256), dim3(256), 0, s, row_ptr, rank, m,
                       nnz, T, ntiles, c.n_ne, c.tinfo, c.flags);
    hipLaunchKernelGGL(k_row_markers, dim3(std::min(8192, (m + 255) / 256)), dim3(256), 0, s, row_ptr,
                       m, T, S, c.flags);
    const long long total = (long long)ntiles * T;
    hipLaunchKernelGGL(k_csr5_transpose, dim3((unsigned)std::min<long long>(8192, (total + 255) / 256)),
                       dim3(256), 0, s, col, val, nnz, T, S, total, c.cperm, item, c.ccol, c.cval);
  }
  *out = c;
  return hipGetLastError();
}

void launch_csr5(hipStream_t s, const Csr5& c, bool masked, bool hits, const float* x,
                 const unsigned* xbits, int m, int nnz, float* y, bool y_dense_rows,
                 unsigned char* marks, float** vals_out, const int** keymap_out, int* n_out,
                 hipEvent_t ev_begin, hipEvent_t ev_end, Gate g, float* clear) {
  const bool compressed = c.n_empty > 0;
  float* yk = compressed ? c.y_ne : y;  // results indexed by non-empty row
  *vals_out = yk;
  *keymap_out = compressed ? c.rows_ne : nullptr;
  *n_out = compressed ? c.n_ne : m;
  if (m <= 0 || c.ntiles == 0) {
    // no tile kernel: the buffer is cleared by a fill (gated alike)
    if (clear && c.n > 0) launch_fill_f32(s, clear, c.n, 0.f, g);
    if (m > 0 && y_dense_rows) launch_fill_f32(s, y, m, 0.f, g);
    return;
  }
  // the LDS copy holds the hot columns once they come first (column order),
  // or the whole of x when it fits
  const bool lds_hot = c.n >= 1 && (c.cperm != nullptr || c.n <= kCsr5Hot);
#define SPMSPV_CSR5(SS, XI, MK, HT, HOTN)                                                        \
  {                                                                                              \
    const void* fn = (const void*)k_csr5<SS, XI, MK, HT, HOTN>;                                  \
    constexpr int nt = csr5_nt(SS, HOTN), wpb = nt / kWave;                                      \
    const int waves = std::max(wpb, std::min(c.ntiles, resident_blocks(fn, nt) * wpb));          \
    hipLaunchKernelGGL((k_csr5<SS, XI, MK, HT, HOTN>), dim3((waves + wpb - 1) / wpb), dim3(nt), 0, \
                       s,                                                                        \
                       c.ccol, c.cval, c.flags, c.tinfo, x, xbits, nnz, c.ntiles, yk, marks,     \
                       c.carry_val, c.carry_row, c.carry_hit, std::min(c.n, HOTN), c.n, clear, g); \
  }
#define SPMSPV_CSR5_H(SS, XI, MK, HT)      \
  if (lds_hot) SPMSPV_CSR5(SS, XI, MK, HT, kCsr5Hot) \
  else SPMSPV_CSR5(SS, XI, MK, HT, 0)
#define SPMSPV_CSR5_S(SS)                                          \
  if (!x) SPMSPV_CSR5(SS, kXBits5, true, true, 0)                  \
  else if (masked && hits) SPMSPV_CSR5_H(SS, kXEnc5, true, true)   \
  else if (masked) SPMSPV_CSR5_H(SS, kXEnc5, true, false)          \
  else if (hits) SPMSPV_CSR5_H(SS, kXEnc5, false, true)            \
  else SPMSPV_CSR5_H(SS, kXPlain5, false, false)
  if (ev_begin) (void)hipEventRecord(ev_begin, s);
  if (c.S == 8) {
    SPMSPV_CSR5_S(8)
  } else if (c.S == 16) {
    SPMSPV_CSR5_S(16)
  } else if (c.S == 24) {
    SPMSPV_CSR5_S(24)
  } else {
    SPMSPV_CSR5_S(32)
  }
#undef SPMSPV_CSR5_S
#undef SPMSPV_CSR5_H
#undef SPMSPV_CSR5
  if (ev_end) (void)hipEventRecord(ev_end, s);
  hipLaunchKernelGGL(k_csr5_fixup, dim3(std::min(1024, (c.ntiles + 255) / 256)), dim3(256), 0, s,
                     c.carry_val, c.carry_row, c.carry_hit, c.ntiles, yk, hits ? marks : nullptr, g);
  if (compressed && y_dense_rows)
    hipLaunchKernelGGL(k_csr5_expand, dim3(std::min(8192, (m + 255) / 256)), dim3(256), 0, s, c.row_ptr,
                       c.rank, m, c.y_ne, y, g);
}

}  // namespace spmspv
