// Column-driven SpMSpV building blocks for gfx950:
//  * k_prep: column-length gather + single-pass exclusive scan (decoupled
//    look-back), bin_len/max/min, merge-path splits, and (fused in the last
//    workgroup) the selector features + MLP inference.
//    Replaces ScanFuseDriver/scanFuseKernel/scanFuseTail
//    (la/include/spmspv/csc-spmspv/detail/device/csc_sort_scan_fuse.cuh:8-80,
//    scan_kernel.cuh:12-269) and reduceBinlen/computeBinlenVer2
//    (spmspv.h:2404-2721). Unlike the reference's spin on a predecessor sum
//    of 0 (scan_kernel.cuh:99), the look-back waits on a status flag, so
//    empty columns cannot hang it.
//  * k_naive_col: team-per-x-entry gather (ExtractFuseKernel / ExtractKernel,
//    csc_sort_spmspv.cuh:35-119), atomics or bins.
//  * k_lb_gather: merge-path load-balanced gather over (x entries + products)
//    (KernelIntervalMoveIndirectFuseAndReductionModify[_write_back],
//    csc_sort_load_balance_fuse_reduction_modify.cuh:6-173).
//  * k_compact: hit bitmap -> sorted (key, val), resetting the accumulator
//    (replaces dense2sparse_inner's flag/scan/compact over m,
//    spmspv.h:1882-1928, and keeps the structural support).
#include <climits>

#include "kernels.hpp"
#include "selector.hpp"

namespace spmspv {

// ------------------------------------------------------------------ prep
template <int NT, int IPT>
__global__ __launch_bounds__(NT) void k_prep(const int* __restrict__ colptr,
                                             const int* __restrict__ xkey, int ncols, int xnnz_arg,
                                             const int* __restrict__ xnnz_dev, int* __restrict__ ptr,
                                             int* __restrict__ xbeg, int* __restrict__ split,
                                             unsigned long long* state, int* tile_mm, Ctl* ctl,
                                             SelectorArgs sel) {
  __shared__ unsigned s_tile;
  __shared__ unsigned long long s_pre;
  __shared__ long long s_l[NT / kWave];
  __shared__ int s_i[NT / kWave];
  __shared__ unsigned s_flag;
  const int xnnz = xnnz_dev ? *xnnz_dev : xnnz_arg;
  const unsigned ntiles = (unsigned)((xnnz + NT * IPT - 1) / (NT * IPT));
  const unsigned epoch = ctl->epoch;
  while (true) {
    const unsigned t = next_tile(ctl, kSlotPrep, &s_tile);
    if (t >= ntiles) break;
    const long long base = (long long)t * (NT * IPT) + (long long)threadIdx.x * IPT;
    int len[IPT], beg[IPT];
    long long tsum = 0;
    int mx = 0, mn = INT_MAX;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const long long i = base + j;
      len[j] = 0;
      beg[j] = 0;
      if (i < xnnz) {
        const int c = xkey[i];
        // an out-of-range key contributes nothing instead of a wild read
        const bool ok = (unsigned)c < (unsigned)ncols;
        const int b = ok ? colptr[c] : 0;
        const int e = ok ? colptr[c + 1] : 0;
        beg[j] = b;
        len[j] = e - b;
        mx = max(mx, len[j]);
        mn = min(mn, len[j]);
      }
      tsum += len[j];
    }
    long long btot;
    const long long bex = block_excl_sum<NT, long long>(tsum, s_l, &btot);
    const unsigned long long pre = tile_prefix(state, t, epoch, (unsigned long long)btot, &s_pre);
    long long run = (long long)pre + bex;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const long long i = base + j;
      if (i < xnnz) {
        ptr[i] = (int)run;
        xbeg[i] = beg[j];
        if (split) {
          const long long Mi = i + run;
          const long long Mn = Mi + len[j] + 1;
          for (long long d = (Mi / kLbTile + 1) * kLbTile; d <= Mn; d += kLbTile)
            split[d / kLbTile] = (int)(i + 1);
        }
        run += len[j];
        if (i == xnnz - 1) {
          ptr[xnnz] = (int)run;
          ctl->bin_len = run;
        }
      }
    }
    mx = block_max<NT, int>(mx, s_i);
    mn = block_min<NT, int>(mn, s_i);
    if (threadIdx.x == 0) {
      tile_mm[2 * t] = mx;
      tile_mm[2 * t + 1] = mn;
    }
  }
  const bool last = finish_persistent(ctl, kSlotPrep, &s_flag);
  if (!last) return;
  // Final reductions by the last workgroup; every other workgroup's writes
  // are visible here (finish_persistent fences).
  int mx = 0, mn = INT_MAX;
  for (unsigned t = threadIdx.x; t < ntiles; t += NT) {
    mx = max(mx, tile_mm[2 * t]);
    mn = min(mn, tile_mm[2 * t + 1]);
  }
  mx = block_max<NT, int>(mx, s_i);
  mn = block_min<NT, int>(mn, s_i);
  if (threadIdx.x == 0) {
    long long bl = 0;
    if (xnnz == 0) {
      ptr[0] = 0;
      ctl->bin_len = 0;
    } else {
      bl = __hip_atomic_load(&ctl->bin_len, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    ctl->max_len = xnnz ? mx : 0;
    ctl->min_len = mn;
    ctl->xnnz = xnnz;
    const long long W = (long long)xnnz + bl;
    const long long nt = (W + kLbTile - 1) / kLbTile;
    ctl->n_tiles = nt;
    if (split) {
      split[0] = 0;
      split[nt] = xnnz;
    }
  }
  __syncthreads();
  if (sel.mat) select_in_block<NT>(ctl, sel);
}

void launch_prep(hipStream_t s, const int* colptr, const int* xkey, int ncols, int xnnz,
                 const int* xnnz_dev,
                 int* ptr, int* xbeg, int* split, unsigned long long* state, int* tile_mm, Ctl* ctl,
                 const SelectorArgs* sel, int grid_hint) {
  int grid = kPersistGrid;
  if (!xnnz_dev) grid = max(1, min(kPersistGrid, (xnnz + kPrepTile - 1) / kPrepTile));
  if (grid_hint > 0) grid = min(grid, grid_hint);
  SelectorArgs a{};
  if (sel) a = *sel;
  hipLaunchKernelGGL((k_prep<kPrepNT, kPrepIPT>), dim3(grid), dim3(kPrepNT), 0, s, colptr, xkey,
                     ncols, xnnz, xnnz_dev, ptr, xbeg, split, state, tile_mm, ctl, a);
}

// ------------------------------------------------------------- naive col
// mode 0: atomics into acc + hit bitmap; mode 1: bins at ptr[e] + offset.
template <int NT, int MODE>
__global__ __launch_bounds__(NT) void k_naive_col(const int* __restrict__ colptr,
                                                  const int* __restrict__ crow,
                                                  const float* __restrict__ cval,
                                                  const int* __restrict__ xkey,
                                                  const float* __restrict__ xval,
                                                  const int* __restrict__ ptr, int ncols, int xnnz_arg,
                                                  const int* __restrict__ xnnz_dev, int lg_team,
                                                  float* acc, unsigned* bitmap, int* __restrict__ bkey,
                                                  float* __restrict__ bval) {
  const int xnnz = xnnz_dev ? *xnnz_dev : xnnz_arg;
  const int team = 1 << lg_team;
  const int tpb = NT >> lg_team;
  const int tid_t = threadIdx.x & (team - 1);
  const int tm = threadIdx.x >> lg_team;
  for (long long e = (long long)blockIdx.x * tpb + tm; e < xnnz; e += (long long)gridDim.x * tpb) {
    const int c = xkey[e];
    if ((unsigned)c >= (unsigned)ncols) continue;
    const float xv = xval[e];
    const int b = colptr[c];
    const int en = colptr[c + 1];
    if (MODE == 0) {
      for (int j = b + tid_t; j < en; j += team) {
        const int r = crow[j];
        atomic_add_f32(acc + r, cval[j] * xv);
        atomicOr(bitmap + (r >> 5), 1u << (r & 31));
      }
    } else {
      const int o = ptr[e] - b;
      for (int j = b + tid_t; j < en; j += team) {
        bkey[o + j] = crow[j];
        bval[o + j] = cval[j] * xv;
      }
    }
  }
}

void launch_naive_col(hipStream_t s, int mode, const int* colptr, const int* crow, const float* cval,
                      const int* xkey, const float* xval, const int* ptr, int ncols, int xnnz,
                      const int* xnnz_dev, int team, float* acc, unsigned* bitmap, int* bkey,
                      float* bval) {
  int lg = 0;
  while ((1 << lg) < team && lg < 6) ++lg;
  const int tpb = 256 >> lg;
  long long need = xnnz_dev ? (long long)kPersistGrid * tpb : xnnz;
  int grid = (int)std::min<long long>(4096, (need + tpb - 1) / tpb);
  if (grid < 1) grid = 1;
  if (mode == 0)
    hipLaunchKernelGGL((k_naive_col<256, 0>), dim3(grid), dim3(256), 0, s, colptr, crow, cval, xkey,
                       xval, ptr, ncols, xnnz, xnnz_dev, lg, acc, bitmap, bkey, bval);
  else
    hipLaunchKernelGGL((k_naive_col<256, 1>), dim3(grid), dim3(256), 0, s, colptr, crow, cval, xkey,
                       xval, ptr, ncols, xnnz, xnnz_dev, lg, acc, bitmap, bkey, bval);
}

// ------------------------------------------------------------ LB gather
// Merged sequence: for x entry i a marker at M_i = i + ptr[i], then its
// ptr[i+1]-ptr[i] products. Tile t owns merged positions [t*T, (t+1)*T);
// split[t] = number of markers before t*T (written by k_prep).
template <int NT, int VT, int MODE>
__global__ __launch_bounds__(NT) void k_lb_gather(const int* __restrict__ ptr,
                                                  const int* __restrict__ xbeg,
                                                  const float* __restrict__ xval,
                                                  const int* __restrict__ split,
                                                  const int* __restrict__ crow,
                                                  const float* __restrict__ cval,
                                                  const Ctl* __restrict__ ctl, int xnnz_arg,
                                                  const int* __restrict__ xnnz_dev, float* acc,
                                                  unsigned* bitmap, int* __restrict__ bkey,
                                                  float* __restrict__ bval) {
  constexpr int T = NT * VT;
  __shared__ int s_ptr[T + 2];
  __shared__ int s_beg[T + 2];
  __shared__ float s_xv[T + 2];
  __shared__ int s_seg[T];
  const int xnnz = xnnz_dev ? *xnnz_dev : xnnz_arg;
  const long long ntiles = ctl->n_tiles;
  const long long W = (long long)xnnz + ctl->bin_len;
  for (long long t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const long long d0 = t * T;
    const long long d1 = min(d0 + T, W);
    const int a0 = split[t];
    const int a1 = split[t + 1];
    const long long b0 = d0 - a0;
    const long long b1 = d1 - a1;
    const int sa = a0 > 0 ? a0 - 1 : 0;
    const int nseg = a1 - sa + 1;
    for (int k = threadIdx.x; k < nseg; k += NT) {
      const int i = sa + k;
      s_ptr[k] = ptr[i];
      if (i < xnnz) {
        s_beg[k] = xbeg[i];
        s_xv[k] = xval[i];
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
    for (int q = threadIdx.x; q < np; q += NT) {
      const int sg = s_seg[q];
      const long long p = b0 + q;
      const int idx = s_beg[sg] + (int)(p - s_ptr[sg]);
      const int r = crow[idx];
      const float v = cval[idx] * s_xv[sg];
      if (MODE == 0) {
        atomic_add_f32(acc + r, v);
        atomicOr(bitmap + (r >> 5), 1u << (r & 31));
      } else {
        bkey[p] = r;
        bval[p] = v;
      }
    }
    __syncthreads();
  }
}

void launch_lb_gather(hipStream_t s, int mode, const int* ptr, const int* xbeg, const float* xval,
                      const int* split, const int* crow, const float* cval, const Ctl* ctl,
                      int xnnz, const int* xnnz_dev, float* acc, unsigned* bitmap, int* bkey,
                      float* bval, long long tiles_hint) {
  long long grid = tiles_hint > 0 ? std::min<long long>(tiles_hint, 8192) : 2048;
  if (grid < 1) grid = 1;
  if (mode == 0)
    hipLaunchKernelGGL((k_lb_gather<kLbNT, kLbVT, 0>), dim3((unsigned)grid), dim3(kLbNT), 0, s, ptr,
                       xbeg, xval, split, crow, cval, ctl, xnnz, xnnz_dev, acc, bitmap, bkey, bval);
  else
    hipLaunchKernelGGL((k_lb_gather<kLbNT, kLbVT, 1>), dim3((unsigned)grid), dim3(kLbNT), 0, s, ptr,
                       xbeg, xval, split, crow, cval, ctl, xnnz, xnnz_dev, acc, bitmap, bkey, bval);
}

// -------------------------------------------------------------- compact
template <int NT, int WPT>
__global__ __launch_bounds__(NT) void k_compact(unsigned* __restrict__ bitmap, int m,
                                                float* __restrict__ acc, int reset_acc,
                                                int* __restrict__ ykey, float* __restrict__ yval,
                                                unsigned long long* state, Ctl* ctl) {
  __shared__ unsigned s_tile;
  __shared__ unsigned long long s_pre;
  __shared__ int s_i[NT / kWave];
  __shared__ unsigned s_flag;
  const int nwords = (m + 31) >> 5;
  const unsigned ntiles = (unsigned)((nwords + NT * WPT - 1) / (NT * WPT));
  const unsigned epoch = ctl->epoch;
  while (true) {
    const unsigned t = next_tile(ctl, kSlotCompact, &s_tile);
    if (t >= ntiles) break;
    const int w0 = (int)t * (NT * WPT) + threadIdx.x * WPT;
    unsigned wd[WPT];
    int cnt = 0;
#pragma unroll
    for (int j = 0; j < WPT; ++j) {
      wd[j] = (w0 + j < nwords) ? bitmap[w0 + j] : 0u;
      cnt += __popc(wd[j]);
    }
    int btot;
    const int bex = block_excl_sum<NT, int>(cnt, s_i, &btot);
    const unsigned long long pre = tile_prefix(state, t, epoch, (unsigned long long)btot, &s_pre);
    int pos = (int)pre + bex;
#pragma unroll
    for (int j = 0; j < WPT; ++j) {
      unsigned b = wd[j];
      if (!b) continue;
      bitmap[w0 + j] = 0u;
      while (b) {
        const int bit = __ffs(b) - 1;
        b &= b - 1;
        const int r = ((w0 + j) << 5) + bit;
        if (ykey) ykey[pos] = r;
        if (yval) yval[pos] = acc[r];
        if (reset_acc) acc[r] = 0.0f;
        ++pos;
      }
    }
    if (t == ntiles - 1 && threadIdx.x == 0) ctl->ynnz = (int)(pre + btot);
  }
  const bool last = finish_persistent(ctl, kSlotCompact, &s_flag);
  if (last && threadIdx.x == 0 && ntiles == 0) ctl->ynnz = 0;
}

void launch_compact(hipStream_t s, unsigned* bitmap, int m, float* acc, int reset_acc, int* ykey,
                    float* yval, unsigned long long* state, Ctl* ctl) {
  const int nwords = (m + 31) >> 5;
  int grid = std::max(1, std::min(kPersistGrid, (nwords + kCmpTile - 1) / kCmpTile));
  hipLaunchKernelGGL((k_compact<kCmpNT, kCmpWPT>), dim3(grid), dim3(kCmpNT), 0, s, bitmap, m, acc,
                     reset_acc, ykey, yval, state, ctl);
}

// ------------------------------------------------------------ utilities
__global__ void k_bitmap_from_dense(const float* __restrict__ v, int n, unsigned* __restrict__ bits) {
  // one wave per 64 elements; element e -> bit e of words 2w, 2w+1
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool nz = (e < n) && (v[e] != 0.0f);
  const uint64_t b = __ballot(nz);
  const long long wbase = (e - lane()) >> 5;
  const int nwords = (n + 31) >> 5;
  if (lane() == 0 && wbase < nwords) bits[wbase] = (unsigned)(b & 0xffffffffu);
  if (lane() == 32 && wbase + 1 < nwords) bits[wbase + 1] = (unsigned)(b >> 32);
}

void launch_bitmap_from_dense(hipStream_t s, const float* v, int n, unsigned* bits) {
  if (n <= 0) return;
  const int grid = (n + 255) / 256;
  hipLaunchKernelGGL(k_bitmap_from_dense, dim3(grid), dim3(256), 0, s, v, n, bits);
}

__global__ void k_fill_f32(float* p, long long n, float v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_i32(int* p, long long n, int v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}
void launch_fill_f32(hipStream_t s, float* p, long long n, float v) {
  if (n <= 0) return;
  const long long grid = std::min<long long>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)grid), dim3(256), 0, s, p, n, v);
}
void launch_fill_i32(hipStream_t s, int* p, long long n, int v) {
  if (n <= 0) return;
  const long long grid = std::min<long long>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)grid), dim3(256), 0, s, p, n, v);
}

__global__ void k_scatter(const int* __restrict__ key, const float* __restrict__ val, int nnz_arg,
                          const int* nnz_dev, int len, float* __restrict__ dense) {
  const int nnz = nnz_dev ? *nnz_dev : nnz_arg;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += gridDim.x * blockDim.x) {
    const int k = key[i];
    if ((unsigned)k < (unsigned)len) dense[k] = val[i];
  }
}
void launch_scatter(hipStream_t s, const int* key, const float* val, int nnz, const int* nnz_dev,
                    int len, float* dense) {
  int grid = nnz_dev ? 2048 : std::max(1, std::min(8192, (nnz + 255) / 256));
  if (!nnz_dev && nnz <= 0) return;
  hipLaunchKernelGGL(k_scatter, dim3(grid), dim3(256), 0, s, key, val, nnz, nnz_dev, len, dense);
}

__global__ void k_set_bits(const int* __restrict__ key, int nnz_arg, const int* nnz_dev, int len,
                           unsigned* bits) {
  const int nnz = nnz_dev ? *nnz_dev : nnz_arg;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += gridDim.x * blockDim.x) {
    const int k = key[i];
    if ((unsigned)k >= (unsigned)len) continue;
    atomicOr(bits + (k >> 5), 1u << (k & 31));
  }
}
void launch_set_bits(hipStream_t s, const int* key, int nnz, const int* nnz_dev, int len,
                     unsigned* bits) {
  int grid = nnz_dev ? 2048 : std::max(1, std::min(8192, (nnz + 255) / 256));
  if (!nnz_dev && nnz <= 0) return;
  hipLaunchKernelGGL(k_set_bits, dim3(grid), dim3(256), 0, s, key, nnz, nnz_dev, len, bits);
}

// Undo a sparse x written into a dense vector and bit array (restores the
// all-zero invariant of the handle's internal buffers in O(xnnz)).
__global__ void k_clear_keys(const int* __restrict__ key, int nnz, float* dense, unsigned* bits) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += gridDim.x * blockDim.x) {
    const int k = key[i];
    if (dense) dense[k] = 0.0f;
    if (bits) bits[k >> 5] = 0u;
  }
}
void launch_clear_keys(hipStream_t s, const int* key, int nnz, float* dense, unsigned* bits) {
  if (nnz <= 0) return;
  int grid = std::max(1, std::min(8192, (nnz + 255) / 256));
  hipLaunchKernelGGL(k_clear_keys, dim3(grid), dim3(256), 0, s, key, nnz, dense, bits);
}

}  // namespace spmspv
