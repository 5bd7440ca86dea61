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

// Selector inputs of a large x from kSampleKeys evenly spaced keys (the
// definition in kernels.hpp), by the whole workgroup; every thread gets
// the results.
template <int NT>
__device__ void sample_x(const int* __restrict__ colptr, const int* __restrict__ xkey, int ncols, int xnnz,
                         long long* s_l, int* s_i, long long* bl, int* mx) {
  static_assert(kSampleKeys % NT == 0, "whole samples per thread");
  constexpr int PER = kSampleKeys / NT;
  int c[PER], b[PER], e[PER];
#pragma unroll
  for (int j = 0; j < PER; ++j) c[j] = xkey[(long long)(threadIdx.x + j * NT) * xnnz / kSampleKeys];
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int ok = (unsigned)c[j] < (unsigned)ncols ? 1 : 0;
    const int ci = c[j] & -ok;
    b[j] = colptr[ci];
    e[j] = colptr[ci + ok];
  }
  long long sum = 0;
  int m = 0;
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    sum += e[j] - b[j];
    m = max(m, e[j] - b[j]);
  }
  sum = block_sum<NT, long long>(sum, s_l);
  m = block_max<NT, int>(m, s_i);
  *bl = (long long)llrint((double)sum * (double)xnnz / (double)kSampleKeys);
  *mx = m;
}

// ------------------------------------------------------------------ prep
// Ordered tiles (tile = blockIdx.x). Each tile publishes its max/min column
// length before its look-back aggregate; the last tile, whose look-back has
// seen every earlier tile publish, finalises bin_len, max, min, the merge
// tile count and the selector.
template <int NT, int IPT>
__global__ __launch_bounds__(NT) void k_prep(const int* __restrict__ colptr,
                                             const int* __restrict__ xkey, int ncols, int xnnz_arg,
                                             const int* __restrict__ xnnz_dev, int* __restrict__ ptr,
                                             int* __restrict__ xbeg, int* __restrict__ split,
                                             unsigned long long* state, int* tile_mm, Ctl* ctl,
                                             Ep ep, SelectorArgs sel, long long* samp) {
  __shared__ unsigned long long s_pre;
  __shared__ long long s_l[NT / kWave];
  __shared__ int s_i[NT / kWave];
  __shared__ __align__(16) float s_w[kNWeights];  // selector weights (selecting workgroup only)
  __shared__ MatFeat s_mf;
  const int xnnz = xnnz_dev ? *xnnz_dev : xnnz_arg;
  const unsigned ntiles = (unsigned)((xnnz + NT * IPT - 1) / (NT * IPT));
  const unsigned t = blockIdx.x;
  // The workgroup that will run the selector loads its weights and the
  // matrix features now, so that their memory latency overlaps the scan
  // instead of following it.
  const bool selects = sel.mat != nullptr && t == (ntiles == 0 ? 0u : ntiles - 1);
  constexpr int kWPT = (kNWeights + NT - 1) / NT;
  constexpr int kMF = sizeof(MatFeat) / sizeof(double);
  static_assert(sizeof(MatFeat) % sizeof(double) == 0 && kMF <= NT, "MatFeat is copied as doubles");
  float wr[kWPT];
  double mfr = 0.0;  // field threadIdx.x of MatFeat (lanes, not a struct copy)
  if (selects) {
    if (sel.weights) {
#pragma unroll
      for (int i = 0; i < kWPT; ++i) {
        const int idx = threadIdx.x + i * NT;
        wr[i] = idx < kNWeights ? sel.weights[idx] : 0.f;
      }
    }
    if (threadIdx.x < kMF) mfr = reinterpret_cast<const double*>(sel.mat)[threadIdx.x];
  }
  // weights and matrix features into LDS (visible after the next barrier)
  auto stage = [&]() {
    if (sel.weights) {
#pragma unroll
      for (int i = 0; i < kWPT; ++i) {
        const int idx = threadIdx.x + i * NT;
        if (idx < kNWeights) s_w[wlds_index(idx)] = wr[i];
      }
    }
    if (threadIdx.x < kMF) reinterpret_cast<double*>(&s_mf)[threadIdx.x] = mfr;
  };
  // after a barrier that follows stage()
  auto run_selector = [&](long long xn, long long bl, int mxl) {
    if (!selects) return;
    SelectorArgs sl = sel;
    if (sel.weights) sl.weights = s_w;
    sl.mat = &s_mf;
    select_in_wave(ctl, sl, (double)xn, (double)bl, (double)mxl);
  };
  if (ntiles == 0) {
    // empty x: block 0 alone writes the results
    if (t != 0) return;
    if (threadIdx.x == 0) {
      ptr[0] = 0;
      ctl->bin_len = 0;
      ctl->max_len = 0;
      ctl->min_len = INT_MAX;
      ctl->xnnz = 0;
      ctl->n_tiles = 0;
      if (split) split[0] = 0;
    }
    if (selects) stage();
    __syncthreads();
    run_selector(0, 0, 0);
    if (sel.stash && threadIdx.x == 0) *sel.stash = sel_word(0, sel.mat ? ctl->kernel_choice : -1);
    return;
  }
  if (t >= ntiles) return;
  const unsigned epoch = ep_of(ctl, ep);
  // a large x (ntiles > kSampT) samples in its first kSampT tiles, one key
  // per thread, beside the scan (the positions sample_x takes)
  constexpr int kSampT = kSampleKeys / NT;
  const bool sampled = sel.mat != nullptr && xnnz > kSampleMin;
  const bool my_samp = sampled && samp && t < (unsigned)kSampT;
  int slen = 0;
  if (my_samp) {
    const long long sidx = (long long)t * NT + threadIdx.x;
    const int sc = xkey[sidx * xnnz / kSampleKeys];
    const int sok = (unsigned)sc < (unsigned)ncols ? 1 : 0;
    const int sci = sc & -sok;
    slen = colptr[sci + sok] - colptr[sci];
  }
  const long long base = (long long)t * (NT * IPT) + (long long)threadIdx.x * IPT;
  int len[IPT], beg[IPT];
  long long tsum = 0;
  int mx = 0, mn = INT_MAX;
  // all keys, then all column bounds, are loaded unconditionally (indices
  // clamped) so that each group is one round trip to memory, not one per item
  int c[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const long long i = base + j;
    c[j] = xkey[i < xnnz ? i : (long long)xnnz - 1];
  }
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    // an out-of-range key contributes nothing instead of a wild read
    // (index masked arithmetically: a select here becomes a branch with a
    // scalar load on the other side, and every item then waits for memory)
    const int ok = (base + j < xnnz && (unsigned)c[j] < (unsigned)ncols) ? 1 : 0;
    const int ci = c[j] & -ok;
    beg[j] = colptr[ci];
    len[j] = colptr[ci + ok];
    c[j] = ok;
  }
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    len[j] = c[j] ? len[j] - beg[j] : 0;
    beg[j] = c[j] ? beg[j] : 0;
    if (c[j]) {
      mx = max(mx, len[j]);
      mn = min(mn, len[j]);
    }
    tsum += len[j];
  }
  // the selecting tile stages the selector's inputs now, while the loads
  // above are in, instead of after its look-back (the barrier below
  // publishes them)
  if (selects) stage();
  mx = block_max<NT, int>(mx, s_i);
  mn = block_min<NT, int>(mn, s_i);
  long long ssum = 0;
  int smx = 0;
  if (my_samp) {
    ssum = block_sum<NT, long long>((long long)slen, s_l);
    smx = block_max<NT, int>(slen, s_i);
  }
  // a single tile (small x) needs neither the look-back nor the tile max/min
  const bool single = ntiles == 1;
  if (!single && threadIdx.x == 0) {
    // device-coherent stores, complete before this tile's aggregate is
    // published; the last tile reads them with device-coherent loads once
    // the look-back has seen every earlier tile (no cache write-back fence)
    __hip_atomic_store(&tile_mm[2 * t], mx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&tile_mm[2 * t + 1], mn, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (my_samp) {
      __hip_atomic_store(&samp[2 * t], ssum, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&samp[2 * t + 1], (long long)smx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  }
  long long btot;
  const long long bex = block_excl_sum<NT, long long>(tsum, s_l, &btot);
  const unsigned long long pre =
      single ? 0ull : tile_prefix(state, t, epoch, (unsigned long long)btot, &s_pre);
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
      if (i == xnnz - 1) ptr[xnnz] = (int)run;
    }
  }
  if (t != ntiles - 1) return;
  int gmx = mx, gmn = mn;
  if (!single) {
    // last tile: every earlier tile has published, so its tile_mm is in place
    gmx = 0;
    gmn = INT_MAX;
    for (unsigned u = threadIdx.x; u < ntiles; u += NT) {
      gmx = max(gmx, __hip_atomic_load(&tile_mm[2 * u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
      gmn = min(gmn, __hip_atomic_load(&tile_mm[2 * u + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    }
    gmx = block_max<NT, int>(gmx, s_i);
    gmn = block_min<NT, int>(gmn, s_i);
  }
  if (threadIdx.x == 0) {
    const long long bl = (long long)pre + btot;
    ctl->bin_len = bl;
    ctl->max_len = gmx;
    ctl->min_len = gmn;
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
  // a large x's selector inputs are sampled (the same definition in every
  // launch that selects); the selector word keeps the exact bin_len
  long long fbl = (long long)pre + btot;
  int fmx = gmx;
  if (selects && sampled) {
    if (samp && ntiles > (unsigned)kSampT) {
      // the first tiles' sample partials, in place since the look-back
      // has seen those tiles publish
      long long os = 0;
      int om = 0;
      for (int u = threadIdx.x; u < kSampT; u += NT) {
        os += __hip_atomic_load(&samp[2 * u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        om = max(om, (int)__hip_atomic_load(&samp[2 * u + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
      }
      const long long st = block_sum<NT, long long>(os, s_l);
      fmx = block_max<NT, int>(om, s_i);
      fbl = (long long)llrint((double)st * (double)xnnz / (double)kSampleKeys);
    } else {
      sample_x<NT>(colptr, xkey, ncols, xnnz, s_l, s_i, &fbl, &fmx);
    }
  }
  run_selector(xnnz, fbl, fmx);
  // the selector's choice was written by thread 0 (lane 0 of wave 0)
  if (sel.stash && threadIdx.x == 0)
    *sel.stash = sel_word((long long)pre + btot, sel.mat ? ctl->kernel_choice : -1);
}

// Features-only selector launch: bin_len and the largest and smallest
// column length over x's support (a plain reduction: no scan, no tile
// chain), then the selector. For calls whose slot is a row slot, which
// needs none of the column scan. Every workgroup publishes its partials and
// counts itself done; the one that finishes last (the count it reads is
// the grid size less one) reduces the partials and runs the selector, so no
// workgroup waits for another. Every workgroup loads the selector weights
// at its start, since any of them may be the last.
template <int NT, int KPT>
__global__ __launch_bounds__(NT) void k_select_feat(const int* __restrict__ colptr,
                                                    const int* __restrict__ xkey,
                                                    const float* __restrict__ xval, int ncols, int xnnz,
                                                    long long* part, Ctl* ctl, SelectorArgs sel,
                                                    float* __restrict__ enc,
                                                    const int* __restrict__ enc_perm) {
  __shared__ __align__(16) float s_w[kNWeights];
  __shared__ MatFeat s_mf;
  __shared__ long long s_l[NT / kWave];
  __shared__ int s_i[NT / kWave];
  __shared__ unsigned s_last;
  const bool selects = sel.mat != nullptr;
  constexpr int kWPT = (kNWeights + NT - 1) / NT;
  constexpr int kMF = sizeof(MatFeat) / sizeof(double);
  float wr[kWPT];
  double mfr = 0.0;
  if (selects) {
    if (sel.weights) {
#pragma unroll
      for (int i = 0; i < kWPT; ++i) {
        const int idx = threadIdx.x + i * NT;
        wr[i] = idx < kNWeights ? sel.weights[idx] : 0.f;
      }
    }
    if (threadIdx.x < kMF) mfr = reinterpret_cast<const double*>(sel.mat)[threadIdx.x];
  }
  const int w = threadIdx.x / kWave;
  // for a large x, this thread's share of the sampled keys (the positions
  // sample_x takes), whose loads go out with the scan's
  const bool sampled = selects && xnnz > kSampleMin;
  long long ssum = 0;
  int smx = 0;
  if (sampled) {
    for (long long sidx = (long long)blockIdx.x * NT + threadIdx.x; sidx < kSampleKeys;
         sidx += (long long)gridDim.x * NT) {
      const int c = xkey[sidx * xnnz / kSampleKeys];
      const int ok = (unsigned)c < (unsigned)ncols ? 1 : 0;
      const int ci = c & -ok;
      const int len = colptr[ci + ok] - colptr[ci];
      ssum += len;
      smx = max(smx, len);
    }
  }
  long long sum = 0;
  int mx = 0, mn = INT_MAX;
  // the support encoding of a round is scattered at the start of the next
  // one, and that of the last round after this workgroup has published its
  // partials: the publication then waits for no scatter store, and the
  // selector runs while the stores drain
  int ok[KPT], pc[KPT];
  float xv[KPT];
  bool pending = false;
  auto scatter = [&]() {
    // one wait for every load here; otherwise each store behind a branch
    // below waits for the whole memory counter, the stores before it
    // included
#pragma unroll
    for (int j = 0; j < KPT; ++j) asm volatile("" ::"v"(pc[j]), "v"(xv[j]));
    // support encoding (as launch_scatter with encode): -0.0 marks a zero
    // value inside the support
#pragma unroll
    for (int j = 0; j < KPT; ++j)
      if (ok[j]) enc[pc[j]] = xv[j] == 0.f ? -0.f : xv[j];
  };
  for (long long t0 = (long long)blockIdx.x * NT * KPT; t0 < xnnz; t0 += (long long)gridDim.x * NT * KPT) {
    if (pending) scatter();
    const long long b0 = t0 + (long long)w * kWave * KPT + lane();
    int c[KPT], beg[KPT], end[KPT];
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
      const long long i = b0 + (long long)j * kWave;
      const long long ic = i < xnnz ? i : (long long)xnnz - 1;
      c[j] = xkey[ic];
      xv[j] = enc ? xval[ic] : 0.f;
    }
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
      ok[j] = (b0 + (long long)j * kWave < xnnz && (unsigned)c[j] < (unsigned)ncols) ? 1 : 0;
      const int ci = c[j] & -ok[j];
      beg[j] = colptr[ci];
      end[j] = colptr[ci + ok[j]];
      // the permuted position, loaded unconditionally beside the column
      // bounds (a load behind the store's branch would be waited for alone)
      pc[j] = enc_perm ? enc_perm[ci] : ci;
    }
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
      if (ok[j]) {
        const int len = end[j] - beg[j];
        sum += len;
        mx = max(mx, len);
        mn = min(mn, len);
      }
    }
    pending = enc != nullptr;
  }
  // the selector's inputs into LDS in every workgroup, any of which may be
  // the last (the reduction's barriers publish them)
  if (selects) {
    if (sel.weights) {
#pragma unroll
      for (int i = 0; i < kWPT; ++i) {
        const int idx = threadIdx.x + i * NT;
        if (idx < kNWeights) s_w[wlds_index(idx)] = wr[i];
      }
    }
    if (threadIdx.x < kMF) reinterpret_cast<double*>(&s_mf)[threadIdx.x] = mfr;
  }
  sum = block_sum<NT, long long>(sum, s_l);
  mx = block_max<NT, int>(mx, s_i);
  mn = block_min<NT, int>(mn, s_i);
  if (sampled) {
    ssum = block_sum<NT, long long>(ssum, s_l);
    smx = block_max<NT, int>(smx, s_i);
  }
  if (threadIdx.x == 0) {
    __hip_atomic_store(&part[blockIdx.x], sum, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&part[kSelMaxGrid + blockIdx.x], (long long)mx, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&part[2 * kSelMaxGrid + blockIdx.x], (long long)mn, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    if (sampled) {
      __hip_atomic_store(&part[3 * kSelMaxGrid + blockIdx.x], ssum, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&part[4 * kSelMaxGrid + blockIdx.x], (long long)smx, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // partials in place before the count
    const unsigned before =
        __hip_atomic_fetch_add(&ctl->done[kSlotSelect], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_last = before == gridDim.x - 1 ? 1u : 0u;
    // the last one resets the count for the next launch (stream-ordered)
    if (before == gridDim.x - 1)
      __hip_atomic_store(&ctl->done[kSlotSelect], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  if (!s_last) {
    if (pending) scatter();
    return;
  }
  // the other workgroups' partials (this one's own are in registers)
  long long osum = 0;
  int omx = 0, omn = INT_MAX;
  for (unsigned u = threadIdx.x; u < gridDim.x; u += NT) {
    if (u == blockIdx.x) continue;
    osum += __hip_atomic_load(&part[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    omx = max(omx, (int)__hip_atomic_load(&part[kSelMaxGrid + u], __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT));
    omn = min(omn, (int)__hip_atomic_load(&part[2 * kSelMaxGrid + u], __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT));
  }
  sum += block_sum<NT, long long>(osum, s_l);
  mx = max(mx, block_max<NT, int>(omx, s_i));
  mn = min(mn, block_min<NT, int>(omn, s_i));
  if (threadIdx.x == 0) {
    ctl->bin_len = sum;
    ctl->max_len = mx;
    ctl->min_len = mn;
    ctl->xnnz = xnnz;
  }
  if (!selects) {
    if (pending) scatter();
    return;
  }
  SelectorArgs sl = sel;
  if (sel.weights) sl.weights = s_w;
  sl.mat = &s_mf;
  long long fbl = sum;
  int fmx = mx;
  if (sampled) {
    // the sampled inputs: the other workgroups' partials and this one's
    long long os = 0;
    int om = 0;
    for (unsigned u = threadIdx.x; u < gridDim.x; u += NT) {
      if (u == blockIdx.x) continue;
      os += __hip_atomic_load(&part[3 * kSelMaxGrid + u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      om = max(om, (int)__hip_atomic_load(&part[4 * kSelMaxGrid + u], __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT));
    }
    const long long st = ssum + block_sum<NT, long long>(os, s_l);
    fmx = max(smx, block_max<NT, int>(om, s_i));
    fbl = (long long)llrint((double)st * (double)xnnz / (double)kSampleKeys);
  }
  select_in_wave(ctl, sl, (double)xnnz, (double)fbl, (double)fmx);
  if (sel.stash && threadIdx.x == 0) *sel.stash = sel_word(sum, ctl->kernel_choice);
  if (pending) scatter();
}

// Features-only selector for a large x (xnnz > kSampleMin) when the caller
// takes an estimated bin_len: workgroup 0 (NT = kSampleKeys threads) loads
// one sampled key per thread with the selector's weights, then the keys'
// column bounds, reduces them and runs the selector, with no step that waits
// for another workgroup; workgroups 1.. scatter x's support encoding
// meanwhile. (Spreading 4,096 samples over 16 workgroups needed their
// partials to meet through memory, two more dependent round trips; 4,096
// samples in one workgroup were 12,000 cache-line fills through one L1.)
// The selector word carries the estimate, flagged.
template <int NT, int KPT>
__global__ __launch_bounds__(NT) void k_select_sampled(const int* __restrict__ colptr,
                                                       const int* __restrict__ xkey,
                                                       const float* __restrict__ xval, int ncols, int xnnz,
                                                       Ctl* ctl, SelectorArgs sel, float* __restrict__ enc,
                                                       const int* __restrict__ enc_perm) {
  static_assert(NT == kSampleKeys, "one sampled key per thread of workgroup 0");
  __shared__ __align__(16) float s_w[kNWeights];
  __shared__ MatFeat s_mf;
  __shared__ long long s_l[NT / kWave];
  __shared__ int s_i[NT / kWave];
  if (blockIdx.x == 0) {
    constexpr int kWPT = (kNWeights + NT - 1) / NT;
    constexpr int kMF = sizeof(MatFeat) / sizeof(double);
    static_assert(kMF <= NT, "MatFeat is copied as doubles");
    float wr[kWPT];
    if (sel.weights) {
#pragma unroll
      for (int i = 0; i < kWPT; ++i) {
        const int idx = threadIdx.x + i * NT;
        wr[i] = idx < kNWeights ? sel.weights[idx] : 0.f;
      }
    }
    const double mfr = threadIdx.x < kMF ? reinterpret_cast<const double*>(sel.mat)[threadIdx.x] : 0.0;
    // this thread's sample (the positions sample_x takes)
    const int sc = xkey[(long long)threadIdx.x * xnnz / kSampleKeys];
    const int sok = (unsigned)sc < (unsigned)ncols ? 1 : 0;
    const int sci = sc & -sok;
    const int len = colptr[sci + sok] - colptr[sci];
    if (sel.weights) {
#pragma unroll
      for (int i = 0; i < kWPT; ++i) {
        const int idx = threadIdx.x + i * NT;
        if (idx < kNWeights) s_w[wlds_index(idx)] = wr[i];
      }
    }
    if (threadIdx.x < kMF) reinterpret_cast<double*>(&s_mf)[threadIdx.x] = mfr;
    const long long sum = block_sum<NT, long long>((long long)len, s_l);  // its barriers publish s_w, s_mf
    const int mx = block_max<NT, int>(len, s_i);
    const long long bl = (long long)llrint((double)sum * (double)xnnz / (double)kSampleKeys);
    if (threadIdx.x == 0) {
      ctl->bin_len = bl;
      ctl->max_len = mx;
      ctl->min_len = 0;
      ctl->xnnz = xnnz;
    }
    SelectorArgs sl = sel;
    if (sel.weights) sl.weights = s_w;
    sl.mat = &s_mf;
    select_in_wave(ctl, sl, (double)xnnz, (double)bl, (double)mx);
    if (sel.stash && threadIdx.x == 0) *sel.stash = sel_word(bl, ctl->kernel_choice | kSelEstimated);
    return;
  }
  // workgroups 1..: the support encoding
  if (!enc) return;
  const int w = threadIdx.x / kWave;
  const long long sb = blockIdx.x - 1, sg = gridDim.x - 1;
  for (long long t0 = sb * NT * KPT; t0 < xnnz; t0 += sg * NT * KPT) {
    const long long b0 = t0 + (long long)w * kWave * KPT + lane();
    int c[KPT], pc[KPT];
    float xv[KPT];
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
      const long long i = b0 + (long long)j * kWave;
      const long long ic = i < xnnz ? i : (long long)xnnz - 1;
      c[j] = xkey[ic];
      xv[j] = xval[ic];
    }
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
      const int ok = (b0 + (long long)j * kWave < xnnz && (unsigned)c[j] < (unsigned)ncols) ? 1 : 0;
      const int ci = c[j] & -ok;
      pc[j] = enc_perm ? enc_perm[ci] : ci;
      c[j] = ok;
    }
#pragma unroll
    for (int j = 0; j < KPT; ++j) asm volatile("" ::"v"(pc[j]), "v"(xv[j]));
#pragma unroll
    for (int j = 0; j < KPT; ++j)
      if (c[j]) enc[pc[j]] = xv[j] == 0.f ? -0.f : xv[j];
  }
}

void launch_select_feat(hipStream_t s, const int* colptr, const int* xkey, const float* xval,
                        int ncols, int xnnz, long long* part, Ctl* ctl, const SelectorArgs& sel,
                        float* enc, const int* enc_perm) {
  if (sel.estimate && sel.mat && xnnz > kSampleMin) {
    constexpr int kPer = kSampleKeys * 4;
    // the sampling workgroup, then with the encoding one round of 4 entries
    // per thread for the others, 512 workgroups at most
    const int grid = enc ? 1 + (int)std::max<long long>(1, std::min<long long>(511, ((long long)xnnz + kPer - 1) / kPer))
                         : 1;
    hipLaunchKernelGGL((k_select_sampled<kSampleKeys, 4>), dim3(grid), dim3(kSampleKeys), 0, s, colptr, xkey, xval,
                       ncols, xnnz, ctl, sel, enc, enc_perm);
    return;
  }
  // 16 keys per thread fill at most 512 workgroups up to 2 M keys; below
  // 256 K keys, 8 per thread give twice the workgroups and half the chain
  // per thread (3 µs less at 100 K keys, 4 µs more at 1 M)
  if (xnnz < (1 << 18)) {
    constexpr int kPer = 256 * 8;
    const int grid = std::max(1, (xnnz + kPer - 1) / kPer);
    hipLaunchKernelGGL((k_select_feat<256, 8>), dim3(grid), dim3(256), 0, s, colptr, xkey, xval, ncols,
                       xnnz, part, ctl, sel, enc, enc_perm);
    return;
  }
  constexpr int kPer = 256 * 16;
  const int grid = (int)std::max<long long>(
      1, std::min<long long>(kSelMaxGrid, ((long long)xnnz + kPer - 1) / kPer));
  hipLaunchKernelGGL((k_select_feat<256, 16>), dim3(grid), dim3(256), 0, s, colptr, xkey, xval, ncols,
                     xnnz, part, ctl, sel, enc, enc_perm);
}

void launch_prep(hipStream_t s, const int* colptr, const int* xkey, int ncols, int xnnz,
                 const int* xnnz_dev, int* ptr, int* xbeg, int* split, unsigned long long* state,
                 int* tile_mm, Ctl* ctl, const SelectorArgs* sel, Ep ep, long long* samp) {
  // exact grid when the host knows xnnz, else the bound from ncols
  const long long cnt = xnnz_dev ? ncols : xnnz;
  const int grid = (int)std::max<long long>(1, (cnt + kPrepTile - 1) / kPrepTile);
  SelectorArgs a{};
  if (sel) a = *sel;
  hipLaunchKernelGGL((k_prep<kPrepNT, kPrepIPT>), dim3(grid), dim3(kPrepNT), 0, s, colptr, xkey,
                     ncols, xnnz, xnnz_dev, ptr, xbeg, split, state, tile_mm, ctl, ep, a, samp);
}

// ------------------------------------------------------------- naive col
// mode 0: atomics into acc + hit bitmap; mode 1: bins at ptr[e] + offset;
// mode 2: row keys only at ptr[e] + offset (boolean: no value is read).
template <int NT, int MODE>
__global__ __launch_bounds__(NT) void k_naive_col(const int* __restrict__ colptr,
                                                  const int* __restrict__ crow,
                                                  const float* __restrict__ cval,
                                                  const int* __restrict__ xkey,
                                                  const float* __restrict__ xval,
                                                  const int* __restrict__ ptr, int ncols, int xnnz_arg,
                                                  const int* __restrict__ xnnz_dev, int lg_team,
                                                  float* acc, unsigned char* marks,
                                                  int* __restrict__ bkey, float* __restrict__ bval,
                                                  Gate g) {
  if (gate_shut(g)) return;
  const int xnnz = xnnz_dev ? *xnnz_dev : xnnz_arg;
  const int team = 1 << lg_team;
  const int tpb = NT >> lg_team;
  const int tid_t = threadIdx.x & (team - 1);
  const int tm = threadIdx.x >> lg_team;
  for (long long e = (long long)blockIdx.x * tpb + tm; e < xnnz; e += (long long)gridDim.x * tpb) {
    const int c = xkey[e];
    if ((unsigned)c >= (unsigned)ncols) continue;
    const float xv = MODE == 2 ? 0.f : xval[e];
    const int b = colptr[c];
    const int en = colptr[c + 1];
    if (MODE == 2) {
      const int o = ptr[e] - b;
      for (int j = b + tid_t; j < en; j += team) bkey[o + j] = crow[j];
    } else if (MODE == 0) {
      for (int j = b + tid_t; j < en; j += team) {
        const int r = crow[j];
        atomic_add_f32(acc + r, cval[j] * xv);
        marks[r] = 1;  // plain byte store: idempotent, no atomic needed
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
                      const int* xnnz_dev, int team, float* acc, unsigned char* marks, int* bkey,
                      float* bval, Gate g) {
  int lg = 0;
  while ((1 << lg) < team && lg < 6) ++lg;
  const int tpb = 256 >> lg;
  long long need = xnnz_dev ? (long long)kPersistGrid * tpb : xnnz;
  int grid = (int)std::min<long long>(4096, (need + tpb - 1) / tpb);
  if (grid < 1) grid = 1;
  if (mode == 0)
    hipLaunchKernelGGL((k_naive_col<256, 0>), dim3(grid), dim3(256), 0, s, colptr, crow, cval, xkey,
                       xval, ptr, ncols, xnnz, xnnz_dev, lg, acc, marks, bkey, bval, g);
  else if (mode == 2)
    hipLaunchKernelGGL((k_naive_col<256, 2>), dim3(grid), dim3(256), 0, s, colptr, crow, cval, xkey,
                       xval, ptr, ncols, xnnz, xnnz_dev, lg, acc, marks, bkey, bval, g);
  else
    hipLaunchKernelGGL((k_naive_col<256, 1>), dim3(grid), dim3(256), 0, s, colptr, crow, cval, xkey,
                       xval, ptr, ncols, xnnz, xnnz_dev, lg, acc, marks, bkey, bval, g);
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
                                                  unsigned char* marks, int* __restrict__ bkey,
                                                  float* __restrict__ bval, Gate g) {
  if (gate_shut(g)) return;
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
    const int nseg = a1 - sa + 1;  // <= T + 2
    {
      // every load of the segment table issued at once (clamped indices)
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
        xv[u] = MODE == 2 ? 0.f : xval[ic];
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
    const int np = (int)(b1 - b0);  // <= T
    if (np > 0) {
      // the VT products of a thread: indices from LDS, then all their
      // matrix loads issued together
      int idx[VT];
      float xs[VT];
#pragma unroll
      for (int k = 0; k < VT; ++k) {
        const int q0 = threadIdx.x + k * NT;
        const int q = q0 < np ? q0 : 0;
        const int sg = s_seg[q];
        idx[k] = s_beg[sg] + (int)(b0 + q - s_ptr[sg]);
        xs[k] = s_xv[sg];
      }
      int r[VT];
      float cv[VT];
#pragma unroll
      for (int k = 0; k < VT; ++k) {
        r[k] = crow[idx[k]];
        cv[k] = MODE == 2 ? 0.f : cval[idx[k]];
      }
#pragma unroll
      for (int k = 0; k < VT; ++k) {
        const int q = threadIdx.x + k * NT;
        if (q < np) {
          const float v = cv[k] * xs[k];
          if (MODE == 2) {
            bkey[b0 + q] = r[k];
          } else if (MODE == 0) {
            atomic_add_f32(acc + r[k], v);
            marks[r[k]] = 1;
          } else {
            bkey[b0 + q] = r[k];
            bval[b0 + q] = v;
          }
        }
      }
    }
    __syncthreads();
  }
}

void launch_lb_gather(hipStream_t s, int mode, const int* ptr, const int* xbeg, const float* xval,
                      const int* split, const int* crow, const float* cval, const Ctl* ctl,
                      int xnnz, const int* xnnz_dev, float* acc, unsigned char* marks, int* bkey,
                      float* bval, long long tiles_hint, Gate g) {
  long long grid = tiles_hint > 0 ? std::min<long long>(tiles_hint, 8192) : 2048;
  if (grid < 1) grid = 1;
  if (mode == 0)
    hipLaunchKernelGGL((k_lb_gather<kLbNT, kLbVT, 0>), dim3((unsigned)grid), dim3(kLbNT), 0, s, ptr,
                       xbeg, xval, split, crow, cval, ctl, xnnz, xnnz_dev, acc, marks, bkey, bval, g);
  else if (mode == 2)
    hipLaunchKernelGGL((k_lb_gather<kLbNT, kLbVT, 2>), dim3((unsigned)grid), dim3(kLbNT), 0, s, ptr,
                       xbeg, xval, split, crow, cval, ctl, xnnz, xnnz_dev, acc, marks, bkey, bval, g);
  else
    hipLaunchKernelGGL((k_lb_gather<kLbNT, kLbVT, 1>), dim3((unsigned)grid), dim3(kLbNT), 0, s, ptr,
                       xbeg, xval, split, crow, cval, ctl, xnnz, xnnz_dev, acc, marks, bkey, bval, g);
}

// ----------------------------------------------------- two-layer gather
// The reference's two-layer load balance (CscBasedMyNoSortSpmspv /
// CscBasedSortMySpmspv, spmspv.h:1823-1878, 2000-2029;
// csc_two_layer_load_balance_fuse_reduction_modify.cuh:31-394): tiles of
// exactly T products (the merge path counts x entries as items too).
// Layer 1, k_tl_starts: per tile, the x entry holding its first product
// (the last i with ptr[i] <= t*T). Layer 2, k_tl_gather: the tile's x
// entries are staged in LDS when they fit (empty columns can make a tile
// span many; then every product searches ptr in memory) and each product
// finds its entry there by binary search. Products go to the accumulator
// (atomics + row marks) or to their bin at the global product index (the
// reference's write-back dropped the block offset, :146-149).
__global__ void k_tl_starts(const int* __restrict__ ptr, const Ctl* __restrict__ ctl, int xnnz, int tile,
                            int* __restrict__ starts) {
  const long long bl = ctl->bin_len;
  const long long ntiles = (bl + tile - 1) / tile;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t <= ntiles;
       t += (long long)gridDim.x * blockDim.x) {
    if (t == ntiles) {
      starts[t] = xnnz;
      continue;
    }
    const long long p = t * tile;
    int lo = 0, hi = xnnz;  // ptr[lo] <= p < ptr[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (ptr[mid] <= p) lo = mid;
      else hi = mid;
    }
    starts[t] = lo;
  }
}

template <int NT, int VT, int MODE>
__global__ __launch_bounds__(NT) void k_tl_gather(const int* __restrict__ ptr, const int* __restrict__ xbeg,
                                                  const float* __restrict__ xval,
                                                  const int* __restrict__ starts,
                                                  const int* __restrict__ crow,
                                                  const float* __restrict__ cval,
                                                  const Ctl* __restrict__ ctl, float* acc,
                                                  unsigned char* marks, int* __restrict__ bkey,
                                                  float* __restrict__ bval, Gate g) {
  if (gate_shut(g)) return;
  constexpr int T = NT * VT;
  __shared__ int s_ptr[T + 1];
  __shared__ int s_beg[T + 1];
  __shared__ float s_xv[T + 1];
  const long long bl = ctl->bin_len;
  const long long ntiles = (bl + T - 1) / T;
  for (long long t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const long long p0 = t * T;
    const long long p1 = min(p0 + (long long)T, bl);
    const int e0 = starts[t];
    // the last product's entry is at most the next tile's first one
    const int e1 = t + 1 < ntiles ? starts[t + 1] : starts[ntiles] - 1;
    const int span = e1 - e0 + 1;
    const bool fits = span <= T + 1;
    if (fits) {
      for (int k = threadIdx.x; k < span; k += NT) {
        s_ptr[k] = ptr[e0 + k];
        s_beg[k] = xbeg[e0 + k];
        s_xv[k] = xval[e0 + k];
      }
    }
    __syncthreads();
    int idx[VT], ent[VT];
#pragma unroll
    for (int j = 0; j < VT; ++j) {
      const long long p = p0 + (long long)j * NT + threadIdx.x;
      int lo = 0, hi = span;  // (ptr of entry e0 + lo) <= p < (ptr of e0 + hi)
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        const int pm = fits ? s_ptr[mid] : ptr[e0 + mid];
        if (pm <= p) lo = mid;
        else hi = mid;
      }
      ent[j] = lo;
      const int base = fits ? s_ptr[lo] : ptr[e0 + lo];
      const int beg = fits ? s_beg[lo] : xbeg[e0 + lo];
      idx[j] = p < p1 ? beg + (int)(p - base) : -1;
    }
#pragma unroll
    for (int j = 0; j < VT; ++j) {
      if (idx[j] < 0) continue;
      const long long p = p0 + (long long)j * NT + threadIdx.x;
      const float xv = fits ? s_xv[ent[j]] : xval[e0 + ent[j]];
      const int r = crow[idx[j]];
      const float v = cval[idx[j]] * xv;
      if (MODE == 0) {
        atomic_add_f32(acc + r, v);
        marks[r] = 1;
      } else {
        bkey[p] = r;
        bval[p] = v;
      }
    }
    __syncthreads();
  }
}

void launch_tl_gather(hipStream_t s, int mode, const int* ptr, const int* xbeg, const float* xval, int xnnz,
                      int* starts, const int* crow, const float* cval, const Ctl* ctl, float* acc,
                      unsigned char* marks, int* bkey, float* bval, long long tiles_hint, Gate g) {
  constexpr int NT = 256, VT = 7, T = NT * VT;
  const long long tiles = tiles_hint >= 0 ? tiles_hint : 8192;
  const int sgrid = (int)std::max<long long>(1, std::min<long long>(1024, (tiles + 1 + 255) / 256));
  hipLaunchKernelGGL(k_tl_starts, dim3(sgrid), dim3(256), 0, s, ptr, ctl, xnnz, T, starts);
  const int grid = (int)std::max<long long>(1, std::min<long long>(4096, tiles));
  if (mode == 0)
    hipLaunchKernelGGL((k_tl_gather<NT, VT, 0>), dim3(grid), dim3(NT), 0, s, ptr, xbeg, xval, starts, crow, cval,
                       ctl, acc, marks, bkey, bval, g);
  else
    hipLaunchKernelGGL((k_tl_gather<NT, VT, 1>), dim3(grid), dim3(NT), 0, s, ptr, xbeg, xval, starts, crow, cval,
                       ctl, acc, marks, bkey, bval, g);
}

// -------------------------------------------------------------- compact
// 8,192 rows per tile (kCmpNT bit-array words), 8 rows per thread, ordered
// tiles. The hits come either from a bit array (row kernels) or from one
// mark byte per row (column atomic kernels); both are cleared on the way.
// Eight rows per thread (not one word of 32) puts 1,024 threads on a tile,
// so that at high hit rates each thread writes at most 8 outputs, whose
// loads go out in one round; 8,192-row tiles keep the look-back short (a
// tile reads 64 predecessors per round: 2,048-row tiles took up to 8 rounds
// at a million rows).
__device__ __forceinline__ unsigned bytes_to_bits8(uint2 a) {
  // byte k (0 or 1) of a 32-bit lane -> bit k: the four partial products of
  // x * 0x01020408 land on distinct bits, bits 24..27 hold bytes 0..3
  auto nib = [](unsigned x) { return (((x & 0x01010101u) * 0x01020408u) >> 24) & 0xFu; };
  return nib(a.x) | (nib(a.y) << 4);
}

template <int NT, bool BYTES>
__global__ __launch_bounds__(NT) void k_compact(unsigned* __restrict__ bitmap,
                                                unsigned char* __restrict__ marks, int m,
                                                float* __restrict__ acc, int reset_acc,
                                                int* __restrict__ ykey, float* __restrict__ yval,
                                                unsigned long long* state, Ctl* ctl, Ep ep,
                                                const int* __restrict__ keymap, Gate g,
                                                long long* sig, unsigned* __restrict__ clr, int nclr) {
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < nclr; i += NT) clr[(long long)i * kBinCurStride] = 0u;
  if (gate_shut(g)) return;
  constexpr int kRows = 8, kTile = NT * kRows;
  static_assert(kTile == kCmpNT * 32, "a tile is kCmpNT bit-array words");
  __shared__ unsigned long long s_pre;
  __shared__ int s_i[NT / kWave];
  __shared__ int s_rows[kTile];
  __shared__ float s_vals[kTile];
  const unsigned ntiles = (unsigned)(((long long)m + kTile - 1) / kTile);
  const unsigned t = blockIdx.x;
  if (t >= ntiles) {
    // no rows: an empty result
    if (ntiles == 0 && t == 0 && threadIdx.x == 0) {
      ctl->ynnz = 0;
      if (sig) *sig = 0;
    }
    return;
  }
  const unsigned epoch = ep_of(ctl, ep);
  // first row of this thread; the mark bytes are padded to 32 rows and the
  // bit array to whole words, so the 8 rows are readable whenever r0 < m
  const int r0 = (int)t * kTile + (int)threadIdx.x * kRows;
  unsigned b = 0u;
  if (r0 < m) {
    if (BYTES) b = bytes_to_bits8(*reinterpret_cast<const uint2*>(marks + r0));
    else b = (bitmap[r0 >> 5] >> (r0 & 31)) & 0xFFu;
  }
  // the thread's 8 sums, loaded while the scan and the look-back run (the
  // 8 rows are consecutive: two 16-byte loads when they all exist)
  float av[kRows];
#pragma unroll
  for (int u = 0; u < kRows; ++u) av[u] = 0.f;
  const bool whole = r0 + kRows <= m && (reinterpret_cast<uintptr_t>(acc + r0) & 15u) == 0;
  if (b && yval) {
    if (whole) {
      const float4 lo = *reinterpret_cast<const float4*>(acc + r0);
      const float4 hi = *reinterpret_cast<const float4*>(acc + r0 + 4);
      av[0] = lo.x; av[1] = lo.y; av[2] = lo.z; av[3] = lo.w;
      av[4] = hi.x; av[5] = hi.y; av[6] = hi.z; av[7] = hi.w;
    } else {
#pragma unroll
      for (int u = 0; u < kRows; ++u)
        if ((b >> u) & 1u) av[u] = acc[r0 + u];
    }
  }
  const int cnt = __popc(b);
  int btot;
  const int bex = block_excl_sum<NT, int>(cnt, s_i, &btot);
  const unsigned long long pre = tile_prefix(state, t, epoch, (unsigned long long)btot, &s_pre);
  // rows of the tile and their sums land in LDS at their local rank, then
  // leave as contiguous, coalesced writes
  int loc = bex;
  if (b) {
    if (BYTES) {
      *reinterpret_cast<uint2*>(marks + r0) = make_uint2(0u, 0u);
    } else {
      // the four threads of a word are lanes of one wave, which all read
      // the word before any of them clears it
      bitmap[r0 >> 5] = 0u;
    }
#pragma unroll
    for (int u = 0; u < kRows; ++u) {
      if ((b >> u) & 1u) {
        s_rows[loc] = r0 + u;
        s_vals[loc] = av[u];
        ++loc;
        if (reset_acc && !whole) acc[r0 + u] = 0.0f;
      }
    }
    if (reset_acc && whole) {
      *reinterpret_cast<float4*>(acc + r0) = make_float4(0.f, 0.f, 0.f, 0.f);
      *reinterpret_cast<float4*>(acc + r0 + 4) = make_float4(0.f, 0.f, 0.f, 0.f);
    }
  }
  __syncthreads();
  // U rows per thread per round, their loads issued together
  // (the loads are unconditional, on a valid row, so that the compiler
  // issues them back to back instead of waiting for each in a branch)
  constexpr int U = 8;
  for (int i0 = threadIdx.x; i0 < btot; i0 += NT * U) {
    int r[U], key[U];
    float v[U];
    unsigned ok = 0u;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = i0 + u * NT;
      const bool in = i < btot;
      r[u] = s_rows[in ? i : 0];
      ok |= in ? (1u << u) : 0u;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = i0 + u * NT;
      key[u] = keymap ? keymap[(unsigned)r[u]] : r[u];
      v[u] = s_vals[i < btot ? i : 0];
    }
    // one wait for all loads here; otherwise every store behind a branch
    // below waits for the whole memory counter, stores included
#pragma unroll
    for (int u = 0; u < U; ++u) asm volatile("" ::"v"(key[u]), "v"(v[u]));
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if ((ok >> u) & 1u) {
        const long long o = (long long)pre + i0 + u * NT;
        if (sig) {
          // device-coherent stores: visible once acknowledged, with no
          // write-back of this workgroup's L2 before the count goes out
          if (ykey) __hip_atomic_store(ykey + o, key[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (yval) __hip_atomic_store(yval + o, v[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
          if (ykey) ykey[o] = key[u];
          if (yval) yval[o] = v[u];
        }
      }
    }
  }
  if (t == ntiles - 1 && threadIdx.x == 0)
    __hip_atomic_store(&ctl->ynnz, (int)(pre + btot), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (sig) {
    // the result count goes to the host word from the workgroup that
    // finishes last, once every workgroup's output stores (device-coherent
    // above) are acknowledged: the count never arrives ahead of the result.
    // The cleared state (marks, accumulator) is read by later launches only,
    // after this one has ended.
    // One 64-bit atomic per workgroup: an arrival in the high half and,
    // from the last tile, the count in the low half (below 2^31: no carry),
    // so the workgroup that completes the arrivals also holds the count.
    static_assert(offsetof(Ctl, done) % 8 == 0 && kSlotCount % 2 == 0, "64-bit count word");
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (threadIdx.x == 0) {
      auto* word = reinterpret_cast<unsigned long long*>(&ctl->done[kSlotCount]);
      const unsigned long long mine =
          (1ull << 32) | (t == ntiles - 1 ? (unsigned long long)(pre + btot) : 0ull);
      const unsigned long long now =
          __hip_atomic_fetch_add(word, mine, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + mine;
      if ((unsigned)(now >> 32) == ntiles) {
        __hip_atomic_store(word, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // every workgroup's device-coherent outputs were acknowledged before
        // its arrival: the count needs no L2 write-back (a system fence would
        // write back the whole L2 of this XCD)
        *sig = (long long)(unsigned)now;
      }
    }
  }
}

void launch_compact(hipStream_t s, unsigned* bitmap, unsigned char* marks, int m, float* acc,
                    int reset_acc, int* ykey, float* yval, unsigned long long* state, Ctl* ctl,
                    Ep ep, const int* keymap, Gate g, long long* sig, unsigned* clr, int nclr) {
  constexpr int nt = kCmpNT * 4;
  const int grid = std::max(1, (int)(((long long)m + kCmpNT * 32 - 1) / (kCmpNT * 32)));
  if (marks)
    hipLaunchKernelGGL((k_compact<nt, true>), dim3(grid), dim3(nt), 0, s, bitmap, marks, m, acc,
                       reset_acc, ykey, yval, state, ctl, ep, keymap, g, sig, clr, nclr);
  else
    hipLaunchKernelGGL((k_compact<nt, false>), dim3(grid), dim3(nt), 0, s, bitmap, marks, m, acc,
                       reset_acc, ykey, yval, state, ctl, ep, keymap, g, sig, clr, nclr);
}

// ------------------------------------------------------------ utilities
// The host reads only the word itself (device results stay in device
// memory, read by later launches), so no L2 write-back precedes it.
__global__ void k_signal_int(const int* src, long long* dst, Gate g) {
  if (gate_shut(g)) return;
  const long long v = *src;
  *dst = v;
}
void launch_signal_int(hipStream_t s, const int* src, long long* dst, Gate g) {
  hipLaunchKernelGGL(k_signal_int, dim3(1), dim3(1), 0, s, src, dst, g);
}

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

__global__ void k_fill_f32(float* p, long long n, float v, Gate g) {
  if (gate_shut(g)) return;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_i32(int* p, long long n, int v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}
void launch_fill_f32(hipStream_t s, float* p, long long n, float v, Gate g) {
  if (n <= 0) return;
  const long long grid = std::min<long long>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)grid), dim3(256), 0, s, p, n, v, g);
}
void launch_fill_i32(hipStream_t s, int* p, long long n, int v) {
  if (n <= 0) return;
  const long long grid = std::min<long long>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)grid), dim3(256), 0, s, p, n, v);
}

__global__ void k_scatter(const int* __restrict__ key, const float* __restrict__ val, int nnz_arg,
                          const int* nnz_dev, int len, float* __restrict__ dense, int encode, Gate g,
                          const int* __restrict__ perm) {
  if (gate_shut(g)) return;
  const int nnz = nnz_dev ? *nnz_dev : nnz_arg;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += gridDim.x * blockDim.x) {
    const int k = key[i];
    const float v = val[i];
    if ((unsigned)k < (unsigned)len) dense[perm ? perm[k] : k] = (encode && v == 0.f) ? -0.f : v;
  }
}
void launch_scatter(hipStream_t s, const int* key, const float* val, int nnz, const int* nnz_dev,
                    int len, float* dense, bool encode, Gate g, const int* perm) {
  int grid = nnz_dev ? 2048 : std::max(1, std::min(8192, (nnz + 255) / 256));
  if (!nnz_dev && nnz <= 0) return;
  hipLaunchKernelGGL(k_scatter, dim3(grid), dim3(256), 0, s, key, val, nnz, nnz_dev, len, dense,
                     encode ? 1 : 0, g, perm);
}

__global__ void k_encode_x(const float* __restrict__ x, const unsigned* __restrict__ bits, int n,
                           float* __restrict__ enc, const int* __restrict__ perm) {
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (long long)gridDim.x * blockDim.x) {
    const bool in = (bits[c >> 5] >> (c & 31)) & 1u;
    const float v = x[c];
    enc[perm ? perm[c] : c] = in ? (v == 0.f ? -0.f : v) : 0.f;
  }
}
void launch_encode_x(hipStream_t s, const float* x, const unsigned* bits, int n, float* enc,
                     const int* perm) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_encode_x, dim3(std::min(8192, (n + 255) / 256)), dim3(256), 0, s, x, bits, n,
                     enc, perm);
}

__global__ void k_set_bits(const int* __restrict__ key, int nnz_arg, const int* nnz_dev, int len,
                           unsigned* bits) {
  const int nnz = nnz_dev ? *nnz_dev : nnz_arg;
  // loop bound is wave-uniform so the shuffles below see every lane
  for (int i0 = blockIdx.x * blockDim.x; i0 < nnz; i0 += gridDim.x * blockDim.x) {
    const int i = i0 + threadIdx.x;
    int k = -1;
    if (i < nnz) k = key[i];
    const bool ok = (unsigned)k < (unsigned)len;
    const int w = ok ? (k >> 5) : -1;
    unsigned bit = ok ? (1u << (k & 31)) : 0u;
    // segmented OR over runs of lanes with the same word (sorted keys make
    // long runs); the last lane of each run issues the one atomic
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const int wo = __shfl_up(w, o, kWave);
      const unsigned bo = __shfl_up(bit, o, kWave);
      if (lane() >= o && wo == w) bit |= bo;
    }
    const int wn = __shfl_down(w, 1, kWave);
    const bool last = lane() == kWave - 1 || wn != w;
    if (ok && last) atomicOr(bits + w, bit);
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
