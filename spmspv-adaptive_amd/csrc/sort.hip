// In-house LSD radix sort of (uint32 key, 32-bit value) pairs and the
// segmented reduce-by-key that follows it: the sort half of the reference's
// CSC sort kernels (thrust::sort_by_key + thrust::reduce_by_key,
// la/include/spmspv/csc-spmspv/detail/device/csc_sort_spmspv.cuh:339-437,
// 1030-1152), written for wave64:
//  * keys are ranked inside a tile with 8 wave ballots per key (peers with
//    the same 8-bit digit) and per-wave digit counters in LDS (stable);
//  * each pass counts digits per tile, scans the digit-major count array
//    (one small look-back scan over 256 x tiles entries) and scatters, so no
//    tile waits on a chain of earlier tiles;
//  * tiles are re-ordered in LDS before the store so that runs of equal
//    digits leave as contiguous, coalesced writes;
//  * n <= one tile takes a single-workgroup path that sorts all passes in
//    LDS and reduces in the same launch.
// The reduce is deterministic: fixed-order segmented scans inside a tile
// and a fixed-order carry fix-up across tiles.
#include "kernels.hpp"
#include "selector.hpp"

namespace spmspv {

// ---------------------------------------------------------------- ranking
// Items of a tile are laid out per wave: wave w owns positions
// [w*64*IPT, (w+1)*64*IPT); item j of lane l sits at w*64*IPT + j*64 + l
// (IPT may be a run-time nj <= IPT: slots j >= nj hold no item). On return
// cnt (this wave's 256 counters) holds the per-digit counts of the wave's
// chunk and rank[j] the stable rank of item j among equal digits inside
// the chunk.
template <int IPT>
__device__ __forceinline__ void wave_rank(const unsigned (&key)[IPT], const bool (&valid)[IPT],
                                          int shift, unsigned* cnt, int (&rank)[IPT],
                                          unsigned (&dig)[IPT], int nj = IPT) {
  for (int k = lane(); k < 256; k += kWave) cnt[k] = 0u;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  const uint64_t lt = lane() == 0 ? 0ull : (~0ull >> (kWave - lane()));
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    if (j >= nj) {  // slots past the tile's fill (wave-uniform): no items
      rank[j] = 0;
      dig[j] = 0u;
      continue;
    }
    const unsigned d = valid[j] ? ((key[j] >> shift) & 255u) : 0u;
    uint64_t peers = __ballot(valid[j]);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    unsigned base = 0;
    if (valid[j]) base = cnt[d];
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (valid[j]) {
      rank[j] = (int)base + __popcll(peers & lt);
      if ((peers >> lane()) == 1ull) cnt[d] = base + (unsigned)__popcll(peers);
    } else {
      rank[j] = 0;
    }
    dig[j] = d;
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
}

// Converts the per-wave counts into exclusive-over-waves offsets (in place)
// and returns, for digit d = threadIdx.x (NT == 256), the tile total.
template <int NW>
__device__ __forceinline__ unsigned digit_wave_offsets(unsigned* s_cnt) {
  const int d = threadIdx.x;
  unsigned tot = 0;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    const unsigned c = s_cnt[w * 256 + d];
    s_cnt[w * 256 + d] = tot;
    tot += c;
  }
  return tot;
}

// -------------------------------------------------------- segmented scan
// Exclusive segmented sum over the threads of a block: the combine of
// (f1,v1) then (f2,v2) is (f1|f2, f2 ? v2 : v1 + v2). Returns for each
// thread the combined value of all earlier threads back to the nearest one
// with f set (inclusive). Also returns the block-inclusive value of the last
// thread in *last_incl (valid in all threads).
template <int NT, typename V>
__device__ __forceinline__ V block_seg_excl(bool f, V v, int* s_f, V* s_v, V* last_incl) {
  constexpr int NW = NT / kWave;
  const int w = threadIdx.x / kWave;
  int fi = f ? 1 : 0;
  V vi = v;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const int fo = __shfl_up(fi, o, kWave);
    const V vo = __shfl_up(vi, o, kWave);
    if (lane() >= o) {
      if (!fi) vi = vo + vi;
      fi |= fo;
    }
  }
  if (lane() == kWave - 1) {
    s_f[w] = fi;
    s_v[w] = vi;
  }
  __syncthreads();
  V pv = V(0);
  for (int k = 0; k < w; ++k) pv = s_f[k] ? s_v[k] : pv + s_v[k];
  V tv = V(0);
  for (int k = 0; k < NW; ++k) tv = s_f[k] ? s_v[k] : tv + s_v[k];
  const V incl = fi ? vi : pv + vi;
  V ex = __shfl_up(incl, 1, kWave);
  if (lane() == 0) ex = pv;
  __syncthreads();
  *last_incl = tv;
  return ex;
}

// ---------------------------------------------------------- tile reduce
// Reduces the sorted keys s_k[0..cnt) / values s_v of one tile. `prev_key`
// is the key just before the tile (ignored when global_start). get_prefix
// (called by all threads with the tile's head count) returns the number of
// heads of all earlier tiles. Segment ends whose head lies in an earlier
// tile go to carry_val/carry_idx[tile] for the fix-up kernel.
// With COHERENT the outputs go out as device-coherent stores (visible to
// every XCD and to copies once acknowledged, no L2 write-back needed).
template <int NT, int IPT, typename PrefixFn, bool COHERENT = false>
__device__ __forceinline__ long long tile_reduce(const unsigned* s_k, const float* s_v, int cnt,
                                                 unsigned prev_key, bool global_start,
                                                 PrefixFn get_prefix, int* ykey, float* yval,
                                                 float* carry_val, int* carry_idx, long long tile,
                                                 int* s_i, int* s_f, float* s_fv, int ipt = IPT) {
  // thread t reduces sorted items t*ipt .. t*ipt+ipt-1 (ipt <= IPT)
  const int i0 = threadIdx.x * ipt;
  bool head[IPT];
  int nh = 0;
  float tail = 0.f;
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const int i = i0 + j;
    head[j] = false;
    if (j < ipt && i < cnt) {
      const unsigned pk = (i == 0) ? prev_key : s_k[i - 1];
      head[j] = (i == 0 && global_start) || s_k[i] != pk;
      if (head[j]) tail = s_v[i];
      else tail += s_v[i];
    }
    nh += head[j] ? 1 : 0;
  }
  int htot;
  const int hex = block_excl_sum<NT, int>(nh, s_i, &htot);
  const long long gpre = get_prefix(htot);
  float last_incl;
  const float cin = block_seg_excl<NT, float>(nh > 0, tail, s_f, s_fv, &last_incl);
  float acc = cin;
  int hc = hex;
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const int i = i0 + j;
    if (j >= ipt || i >= cnt) break;
    if (head[j]) {
      ++hc;
      acc = s_v[i];
      if (COHERENT)
        __hip_atomic_store(ykey + gpre + hc - 1, (int)s_k[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      else
        ykey[gpre + hc - 1] = (int)s_k[i];
    } else {
      acc += s_v[i];
    }
    const bool end = (i == cnt - 1) || (s_k[i + 1] != s_k[i]);
    if (end) {
      if (hc > 0) {
        if (COHERENT)
          __hip_atomic_store(yval + gpre + hc - 1, acc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else
          yval[gpre + hc - 1] = acc;
      } else {
        carry_val[tile] = acc;
        carry_idx[tile] = (int)(gpre - 1);
      }
    }
  }
  if (threadIdx.x == 0 && carry_idx && (cnt == 0 || head[0])) carry_idx[tile] = -1;
  return gpre + htot;
}

// ------------------------------------------------------------- histogram
// Digit counts of every pass into hist[pass][256] (zeroed by the caller's
// preceding fill); each pass kernel scans its row of it.
// ------------------------------------------------------------ one pass
// A pass is three launches, with no look-back chain between tiles:
//   k_digit_count  per-tile digit counts, digit-major: cnt[d * ntiles + t];
//   k_digit_scan   exclusive scan of that array in place, so that
//                  cnt[d * ntiles + t] becomes the global position of the
//                  first item of digit d in tile t (stable: digit, then tile);
//   k_sort_scatter ranks the tile again and writes every item there.
// (With every tile resident at once, a per-digit decoupled look-back
// spends most of the pass waiting for the prefix to crawl along the tile
// chain; the scan of 256 x ntiles counts costs a few microseconds.)

// Keys of tile t in the wave-major layout of wave_rank.
template <int NT, int IPT>
__device__ __forceinline__ void load_tile_keys(const unsigned* __restrict__ kin, long long n,
                                               long long base, unsigned (&key)[IPT],
                                               bool (&valid)[IPT]) {
  const int w = threadIdx.x / kWave;
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const long long gi = base + (long long)w * kWave * IPT + j * kWave + lane();
    valid[j] = gi < n;
    key[j] = kin[valid[j] ? gi : n - 1];
  }
}

template <int NT, int IPT>
__global__ __launch_bounds__(NT) void k_digit_count(const unsigned* __restrict__ kin,
                                                    const long long* __restrict__ n_dev, int shift,
                                                    unsigned* __restrict__ cnt, int small_cap,
                                                    Gate g) {
  if (gate_shut(g)) return;
  constexpr int TILE = NT * IPT;
  constexpr int NW = NT / kWave;
  static_assert(NT == 256, "one thread per digit");
  __shared__ unsigned s_cnt[NW * 256];
  const long long n = *n_dev;
  if (n <= small_cap) return;
  const unsigned ntiles = (unsigned)((n + TILE - 1) / TILE);
  const int w = threadIdx.x / kWave;
  // tiles are independent: any grid covers them. Only counts are needed
  // here (no ranks): plain LDS atomics into a private histogram per wave
  for (unsigned t = blockIdx.x; t < ntiles; t += gridDim.x) {
    unsigned key[IPT];
    bool valid[IPT];
    load_tile_keys<NT, IPT>(kin, n, (long long)t * TILE, key, valid);
#pragma unroll
    for (int q = 0; q < NW; ++q) s_cnt[q * 256 + threadIdx.x] = 0u;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < IPT; ++j)
      if (valid[j]) atomicAdd(&s_cnt[w * 256 + ((key[j] >> shift) & 255u)], 1u);
    __syncthreads();
    unsigned tot = 0;
#pragma unroll
    for (int q = 0; q < NW; ++q) tot += s_cnt[q * 256 + threadIdx.x];
    cnt[(size_t)threadIdx.x * ntiles + t] = tot;
    __syncthreads();
  }
}

// In-place exclusive scan of cnt[0 .. 256 * ntiles) (ntiles from n and the
// sort tile), ordered tiles of NT * IPT entries with decoupled look-back.
template <int NT, int IPT>
__global__ __launch_bounds__(NT) void k_digit_scan(unsigned* __restrict__ cnt,
                                                   const long long* __restrict__ n_dev, int sort_tile,
                                                   int small_cap, unsigned long long* state, Ctl* ctl,
                                                   Ep epa, Gate g) {
  if (gate_shut(g)) return;
  __shared__ unsigned long long s_pre;
  __shared__ unsigned s_t[NT / kWave];
  const long long n = *n_dev;
  if (n <= small_cap) return;
  const long long E = 256 * ((n + sort_tile - 1) / sort_tile);
  const unsigned nb = (unsigned)((E + NT * IPT - 1) / (NT * IPT));
  const unsigned t = blockIdx.x;
  if (t >= nb) return;
  const unsigned epoch = ep_of(ctl, epa);
  const long long base = (long long)t * NT * IPT + (long long)threadIdx.x * IPT;
  unsigned v[IPT];
  unsigned sum = 0;
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const long long i = base + j;
    v[j] = cnt[i < E ? i : E - 1];
    v[j] = i < E ? v[j] : 0u;
    sum += v[j];
  }
  unsigned btot;
  const unsigned bex = block_excl_sum<NT, unsigned>(sum, s_t, &btot);
  const unsigned long long pre = tile_prefix(state, t, epoch, (unsigned long long)btot, &s_pre);
  unsigned run = (unsigned)pre + bex;
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const long long i = base + j;
    if (i < E) cnt[i] = run;
    run += v[j];
  }
}

template <int NT, int IPT>
__global__ __launch_bounds__(NT) void k_sort_scatter(const unsigned* __restrict__ kin,
                                                     const float* __restrict__ vin,
                                                     unsigned* __restrict__ kout,
                                                     float* __restrict__ vout,
                                                     const long long* __restrict__ n_dev, int shift,
                                                     const unsigned* __restrict__ pos0,
                                                     int small_cap, Gate g) {
  if (gate_shut(g)) return;
  constexpr int TILE = NT * IPT;
  constexpr int NW = NT / kWave;
  static_assert(NT == 256, "one thread per digit");
  __shared__ unsigned s_k[TILE];
  __shared__ float s_v[TILE];
  __shared__ unsigned s_cnt[NW * 256];
  __shared__ unsigned s_dstart[256];
  __shared__ unsigned s_gbase[256];
  __shared__ unsigned s_t[NW];
  const long long n = *n_dev;
  if (n <= small_cap) return;
  const unsigned ntiles = (unsigned)((n + TILE - 1) / TILE);
  const int w = threadIdx.x / kWave;
  const int d = threadIdx.x;
  // tiles are independent: any grid covers them
  for (unsigned t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const long long base = (long long)t * TILE;
    // this tile's global start of digit d, from the scan (issued with the keys)
    const unsigned gb = pos0[(size_t)d * ntiles + t];
    unsigned key[IPT];
    float val[IPT];
    bool valid[IPT];
    int rank[IPT];
    unsigned dig[IPT];
    load_tile_keys<NT, IPT>(kin, n, base, key, valid);
      if (vin) {  // null: keys only
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        const long long gi = base + (long long)w * kWave * IPT + j * kWave + lane();
        val[j] = vin[valid[j] ? gi : n - 1];
      }
    } else {
#pragma unroll
      for (int j = 0; j < IPT; ++j) val[j] = 0.f;
    }
    wave_rank<IPT>(key, valid, shift, s_cnt + w * 256, rank, dig);
    __syncthreads();
    const unsigned tot = digit_wave_offsets<NW>(s_cnt);
    unsigned dtot;
    s_dstart[d] = block_excl_sum<NT, unsigned>(tot, s_t, &dtot);
    s_gbase[d] = gb;
    __syncthreads();
  #pragma unroll
    for (int j = 0; j < IPT; ++j) {
      if (valid[j]) {
        const unsigned p = s_dstart[dig[j]] + s_cnt[w * 256 + dig[j]] + rank[j];
        s_k[p] = key[j];
        s_v[p] = val[j];
      }
    }
    __syncthreads();
    const int cnt = (int)min((long long)TILE, n - base);
    for (int i = threadIdx.x; i < cnt; i += NT) {
      const unsigned k = s_k[i];
      const unsigned dd = (k >> shift) & 255u;
      const unsigned g = s_gbase[dd] + (i - s_dstart[dd]);
      kout[g] = k;
      if (vout) vout[g] = s_v[i];
    }
    __syncthreads();
  }
}

// --------------------------------------------------- single-workgroup path
// n <= TILE: all passes in LDS. mode 0: reduce and write (ykey, yval,
// ctl->ynnz). mode 1: write the sorted pairs to (kout, vout).
template <int NT, int IPT>
__global__ __launch_bounds__(NT) void k_sort_small(const unsigned* __restrict__ kin,
                                                   const float* __restrict__ vin,
                                                   const long long* __restrict__ n_dev, int passes,
                                                   int mode, Ctl* ctl, int* ykey, float* yval,
                                                   unsigned* kout, float* vout, int cap, Gate g) {
  if (gate_shut(g)) return;
  constexpr int TILE = NT * IPT;
  constexpr int NW = NT / kWave;
  __shared__ unsigned s_k[TILE];
  __shared__ float s_v[TILE];
  __shared__ unsigned s_cnt[NW * 256];
  __shared__ unsigned s_dstart[256];
  __shared__ unsigned s_t[NW];
  __shared__ int s_i[NW];
  __shared__ int s_f[NW];
  __shared__ float s_fv[NW];
  const long long n = *n_dev;
  if (n > TILE || n > cap) return;
  const int w = threadIdx.x / kWave;
  // items per thread for this n (<= IPT): every wave ranks its share
  const int ie = (int)((n + NT - 1) / NT);
  unsigned key[IPT];
  float val[IPT];
  bool valid[IPT];
  int rank[IPT];
  unsigned dig[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const int gi = w * kWave * ie + j * kWave + lane();
    valid[j] = j < ie && gi < n;
    key[j] = valid[j] ? kin[gi] : 0u;
    val[j] = valid[j] && vin ? vin[gi] : 0.f;  // vin null: keys only
  }
  for (int p = 0; p < passes; ++p) {
    wave_rank<IPT>(key, valid, 8 * p, s_cnt + w * 256, rank, dig, ie);
    __syncthreads();
    const unsigned tot = digit_wave_offsets<NW>(s_cnt);
    unsigned dtot;
    const unsigned dex = block_excl_sum<NT, unsigned>(tot, s_t, &dtot);
    s_dstart[threadIdx.x] = dex;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      if (valid[j]) {
        const unsigned pos = s_dstart[dig[j]] + s_cnt[w * 256 + dig[j]] + rank[j];
        s_k[pos] = key[j];
        s_v[pos] = val[j];
      }
    }
    __syncthreads();
    if (p + 1 < passes) {
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        const int gi = w * kWave * ie + j * kWave + lane();
        if (valid[j]) {
          key[j] = s_k[gi];
          val[j] = s_v[gi];
        }
      }
      __syncthreads();
    }
  }
  const int cnt = (int)n;
  if (mode == 1) {
    for (int i = threadIdx.x; i < cnt; i += NT) {
      kout[i] = s_k[i];
      if (vout) vout[i] = s_v[i];
    }
    return;
  }
  auto no_prefix = [](int) -> long long { return 0; };
  const long long tot = tile_reduce<NT, IPT>(s_k, s_v, cnt, 0u, true, no_prefix, ykey, yval,
                                             nullptr, nullptr, 0, s_i, s_f, s_fv, ie);
  if (threadIdx.x == 0) {
    ctl->ynnz = (int)tot;
    ctl->sort_small = 1;
  }
}

// ------------------------------------------------------- reduce by key
template <int NT, int IPT>
__global__ __launch_bounds__(NT) void k_reduce(const unsigned* __restrict__ keys,
                                               const float* __restrict__ vals,
                                               const long long* __restrict__ n_dev, Ctl* ctl,
                                               int* ykey, float* yval, float* carry_val,
                                               int* carry_idx, unsigned long long* state, Ep epa,
                                               int small_cap, Gate g) {
  if (gate_shut(g)) return;
  constexpr int TILE = NT * IPT;
  constexpr int NW = NT / kWave;
  __shared__ unsigned s_k[TILE];
  __shared__ float s_v[TILE];
  __shared__ int s_i[NW];
  __shared__ int s_f[NW];
  __shared__ float s_fv[NW];
  __shared__ unsigned long long s_pre;
  const long long n = *n_dev;
  if (n <= small_cap) return;
  const unsigned ntiles = (unsigned)((n + TILE - 1) / TILE);
  const unsigned t = blockIdx.x;
  if (t >= ntiles) return;
  const unsigned epoch = ep_of(ctl, epa);
  {
    const long long base = (long long)t * TILE;
    const int cnt = (int)min((long long)TILE, n - base);
    {
      // all of a thread's loads at once (clamped), then into LDS
      unsigned kk[IPT];
      float vv[IPT];
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        const int i = threadIdx.x + j * NT;
        const long long gi = base + (i < cnt ? i : 0);
        kk[j] = keys[gi];
        vv[j] = vals[gi];
      }
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        const int i = threadIdx.x + j * NT;
        if (i < cnt) {
          s_k[i] = kk[j];
          s_v[i] = vv[j];
        }
      }
    }
    const unsigned prev = base > 0 ? keys[base - 1] : 0u;
    __syncthreads();
    auto prefix = [&](int htot) -> long long {
      return (long long)tile_prefix(state, t, epoch, (unsigned long long)htot, &s_pre);
    };
    const long long end = tile_reduce<NT, IPT>(s_k, s_v, cnt, prev, base == 0, prefix, ykey, yval,
                                               carry_val, carry_idx, t, s_i, s_f, s_fv);
    if (t == ntiles - 1 && threadIdx.x == 0) {
      ctl->ynnz = (int)end;
      ctl->sort_small = 0;
    }
  }
}

__global__ void k_carry_fixup(const float* __restrict__ carry_val, const int* __restrict__ carry_idx,
                              const long long* __restrict__ n_dev, int tile, int small_cap,
                              float* yval, Gate g) {
  if (gate_shut(g)) return;
  const long long n = *n_dev;
  if (n <= small_cap) return;
  const long long ntiles = (n + tile - 1) / tile;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < ntiles;
       t += (long long)gridDim.x * blockDim.x) {
    int idx, h;
    float s;
    if (!carry_run<long long>(carry_val, carry_idx, nullptr, t, ntiles, idx, s, h)) continue;
    yval[idx] += s;
  }
}

static int key_passes(int range) {
  int bits = 1;
  while (bits < 32 && (1u << bits) < (unsigned)range) ++bits;
  return (bits + 7) / 8;
}

// Grid for a launch over n items in tiles of `tile`: exact when the host
// knows n (n_hint >= 0), else the bound n_cap.
static int tiles_grid(long long n_hint, long long n_cap, int tile) {
  const long long n = n_hint >= 0 ? n_hint : n_cap;
  return (int)std::max<long long>(1, (n + tile - 1) / tile);
}

// One counted pass (count, scan, scatter) over *n_dev items; tiles beyond
// the device-side count exit at once. small_cap: passes do nothing when
// n <= small_cap (the single-workgroup path handled them).
static void launch_digit_pass(hipStream_t s, const SortBufs& b, const unsigned* kin, const float* vin,
                              unsigned* kout, float* vout, int shift, const long long* n_dev, Ctl* ctl,
                              long long n_hint, long long n_cap, int small_cap, Ep e, Gate g) {
  const int tiles = tiles_grid(n_hint, n_cap, kSortTile);
  const int pgrid = std::min(tiles, 2048);  // count / scatter loop over their tiles
  const int sgrid = (int)std::max<long long>(1, (256LL * tiles + kDScanTile - 1) / kDScanTile);
  hipLaunchKernelGGL((k_digit_count<kSortNT, kSortIPT>), dim3(pgrid), dim3(kSortNT), 0, s, kin, n_dev,
                     shift, b.dcount, small_cap, g);
  hipLaunchKernelGGL((k_digit_scan<kDScanNT, kDScanIPT>), dim3(sgrid), dim3(kDScanNT), 0, s, b.dcount,
                     n_dev, kSortTile, small_cap, b.dstate, ctl, e, g);
  hipLaunchKernelGGL((k_sort_scatter<kSortNT, kSortIPT>), dim3(pgrid), dim3(kSortNT), 0, s, kin, vin, kout,
                     vout, n_dev, shift, b.dcount, small_cap, g);
}

// Sorts (keyA, valA)[*n_dev] by key < range; returns 0 when the result is
// in A, 1 when in B. ep.dev >= 0 (captured graphs) uses device epochs
// ep.dev + pass; otherwise ep.v + pass.
int launch_radix_sort(hipStream_t s, const SortBufs& b, int range, const long long* n_dev, Ctl* ctl,
                      long long n_hint, long long n_cap, bool skip_small, Ep ep, Gate g, int small_cap) {
  const int passes = key_passes(range);
  const int final_b = passes & 1;
  const bool small_known = n_hint >= 0 && n_hint <= small_cap;
  const bool big_known = n_hint > small_cap;
  if (!big_known && !skip_small) {
    hipLaunchKernelGGL((k_sort_small<kSortNT, kSortIPT>), dim3(1), dim3(kSortNT), 0, s, b.keyA,
                       b.valA, n_dev, passes, 1, ctl, nullptr, nullptr, final_b ? b.keyB : b.keyA,
                       final_b ? b.valB : b.valA, small_cap, g);
  }
  if (small_known) return final_b;
  for (int p = 0; p < passes; ++p) {
    const unsigned* kin = (p & 1) ? b.keyB : b.keyA;
    const float* vin = (p & 1) ? b.valB : b.valA;
    unsigned* kout = (p & 1) ? b.keyA : b.keyB;
    float* vout = (p & 1) ? b.valA : b.valB;
    const Ep e{ep.v + (unsigned)p, ep.dev < 0 ? -1 : ep.dev + p};
    launch_digit_pass(s, b, kin, vin, kout, vout, 8 * p, n_dev, ctl, n_hint, n_cap, small_cap, e, g);
  }
  return final_b;
}

// One stable partition pass of (keyA, valA) into (keyB, valB) on the 8-bit
// digit (key >> shift) & 255. Afterwards b.dcount[d * ntiles] (ntiles =
// ceil(n / kSortTile)) is the start of digit d in the output.
void launch_partition_pass(hipStream_t s, const SortBufs& b, int shift, const long long* n_dev, Ctl* ctl,
                           long long n_hint, long long n_cap, Ep ep) {
  launch_digit_pass(s, b, b.keyA, b.valA, b.keyB, b.valB, shift, n_dev, ctl, n_hint, n_cap, -1, ep, Gate{});
}

void launch_sort_reduce(hipStream_t s, const SortBufs& b, int m, Ctl* ctl, int* ykey, float* yval,
                        long long binlen_hint, long long n_cap, Ep ep, Gate g, int small_cap) {
  const long long* n_dev = &ctl->bin_len;
  const int passes = key_passes(m);
  const bool small_known = binlen_hint >= 0 && binlen_hint <= small_cap;
  const bool big_known = binlen_hint > small_cap;
  if (!big_known) {
    hipLaunchKernelGGL((k_sort_small<kSortNT, kSortIPT>), dim3(1), dim3(kSortNT), 0, s, b.keyA,
                       b.valA, n_dev, passes, 0, ctl, ykey, yval, nullptr, nullptr, small_cap, g);
  }
  if (small_known) return;
  const int fb = launch_radix_sort(s, b, m, n_dev, ctl, binlen_hint, n_cap, true, ep, g, small_cap);
  const unsigned* kf = fb ? b.keyB : b.keyA;
  const float* vf = fb ? b.valB : b.valA;
  const Ep er{ep.v + 4u, ep.dev < 0 ? -1 : ep.dev + 4};
  hipLaunchKernelGGL((k_reduce<kRedNT, kRedIPT>), dim3(tiles_grid(binlen_hint, n_cap, kRedTile)),
                     dim3(kRedNT), 0, s, kf, vf, n_dev, ctl, ykey, yval, b.carry_val, b.carry_idx,
                     b.red_state, er, small_cap, g);
  const int fgrid = std::min(2048, std::max(1, tiles_grid(binlen_hint, n_cap, kRedTile) / 256 + 1));
  hipLaunchKernelGGL(k_carry_fixup, dim3(fgrid), dim3(256), 0, s, b.carry_val, b.carry_idx, n_dev,
                     kRedTile, small_cap, yval, g);
}

// ------------------------------------------------ single-launch small x
// A column SpMSpV whose x has at most kFuseKeys keys and whose products fit
// one workgroup's LDS runs in one launch of one workgroup: the column
// scan of x (as k_prep's single tile, including the merge-path splits, so a
// call that does not fit continues with the usual launches), then, when
// bin_len <= kFuseCap and sparse outputs are given, the gather of every
// product (row, a*x) in x order, a stable LDS radix sort by row, the
// in-order reduce and the sorted output with its count. The result is the
// sort-based slots' (the same products, the same order). In auto mode a
// call that fits skips the network (a size rule in front of it: no slot
// reads less than the products of x); one that does not fit runs the
// selector as k_prep does. When it finishes the call it sets the done flag
// in ctl->kernel_choice (which shuts every gated launch queued behind it),
// writes the count to ctl->ynnz, then to *count_sig and the selector word
// (flagged kSelDone, count in the bin_len field) once the outputs are
// complete.
// One workgroup of 1,024 threads (four waves per SIMD, so that each wave's
// LDS and memory waits overlap three others'); KPT keys and IPT products per
// thread.
template <int NT, int KPT, int IPT>
__global__ __launch_bounds__(NT) void k_col_fused(const int* __restrict__ colptr, const int* __restrict__ crow,
                                                  const float* __restrict__ cval, const int* __restrict__ xkey,
                                                  const float* __restrict__ xval, int ncols, int xnnz, int passes,
                                                  int* __restrict__ ptr, int* __restrict__ xbeg,
                                                  int* __restrict__ split, Ctl* ctl, SelectorArgs sel,
                                                  int* __restrict__ ykey, float* __restrict__ yval,
                                                  long long* count_sig, int cap) {
  static_assert(NT >= 256 && NT % 256 == 0, "a thread per radix digit");
  constexpr int NW = NT / kWave, CAP = NT * IPT, KEYS = NT * KPT;
  constexpr int kRawBytes = CAP * 8 > kNWeights * 4 ? CAP * 8 : kNWeights * 4;
  // the sort's keys and values, or the selector's weights (never both);
  // before the sort, the keys' space holds the product -> x entry map
  __shared__ __align__(16) unsigned char s_raw[kRawBytes];
  __shared__ int s_ptr[KEYS + 1];
  __shared__ int s_beg[KEYS];
  __shared__ unsigned s_cnt[NW * 256];
  __shared__ unsigned s_dstart[256];
  __shared__ unsigned s_t[NW];
  __shared__ long long s_l[NW];
  __shared__ int s_i[NW];
  __shared__ int s_f[NW];
  __shared__ float s_fv[NW];
  __shared__ MatFeat s_mf;
  unsigned* s_k = reinterpret_cast<unsigned*>(s_raw);
  float* s_v = reinterpret_cast<float*>(s_raw + CAP * 4);
  float* s_w = reinterpret_cast<float*>(s_raw);
  int* s_seg = reinterpret_cast<int*>(s_raw);
  // the selector's weights and matrix features, loaded beside the column
  // scan (stored to LDS only when the products do not fit, after the scan:
  // their LDS space is the sort's)
  constexpr int kWPT = (kNWeights + NT - 1) / NT;
  constexpr int kMF = sizeof(MatFeat) / sizeof(double);
  float wr[kWPT];
  double mfr = 0.0;
  if (sel.mode != 2 && sel.mat) {
    if (sel.weights) {
#pragma unroll
      for (int i = 0; i < kWPT; ++i) {
        const int idx = threadIdx.x + i * NT;
        wr[i] = idx < kNWeights ? sel.weights[idx] : 0.f;
      }
    }
    if (threadIdx.x < kMF) mfr = reinterpret_cast<const double*>(sel.mat)[threadIdx.x];
  }
  // ---- column scan of x
  const int base = threadIdx.x * KPT;
  int c[KPT], beg[KPT], len[KPT];
#pragma unroll
  for (int j = 0; j < KPT; ++j) c[j] = xnnz > 0 ? xkey[base + j < xnnz ? base + j : xnnz - 1] : 0;
#pragma unroll
  for (int j = 0; j < KPT; ++j) {
    const int ok = (base + j < xnnz && (unsigned)c[j] < (unsigned)ncols) ? 1 : 0;
    const int ci = c[j] & -ok;
    beg[j] = colptr[ci];
    len[j] = colptr[ci + ok];
    c[j] = ok;
  }
  long long tsum = 0;
  int mx = 0, mn = INT_MAX;
#pragma unroll
  for (int j = 0; j < KPT; ++j) {
    len[j] = c[j] ? len[j] - beg[j] : 0;
    beg[j] = c[j] ? beg[j] : 0;
    if (c[j]) {
      mx = max(mx, len[j]);
      mn = min(mn, len[j]);
    }
    tsum += len[j];
  }
  mx = block_max<NT, int>(mx, s_i);
  mn = block_min<NT, int>(mn, s_i);
  long long bl;
  long long run = block_excl_sum<NT, long long>(tsum, s_l, &bl);
#pragma unroll
  for (int j = 0; j < KPT; ++j) {
    const int i = base + j;
    if (i < xnnz) {
      ptr[i] = (int)run;
      xbeg[i] = beg[j];
      s_ptr[i] = (int)run;
      s_beg[i] = beg[j];
      if (split) {
        const long long Mi = i + run;
        const long long Mn = Mi + len[j] + 1;
        for (long long d = (Mi / kLbTile + 1) * kLbTile; d <= Mn; d += kLbTile) split[d / kLbTile] = i + 1;
      }
      run += len[j];
    }
  }
  if (threadIdx.x == 0) {
    ptr[xnnz] = (int)bl;
    s_ptr[xnnz] = (int)bl;
    ctl->bin_len = bl;
    ctl->max_len = xnnz > 0 ? mx : 0;
    ctl->min_len = mn;
    ctl->xnnz = xnnz;
    const long long nt = ((long long)xnnz + bl + kLbTile - 1) / kLbTile;
    ctl->n_tiles = nt;
    if (split) {
      split[0] = 0;
      split[nt] = xnnz;
    }
  }
  const bool fuse = bl <= CAP && bl <= cap && ykey != nullptr && yval != nullptr;
  if (!fuse) {
    if (sel.mode == 2 || !sel.mat) {
      // a fixed slot (or no selector): the gated launches behind run
      if (threadIdx.x == 0) {
        if (sel.mode == 2) ctl->kernel_choice = sel.fixed;
        if (sel.stash) *sel.stash = sel_word(bl, sel.mode == 2 ? sel.fixed : -1);
      }
      return;
    }
    // the selector, as the last tile of k_prep runs it
    {
      if (sel.weights) {
#pragma unroll
        for (int i = 0; i < kWPT; ++i) {
          const int idx = threadIdx.x + i * NT;
          if (idx < kNWeights) s_w[wlds_index(idx)] = wr[i];
        }
      }
      if (threadIdx.x < kMF) reinterpret_cast<double*>(&s_mf)[threadIdx.x] = mfr;
      __syncthreads();
      SelectorArgs sl = sel;
      if (sel.weights) sl.weights = s_w;
      sl.mat = &s_mf;
      select_in_wave(ctl, sl, (double)xnnz, (double)bl, (double)(xnnz > 0 ? mx : 0));
    }
    if (sel.stash && threadIdx.x == 0) *sel.stash = sel_word(bl, ctl->kernel_choice);
    return;
  }
  __syncthreads();  // s_ptr / s_beg complete
  // products per thread for this call (<= IPT): every wave gets its share,
  // so that the ranking's per-slot round trips run on all the waves at once
  // instead of IPT of them in a row on the first few
  const int ie = (int)((bl + NT - 1) / NT);
  // ---- product -> x entry map, merge-path style: thread t fills products
  // t*ie .. t*ie+ie-1 from one search and a forward walk (x entries with
  // empty columns are stepped over)
  {
    const int p0 = threadIdx.x * ie;
    if (p0 < bl) {
      int lo = 0, hi = xnnz;  // s_ptr[lo] <= p0 < s_ptr[hi]
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (s_ptr[mid] <= p0) lo = mid;
        else hi = mid;
      }
      int e = lo, nxt = s_ptr[lo + 1];
#pragma unroll
      for (int k = 0; k < IPT; ++k) {
        const int p = p0 + k;
        if (k < ie && p < bl) {
          while (nxt <= p) nxt = s_ptr[++e + 1];
          s_seg[p] = e;
        }
      }
    }
  }
  __syncthreads();
  // ---- products into registers, in x order (wave-major items as wave_rank takes them)
  const int w = threadIdx.x / kWave;
  unsigned key[IPT];
  float val[IPT];
  bool valid[IPT];
  int rank[IPT];
  unsigned dig[IPT];
  int pos[IPT];
  float xv[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const int p = w * kWave * ie + j * kWave + lane();
    valid[j] = j < ie && p < bl;
    const int e = valid[j] ? s_seg[p] : 0;
    pos[j] = valid[j] ? s_beg[e] + (p - s_ptr[e]) : 0;
    xv[j] = valid[j] ? xval[e] : 0.f;
  }
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    key[j] = valid[j] ? (unsigned)crow[pos[j]] : 0u;
    val[j] = valid[j] ? cval[pos[j]] * xv[j] : 0.f;
  }
  __syncthreads();  // every map entry read before the sort reuses the space
  // ---- LDS radix sort (stable), as k_sort_small; the digit offsets by
  // threads 0..255 (one per digit)
  for (int pss = 0; pss < passes; ++pss) {
    wave_rank<IPT>(key, valid, 8 * pss, s_cnt + w * 256, rank, dig, ie);
    __syncthreads();
    unsigned tot = 0;
    if (threadIdx.x < 256) tot = digit_wave_offsets<NW>(s_cnt);
    unsigned dtot;
    const unsigned dex = block_excl_sum<NT, unsigned>(tot, s_t, &dtot);
    if (threadIdx.x < 256) s_dstart[threadIdx.x] = dex;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      if (valid[j]) {
        const unsigned q = s_dstart[dig[j]] + s_cnt[w * 256 + dig[j]] + rank[j];
        s_k[q] = key[j];
        s_v[q] = val[j];
      }
    }
    __syncthreads();
    if (pss + 1 < passes) {
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        const int gi = w * kWave * ie + j * kWave + lane();
        if (valid[j]) {
          key[j] = s_k[gi];
          val[j] = s_v[gi];
        }
      }
      __syncthreads();
    }
  }
  auto no_prefix = [](int) -> long long { return 0; };
  const long long tot = tile_reduce<NT, IPT, decltype(no_prefix), true>(s_k, s_v, (int)bl, 0u, true, no_prefix, ykey, yval, nullptr,
                                             nullptr, 0, s_i, s_f, s_fv, ie);
  // ---- publish: outputs complete, then the count and the done flag
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  if (threadIdx.x == 0) {
    const int choice = sel.mode == 2 ? sel.fixed : kFusedSlot;
    ctl->ynnz = (int)tot;
    ctl->sort_small = 1;
    ctl->kernel_choice = choice | kSelDone;
    // the outputs are device-coherent and acknowledged (the wait above, in
    // every thread, before the barrier): no L2 write-back before the words
    if (count_sig) *count_sig = tot;
    if (sel.stash) *sel.stash = sel_word(tot, choice | kSelDone);
  }
}

bool launch_col_fused(hipStream_t s, const int* colptr, const int* crow, const float* cval, const int* xkey,
                      const float* xval, int ncols, int xnnz, int m, int* ptr, int* xbeg, int* split, Ctl* ctl,
                      const SelectorArgs& sel, int* ykey, float* yval, long long* count_sig, int cap) {
  if (xnnz > kFuseKeys) return false;
  hipLaunchKernelGGL((k_col_fused<kFuseNT, kFuseKeys / kFuseNT, kFuseCap / kFuseNT>), dim3(1), dim3(kFuseNT), 0, s,
                     colptr, crow, cval, xkey, xval, ncols, xnnz, key_passes(m), ptr, xbeg, split, ctl, sel, ykey,
                     yval, count_sig, cap);
  return true;
}

}  // namespace spmspv
