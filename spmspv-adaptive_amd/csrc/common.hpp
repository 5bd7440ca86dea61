// Shared device primitives for the gfx950 SpMSpV engine: wave64 scans and
// reductions, the single-pass decoupled look-back tile prefix used by every
// scan-like kernel, and the per-handle control block.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spmspv {

constexpr int kWave = 64;

// ---------------------------------------------------------------- control
// One per handle, in device memory. Persistent kernels take tiles from
// `ticket[slot]`; the last workgroup to finish (counted in `done[slot]`)
// resets both and advances `epoch`, so a kernel needs no memset before it
// (which keeps every launch capturable in a hipGraph).
constexpr int kSlots = 16;
struct Ctl {
  unsigned int ticket[kSlots];
  unsigned int done[kSlots];
  unsigned int epoch;
  unsigned int pad0;
  long long bin_len;   // sum of column lengths over x's support
  int max_len;         // max column length over x's support
  int min_len;         // min column length over x's support
  int xnnz;            // current x support size (device copy)
  int ynnz;            // support size of the last result
  int sort_small;      // 1 when the last sort took the single-workgroup path
  int kernel_choice;   // last selector output
  long long n_tiles;   // tiles of the last load-balanced gather
  int frontier;        // BFS: size of the next frontier
  int level;           // BFS: current level
  float feat[16];      // last selector feature vector
  int pad[6];
};

// Look-back status word: [63:62] flag (1 aggregate, 2 inclusive),
// [61:38] epoch (24 bits), [37:0] value.
constexpr int kEpochBits = 24;
constexpr unsigned long long kValMask = (1ull << 38) - 1;
__device__ __forceinline__ unsigned long long pack_state(unsigned flag, unsigned epoch,
                                                         unsigned long long v) {
  return ((unsigned long long)flag << 62) |
         ((unsigned long long)(epoch & ((1u << kEpochBits) - 1)) << 38) | (v & kValMask);
}

__device__ __forceinline__ void st_relaxed(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long ld_relaxed(unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------- wave ops
__device__ __forceinline__ int lane() { return __lane_id(); }

template <typename T>
__device__ __forceinline__ T wave_incl_sum(T v) {
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    T t = __shfl_up(v, o, kWave);
    if (lane() >= o) v += t;
  }
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    T t = __shfl_xor(v, o, kWave);
    v = t > v ? t : v;
  }
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_min(T v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    T t = __shfl_xor(v, o, kWave);
    v = t < v ? t : v;
  }
  return v;
}

// ---- DPP lane moves (gfx9 data-parallel primitives): a lane whose source
// lane is outside the pattern, or whose row is masked off, keeps `old`.
constexpr int kDppRowShr1 = 0x111, kDppRowShr2 = 0x112, kDppRowShr4 = 0x114, kDppRowShr8 = 0x118;
constexpr int kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143, kDppWaveShr1 = 0x138;

template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_i32(int old, int v) {
  return __builtin_amdgcn_update_dpp(old, v, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_f32(float old, float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old),
                                                               __builtin_bit_cast(int, v), CTRL,
                                                               ROWS, 0xf, false));
}

// One step of a segmented inclusive scan: (f, v, h) of this lane combined
// with the running value of the lane CTRL points at. f marks that a
// segment starts inside the lane's range; v sums and h ORs up to it.
template <int CTRL, int ROWS>
__device__ __forceinline__ void seg_step(int& f, float& v, int& h) {
  const int fo = dpp_i32<CTRL, ROWS>(0, f);
  const float vo = dpp_f32<CTRL, ROWS>(0.f, v);
  const int ho = dpp_i32<CTRL, ROWS>(0, h);
  if (!f) {
    v += vo;
    h |= ho;
  }
  f |= fo;
}

// Wave-wide segmented inclusive scan in six DPP steps (rows of 16 lanes,
// then the row broadcasts of lane 15 and lane 31).
__device__ __forceinline__ void wave_seg_scan(int& f, float& v, int& h) {
  seg_step<kDppRowShr1, 0xf>(f, v, h);
  seg_step<kDppRowShr2, 0xf>(f, v, h);
  seg_step<kDppRowShr4, 0xf>(f, v, h);
  seg_step<kDppRowShr8, 0xf>(f, v, h);
  seg_step<kDppRowBcast15, 0xa>(f, v, h);
  seg_step<kDppRowBcast31, 0xc>(f, v, h);
}

// Block-wide exclusive sum of one value per thread; returns the exclusive
// prefix and writes the block total to *total. `lds` needs NT/64 slots.
template <int NT, typename T>
__device__ __forceinline__ T block_excl_sum(T v, T* lds, T* total) {
  constexpr int NW = NT / kWave;
  const int w = threadIdx.x / kWave;
  T inc = wave_incl_sum(v);
  if (lane() == kWave - 1) lds[w] = inc;
  __syncthreads();
  T wpre = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < NW; ++i) {
    T s = lds[i];
    if (i < w) wpre += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return wpre + inc - v;
}

template <int NT, typename T>
__device__ __forceinline__ T block_sum(T v, T* lds) {
  constexpr int NW = NT / kWave;
  v = wave_sum(v);
  if (lane() == 0) lds[threadIdx.x / kWave] = v;
  __syncthreads();
  T tot = 0;
#pragma unroll
  for (int i = 0; i < NW; ++i) tot += lds[i];
  __syncthreads();
  return tot;
}

template <int NT, typename T>
__device__ __forceinline__ T block_max(T v, T* lds) {
  constexpr int NW = NT / kWave;
  v = wave_max(v);
  if (lane() == 0) lds[threadIdx.x / kWave] = v;
  __syncthreads();
  T r = lds[0];
#pragma unroll
  for (int i = 1; i < NW; ++i) r = lds[i] > r ? lds[i] : r;
  __syncthreads();
  return r;
}

template <int NT, typename T>
__device__ __forceinline__ T block_min(T v, T* lds) {
  constexpr int NW = NT / kWave;
  v = wave_min(v);
  if (lane() == 0) lds[threadIdx.x / kWave] = v;
  __syncthreads();
  T r = lds[0];
#pragma unroll
  for (int i = 1; i < NW; ++i) r = lds[i] < r ? lds[i] : r;
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------- tickets
// Persistent-tile helpers. Every workgroup loops: t = next_tile(); if t >=
// ntiles break. The ticket order guarantees that when a workgroup holds tile
// t, every tile < t is held by a running (or finished) workgroup, so the
// look-back below always terminates.
__device__ __forceinline__ unsigned next_tile(Ctl* ctl, int slot, unsigned* s_tile) {
  if (threadIdx.x == 0) *s_tile = atomicAdd(&ctl->ticket[slot], 1u);
  __syncthreads();
  unsigned t = *s_tile;
  __syncthreads();
  return t;
}

// Called once per workgroup after its last tile. The last workgroup resets
// the slot and bumps the epoch. Returns true in the last workgroup (all
// threads), after every other workgroup's writes are visible.
__device__ __forceinline__ bool finish_persistent(Ctl* ctl, int slot, unsigned* s_flag) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    unsigned d = atomicAdd(&ctl->done[slot], 1u);
    unsigned last = (d == gridDim.x - 1) ? 1u : 0u;
    if (last) {
      __threadfence();
      atomicExch(&ctl->ticket[slot], 0u);
      atomicExch(&ctl->done[slot], 0u);
      atomicAdd(&ctl->epoch, 1u);
    }
    *s_flag = last;
  }
  __syncthreads();
  bool last = *s_flag != 0;
  __syncthreads();
  return last;
}

// Epoch of a look-back launch. Host-driven launches pass a fresh value per
// launch (bit 23 clear); launches inside captured graphs, whose arguments
// are frozen, take ctl->epoch + dev, which a small kernel at the start of
// each captured sequence advances (bit 23 set). Stale states of earlier
// launches then never match.
struct Ep {
  unsigned v;
  int dev;
};
__device__ __forceinline__ unsigned ep_of(const Ctl* ctl, Ep e) {
  const unsigned mask = (1u << kEpochBits) - 1;
  return e.dev < 0 ? (e.v & mask) : (((ctl->epoch + (unsigned)e.dev) | 0x800000u) & mask);
}

// Ordered tiles: tile = blockIdx.x. Workgroups are dispatched in
// increasing index order, so every tile a workgroup waits on in the
// look-back below belongs to a workgroup that was dispatched earlier and
// runs to completion without waiting on later ones.
//
// Decoupled look-back: given this tile's aggregate (thread 0's argument is
// used), returns the exclusive prefix of all earlier tiles, to every thread.
// `state` holds one word per tile. Wave 0 does the look-back, 64 tiles at a
// time. Must be called by all threads of the block.
__device__ __forceinline__ unsigned long long tile_prefix(unsigned long long* state, unsigned tile,
                                                          unsigned epoch, unsigned long long agg,
                                                          unsigned long long* s_pre) {
  const unsigned ep = epoch & ((1u << kEpochBits) - 1);
  if (threadIdx.x == 0) {
    if (tile == 0) st_relaxed(&state[0], pack_state(2, ep, agg));
    else st_relaxed(&state[tile], pack_state(1, ep, agg));
  }
  if (threadIdx.x < kWave) {
    unsigned long long excl = 0;
    if (tile > 0) {
      long long base = (long long)tile - 1;  // highest tile index of the window
      while (true) {
        long long idx = base - lane();
        unsigned long long s = 0;
        unsigned flag = 0;
        if (idx >= 0) {
          while (true) {
            s = ld_relaxed(&state[idx]);
            flag = (unsigned)(s >> 62);
            if (((s >> 38) & ((1u << kEpochBits) - 1)) != ep) flag = 0;
            if (flag) break;
            __builtin_amdgcn_s_sleep(1);  // back off instead of hammering the line
          }
        } else {
          flag = 2;  // before tile 0: acts as an inclusive zero
          s = 0;
        }
        unsigned long long v = (idx >= 0) ? (s & kValMask) : 0ull;
        // lanes ordered from nearest (lane 0) to farthest; stop at the first
        // inclusive one.
        uint64_t incl = __ballot(flag == 2);
        int stop = incl ? __builtin_ctzll(incl) : kWave;  // first inclusive lane
        unsigned long long contrib = (lane() <= stop) ? v : 0ull;
        contrib = wave_sum(contrib);
        excl += contrib;
        if (incl) break;
        base -= kWave;
      }
      if (threadIdx.x == 0) st_relaxed(&state[tile], pack_state(2, ep, excl + agg));
    }
    if (threadIdx.x == 0) *s_pre = excl;
  }
  __syncthreads();
  unsigned long long r = *s_pre;
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------- misc
__device__ __forceinline__ int upper_bound_lds(const int* a, int lo, int hi, int key) {
  // first index in [lo,hi) with a[i] > key
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (a[mid] <= key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ void atomic_add_f32(float* p, float v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Carry runs of a tiled segmented reduction: tile t left the partial sum
// cv[t] (hit flag chit[t], nullable) of row crow[t] (-1: none), and
// consecutive tiles with the same row form one run. For the first tile of
// a run, sums the run in tile order into (s, h) and returns true. The
// loads of the tile, its neighbours and its own carry are issued together:
// one round trip to memory for the common run of one or two tiles; longer
// runs (a row over many tiles) are read kCarryBatch tiles per round trip.
constexpr int kCarryBatch = 16;
template <typename Idx>
__device__ __forceinline__ bool carry_run(const float* __restrict__ cv, const int* __restrict__ crow,
                                          const int* __restrict__ chit, Idx t, Idx ntiles, int& r,
                                          float& s, int& h) {
  const Idx tp = t > 0 ? t - 1 : 0;
  const Idx tn = t + 1 < ntiles ? t + 1 : t;
  r = crow[t];
  const int rp = crow[tp];
  const int rn = crow[tn];
  const float v = cv[t];
  h = chit ? chit[t] : 0;
  // keeps the compiler from sinking the loads of v, rn and h below the test
  asm volatile("" ::"v"(v), "v"(rn), "v"(h), "v"(rp));
  if (r < 0 || (t > 0 && rp == r)) return false;
  s = 0.f + v;
  if (!(t + 1 < ntiles && rn == r)) return true;
  // the run goes on at t + 1: batches of tiles, loads issued together,
  // summed in tile order until the row changes
  for (Idx u = t + 1;; u += kCarryBatch) {
    int br[kCarryBatch], bh[kCarryBatch];
    float bv[kCarryBatch];
#pragma unroll
    for (int k = 0; k < kCarryBatch; ++k) {
      const Idx q = u + k < ntiles ? u + k : ntiles - 1;
      br[k] = crow[q];
      bv[k] = cv[q];
      bh[k] = chit ? chit[q] : 0;
    }
#pragma unroll
    for (int k = 0; k < kCarryBatch; ++k) asm volatile("" ::"v"(br[k]), "v"(bv[k]), "v"(bh[k]));
    bool go = true;
#pragma unroll
    for (int k = 0; k < kCarryBatch; ++k) {
      go = go && u + k < ntiles && br[k] == r;
      if (go) {
        s += bv[k];
        h |= bh[k];
      }
    }
    if (!go) return true;
  }
}

}  // namespace spmspv
