// Level-synchronous BFS on the boolean (OR, AND) semiring, one SpMSpV per
// level with the frontier as sparse x. The reference has no BFS loop: it
// replays frontier files recorded elsewhere (la/test/test_application.cu:
// 207-261, frontier format readsvmdata.hpp:308-348), so this driver is new.
//
// A level runs the boolean form of the slot chosen for it (row and column
// indices only: no matrix value is read and no FP32 arithmetic is done):
//   K0/K2  top-down: frontier vertices scan their columns and claim
//          unvisited rows (naive team per vertex / merge-path balanced);
//   K1/K3  top-down gather of all neighbours' row keys, in-house radix sort
//          of the keys alone, then each distinct row claimed;
//   K4     bottom-up: each unvisited row (one lane) scans its neighbours
//          and stops at the first frontier member;
//   K5/K7  merge-path row tiles OR-ing the frontier bits of each row's
//          columns into a hit bit, then claim (on booleans the masked and
//          unmasked row kernels coincide);
//   K6     naive row scan without early exit, hit bits, then claim.
// Levels are int32 and unique, so any order of claims gives the same
// levels. The selector for level L+1 runs in the same stream right after
// level L (features of the new frontier + MLP, one launch), and a last
// one-wave kernel writes the next frontier's size and choice, tagged with
// the level counter, into one host-coherent word that the host polls (no
// stream wait, no copies); with graphs on, each (slot, parity) level
// sequence is captured once as a hipGraph and replayed. Every kernel reads
// the frontier size from device memory, which is what makes the capture
// valid for any level.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>

#include "handle.hpp"
#include "selector.hpp"

using namespace spmspv;

struct BfsState {
  unsigned* visited = nullptr;              // n bits
  unsigned* fbits[2] = {nullptr, nullptr};  // frontier bit arrays
  int* front[2] = {nullptr, nullptr};       // frontier lists
  int* cnt = nullptr;                       // [0],[1] sizes, [2] level, [3] bottom-up queue size
  int* queue = nullptr;                     // m: rows the bottom-up head left undecided
  unsigned long long* reached = nullptr;    // [0] vertices, [1] degree sum
  void* alloc = nullptr;
  hipStream_t own_stream = nullptr;
  int* levels_bound = nullptr;              // levels array the graphs were captured with
  int bound_mode = -9, bound_fixed = -9;    // selector settings captured in the graphs
  const float* bound_weights = nullptr;
  hipGraphExec_t graph[8][2] = {};
};

namespace {

constexpr int kBfsGrid = 2048;
constexpr int kBuBatch = 8;  // neighbours each row tests in the bottom-up head pass
constexpr int kBuTeam = 8;   // lanes per queued row in the bottom-up tail pass

__device__ __forceinline__ int wave_append(int* counter, bool pred) {
  const uint64_t b = __ballot(pred);
  if (!b) return -1;
  const int leader = __builtin_ctzll(b);
  int base = 0;
  if (lane() == leader) base = atomicAdd(counter, __popcll(b));
  base = __shfl(base, leader, kWave);
  const uint64_t lt = lane() == 0 ? 0ull : (~0ull >> (kWave - lane()));
  return pred ? base + __popcll(b & lt) : -1;
}

__device__ __forceinline__ bool try_claim(unsigned* visited, int v) {
  const unsigned bit = 1u << (v & 31);
  if (__hip_atomic_load(visited + (v >> 5), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & bit)
    return false;
  return !(atomicOr(visited + (v >> 5), bit) & bit);
}

// Called by every lane of a wave (convergent): appends the claimed vertices.
__device__ __forceinline__ void record(int v, bool claimed, int level, int* levels, int* nf,
                                       int* ncnt, unsigned* nbits) {
  const int pos = wave_append(ncnt, claimed);
  if (claimed) {
    levels[v] = level;
    nf[pos] = v;
    atomicOr(nbits + (v >> 5), 1u << (v & 31));
  }
}

__global__ void k_bfs_init(int* levels, int n, unsigned* visited, unsigned* fb0, unsigned* fb1,
                           int* front0, int* cnt, unsigned long long* reached, int src) {
  const int words = (n + 31) >> 5;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    levels[i] = i == src ? 0 : -1;
  for (int w = blockIdx.x * blockDim.x + threadIdx.x; w < words; w += gridDim.x * blockDim.x) {
    const unsigned b = (w == (src >> 5)) ? (1u << (src & 31)) : 0u;
    visited[w] = b;
    fb0[w] = b;
    fb1[w] = 0u;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    front0[0] = src;
    cnt[0] = 1;
    cnt[1] = 0;
    cnt[2] = 0;
    cnt[3] = 0;
    reached[0] = 0;
    reached[1] = 0;
  }
}

// Opens a level: empties the next frontier, advances the level number and
// the device epoch that the level's look-back kernels use (their launch
// arguments are frozen inside a captured graph).
// The next level's frontier size and selector choice, tagged with the level
// counter, into one host-coherent word (sig[0]), its selector features into
// sig[1..8] first: the host polls one word per level instead of waiting for
// the stream and three copies.
__global__ void k_bfs_signal(const int* __restrict__ cnt, int nxt, const Ctl* __restrict__ ctl, long long* sig) {
  // the feature words are acknowledged (each writer waits for its store)
  // before the level word goes out; the host reads nothing else, so no L2
  // write-back is needed
  if (threadIdx.x < 16) {
    reinterpret_cast<float*>(sig + 1)[threadIdx.x] = ctl->feat[threadIdx.x];
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const long long w = ((long long)(cnt[2] & 0xffffff) << 40) | ((long long)(ctl->kernel_choice & 0xff) << 32) |
                        (long long)(unsigned)cnt[nxt];
    __hip_atomic_store(sig, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

__global__ void k_bfs_begin(int* cnt, int nxt, unsigned* nbits, int words, Ctl* ctl) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    cnt[nxt] = 0;
    cnt[2] += 1;
    cnt[3] = 0;
    ctl->epoch += 16;
  }
  for (int w = blockIdx.x * blockDim.x + threadIdx.x; w < words; w += gridDim.x * blockDim.x)
    nbits[w] = 0u;
}

// K0: top-down, a team of 2^lg lanes per frontier vertex; loops are
// wave-uniform so that wave_append stays convergent.
__global__ __launch_bounds__(256) void k_td_naive(const int* __restrict__ colptr,
                                                  const int* __restrict__ crow,
                                                  const int* __restrict__ f, int* cnt, int cur,
                                                  int lg, unsigned* visited, int* levels, int* nf,
                                                  unsigned* nbits) {
  const int fsz = cnt[cur];
  const int level = cnt[2];
  int* ncnt = cnt + (cur ^ 1);
  const int team = 1 << lg, tpw = kWave >> lg;
  const int tl = threadIdx.x & (team - 1);
  const int tin = (threadIdx.x & (kWave - 1)) >> lg;
  const long long gw = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const long long nw = ((long long)gridDim.x * blockDim.x) >> 6;
  for (long long base = gw * tpw; base < fsz; base += nw * tpw) {
    const long long e = base + tin;
    int b = 0, len = 0;
    if (e < fsz) {
      const int u = f[e];
      b = colptr[u];
      len = colptr[u + 1] - b;
    }
    const int mx = wave_max(len);
    for (int k0 = 0; k0 < mx; k0 += team) {
      const int k = k0 + tl;
      bool cl = false;
      int v = 0;
      if (k < len) {
        v = crow[b + k];
        cl = try_claim(visited, v);
      }
      record(v, cl, level, levels, nf, ncnt, nbits);
    }
  }
}

// K2: top-down over merge-path tiles of (frontier entries + edges), using
// the splits k_prep wrote for the frontier.
template <int NT, int VT>
__global__ __launch_bounds__(NT) void k_td_lb(const int* __restrict__ ptr,
                                              const int* __restrict__ xbeg,
                                              const int* __restrict__ split,
                                              const int* __restrict__ crow, const Ctl* ctl,
                                              int* cnt, int cur, unsigned* visited, int* levels,
                                              int* nf, unsigned* nbits) {
  constexpr int T = NT * VT;
  __shared__ int s_ptr[T + 2];
  __shared__ int s_beg[T + 2];
  __shared__ int s_seg[T];
  const int xnnz = cnt[cur];
  const int level = cnt[2];
  int* ncnt = cnt + (cur ^ 1);
  const long long ntiles = ctl->n_tiles;
  const long long W = (long long)xnnz + ctl->bin_len;
  for (long long t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const long long d0 = t * T, d1 = min(d0 + T, W);
    const int a0 = split[t], a1 = split[t + 1];
    const long long b0 = d0 - a0, b1 = d1 - a1;
    const int sa = a0 > 0 ? a0 - 1 : 0;
    const int nseg = a1 - sa + 1;
    for (int k = threadIdx.x; k < nseg; k += NT) {
      s_ptr[k] = ptr[sa + k];
      if (sa + k < xnnz) s_beg[k] = xbeg[sa + k];
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
      for (int k = 0; k < VT && D < d1; ++k, ++D) {
        if (i < a1 && (long long)i + s_ptr[i - sa] == D) ++i;
        else s_seg[(p++) - b0] = i - 1 - sa;
      }
    }
    __syncthreads();
    const int np = (int)(b1 - b0);
    for (int q0 = 0; q0 < np; q0 += NT) {
      const int q = q0 + threadIdx.x;
      bool cl = false;
      int v = 0;
      if (q < np) {
        const int sg = s_seg[q];
        v = crow[s_beg[sg] + (int)(b0 + q - s_ptr[sg])];
        cl = try_claim(visited, v);
      }
      record(v, cl, level, levels, nf, ncnt, nbits);
    }
    __syncthreads();
  }
}

// After the sort slots: claims the unique neighbour keys.
// Claims the distinct rows of a sorted key list of *n_dev items (repeats
// are skipped without a test-and-set).
__global__ void k_claim_sorted(const unsigned* __restrict__ keys, const long long* __restrict__ n_dev,
                               int* cnt, int cur, unsigned* visited, int* levels, int* nf, unsigned* nbits) {
  const long long k = *n_dev;
  const int level = cnt[2];
  int* ncnt = cnt + (cur ^ 1);
  for (long long i0 = (long long)blockIdx.x * blockDim.x; i0 < k; i0 += (long long)gridDim.x * blockDim.x) {
    const long long i = i0 + threadIdx.x;
    bool cl = false;
    int v = 0;
    if (i < k) {
      v = (int)keys[i];
      if (i == 0 || keys[i - 1] != (unsigned)v) cl = try_claim(visited, v);
    }
    record(v, cl, level, levels, nf, ncnt, nbits);
  }
}

// K4 as the bottom-up step of a direction-optimising BFS, in two passes.
// Head: one lane per unvisited row tests its first B neighbours (the column
// loads, then the frontier-bit loads, each issued together) and claims the
// row at a hit; a row with no hit among more than B neighbours goes to a
// queue. Tail: a team of kBuTeam lanes per queued row tests the rest, kBuTeam
// neighbours per step, and stops at the first hit. A team per row for every
// row keeps a whole wave on a row until each of its lanes is done (64 ms for
// a small-frontier level of R-MAT s22 with one wave per row), and one lane
// per row for every row leaves a wave waiting on its longest row (86 ms);
// in R-MAT rows the hubs come first (low ids), so once the frontier holds
// them most rows stop in the head.
template <int B>
__global__ __launch_bounds__(256) void k_bu_head(const int* __restrict__ row_ptr,
                                                 const int* __restrict__ col, int m,
                                                 const unsigned* __restrict__ fbits, int* cnt, int cur,
                                                 unsigned* visited, int* levels, int* nf, unsigned* nbits,
                                                 int* queue) {
  const int level = cnt[2];
  int* ncnt = cnt + (cur ^ 1);
  for (long long base = (long long)blockIdx.x * blockDim.x; base < m;
       base += (long long)gridDim.x * blockDim.x) {
    const long long r = base + threadIdx.x;
    bool active = r < m;
    if (active) active = !((visited[r >> 5] >> (r & 31)) & 1u);
    int b = 0, e = 0;
    if (active) {
      b = row_ptr[r];
      e = row_ptr[r + 1];
    }
    const bool go = active && b < e;
    int u[B];
#pragma unroll
    for (int j = 0; j < B; ++j) u[j] = col[go ? min(b + j, e - 1) : 0];
    unsigned w[B];
#pragma unroll
    for (int j = 0; j < B; ++j) w[j] = fbits[(unsigned)u[j] >> 5];
    bool found = false;
#pragma unroll
    for (int j = 0; j < B; ++j) found = found || (go && b + j < e && ((w[j] >> (u[j] & 31)) & 1u));
    const bool cl = go && found;
    if (cl) atomicOr(visited + (r >> 5), 1u << (r & 31));
    record((int)r, cl, level, levels, nf, ncnt, nbits);
    const bool more = go && !found && e - b > B;
    const int q = wave_append(cnt + 3, more);
    if (more) queue[q] = (int)r;
  }
}

template <int B, int TEAM>
__global__ __launch_bounds__(256) void k_bu_tail(const int* __restrict__ row_ptr,
                                                 const int* __restrict__ col,
                                                 const unsigned* __restrict__ fbits, int* cnt, int cur,
                                                 unsigned* visited, int* levels, int* nf, unsigned* nbits,
                                                 const int* __restrict__ queue) {
  const int level = cnt[2];
  int* ncnt = cnt + (cur ^ 1);
  const int nq = cnt[3];
  constexpr int TPW = kWave / TEAM;
  const int tl = threadIdx.x & (TEAM - 1);
  const int tin = (threadIdx.x & (kWave - 1)) / TEAM;
  const long long gw = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const long long nw = ((long long)gridDim.x * blockDim.x) >> 6;
  for (long long q0 = gw * TPW; q0 < nq; q0 += nw * TPW) {
    const long long qi = q0 + tin;
    int r = 0, b = 0, e = 0;
    if (qi < nq) {
      r = queue[qi];
      b = row_ptr[r] + B;
      e = row_ptr[r + 1];
    }
    bool found = false;
    // wave-uniform loop: every team of the wave steps until all are done
    for (int k0 = b; __ballot(!found && k0 < e) != 0ull; k0 += TEAM) {
      const int k = k0 + tl;
      const bool ok = !found && k < e;
      const int v = col[ok ? k : 0];
      int h = ok ? (int)((fbits[(unsigned)v >> 5] >> (v & 31)) & 1u) : 0;
#pragma unroll
      for (int o = TEAM >> 1; o > 0; o >>= 1) h |= __shfl_xor(h, o, kWave);
      found = found || h;
    }
    const bool cl = qi < nq && found && tl == 0;
    if (cl) atomicOr(visited + (r >> 5), 1u << (r & 31));
    record(r, cl, level, levels, nf, ncnt, nbits);
  }
}

// Claims rows whose hit bit a row kernel set (hits & ~visited); clears hits.
__global__ void k_claim_bits(unsigned* hits, int m, int* cnt, int cur, unsigned* visited,
                             int* levels, int* nf, unsigned* nbits) {
  const int level = cnt[2];
  int* ncnt = cnt + (cur ^ 1);
  const int words = (m + 31) >> 5;
  for (long long w0 = (long long)blockIdx.x * blockDim.x; w0 < words; w0 += (long long)gridDim.x * blockDim.x) {
    const long long w = w0 + threadIdx.x;
    unsigned nw = 0;
    if (w < words) {
      const unsigned h = hits[w];
      if (h) {
        hits[w] = 0u;
        nw = h & ~visited[w];
        if (nw) {
          visited[w] |= nw;  // one thread per word: no race
          nbits[w] |= nw;
        }
      }
    }
    const int c = __popc(nw);
    const int inc = wave_incl_sum(c);
    const int tot = __shfl(inc, kWave - 1, kWave);
    int base = 0;
    if (lane() == kWave - 1 && tot) base = atomicAdd(ncnt, tot);
    base = __shfl(base, kWave - 1, kWave);
    int pos = base + inc - c;
    while (nw) {
      const int bit = __ffs(nw) - 1;
      nw &= nw - 1;
      const int v = (int)(w << 5) + bit;
      levels[v] = level;
      nf[pos++] = v;
    }
  }
}

// Graph500 edge count: sum of degrees over reached vertices.
__global__ void k_reached(const int* __restrict__ levels, const int* __restrict__ colptr, int n,
                          unsigned long long* out) {
  unsigned long long c = 0, d = 0;
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    if (levels[v] >= 0) {
      ++c;
      d += (unsigned long long)(colptr[v + 1] - colptr[v]);
    }
  c = wave_sum(c);
  d = wave_sum(d);
  if (lane() == 0) {
    atomicAdd(out, c);
    atomicAdd(out + 1, d);
  }
}

int lg2(int t) {
  int l = 0;
  while ((1 << l) < t && l < 6) ++l;
  return l;
}

// Enqueues one level on stream s: begin, the slot's kernels on frontier
// `cur`, the selector on the new frontier, and the readback of
// (new frontier size, next slot) into pinned host memory.
void enqueue_level(spmspv_handle_t h, hipStream_t s, int slot, int cur, int* levels,
                   const SelectorArgs& sel) {
  BfsState* b = h->bfs;
  const int nxt = cur ^ 1;
  const int words = (h->n + 31) / 32;
  int* f = b->front[cur];
  int* nf = b->front[nxt];
  unsigned* nbits = b->fbits[nxt];
  hipLaunchKernelGGL(k_bfs_begin, dim3(std::min(1024, (words + 255) / 256)), dim3(256), 0, s, b->cnt,
                     nxt, nbits, words, h->ctl);
  switch (slot) {
    case SPMSPV_K_COL_NAIVE_ATOMIC:
      hipLaunchKernelGGL(k_td_naive, dim3(kBfsGrid), dim3(256), 0, s, h->csc_colptr, h->csc_row, f,
                         b->cnt, cur, lg2(h->team_col), b->visited, levels, nf, nbits);
      break;
    case SPMSPV_K_COL_LB_ATOMIC:
      hipLaunchKernelGGL((k_td_lb<kLbNT, kLbVT>), dim3(kBfsGrid), dim3(kLbNT), 0, s, h->ptr, h->xbeg,
                         h->split, h->csc_row, h->ctl, b->cnt, cur, b->visited, levels, nf, nbits);
      break;
    case SPMSPV_K_COL_NAIVE_SORT:
    case SPMSPV_K_COL_LB_SORT: {
      // boolean: the neighbours' row keys alone, sorted, then each distinct
      // row claimed
      int* bk = reinterpret_cast<int*>(h->sb.keyA);
      if (slot == SPMSPV_K_COL_NAIVE_SORT)
        launch_naive_col(s, 2, h->csc_colptr, h->csc_row, nullptr, f, nullptr, h->ptr, h->n, 0,
                         b->cnt + cur, h->team_col, nullptr, nullptr, bk, nullptr);
      else
        launch_lb_gather(s, 2, h->ptr, h->xbeg, nullptr, h->split, h->csc_row, nullptr, h->ctl, 0,
                         b->cnt + cur, nullptr, nullptr, bk, nullptr, -1);
      SortBufs kb = h->sb;
      kb.valA = kb.valB = nullptr;
      const int fb = launch_radix_sort(s, kb, h->m, &h->ctl->bin_len, h->ctl, -1, h->nnz, false, Ep{0, 1},
                                       Gate{}, h->small_sort_cap);
      hipLaunchKernelGGL(k_claim_sorted, dim3(kBfsGrid), dim3(256), 0, s, fb ? kb.keyB : kb.keyA,
                         &h->ctl->bin_len, b->cnt, cur, b->visited, levels, nf, nbits);
      break;
    }
    case SPMSPV_K_ROW_NAIVE:
      if (h->nnz > 0) {
        hipLaunchKernelGGL((k_bu_head<kBuBatch>), dim3(kBfsGrid), dim3(256), 0, s, h->csr_row, h->csr_col,
                           h->m, b->fbits[cur], b->cnt, cur, b->visited, levels, nf, nbits, b->queue);
        hipLaunchKernelGGL((k_bu_tail<kBuBatch, kBuTeam>), dim3(kBfsGrid), dim3(256), 0, s, h->csr_row,
                           h->csr_col, b->fbits[cur], b->cnt, cur, b->visited, levels, nf, nbits, b->queue);
      }
      break;
    default: {  // 5, 6, 7: boolean row kernels (hit bits only), then claim
      if (slot == SPMSPV_K_SPMV_NAIVE)
        launch_row_naive(s, true, true, h->csr_row, h->csr_col, nullptr, nullptr, b->fbits[cur],
                         h->m, h->team_row, nullptr, h->bitmap);
      else
        launch_row_merge(s, true, true, h->csr_row, h->csr_col, nullptr, nullptr, b->fbits[cur],
                         h->m, h->nnz, h->row_split, h->n_row_tiles, h->row_vt, nullptr, h->bitmap,
                         h->row_carry_val, h->row_carry_row, h->row_carry_hit);
      hipLaunchKernelGGL(k_claim_bits, dim3(std::min(kBfsGrid, (words + 255) / 256)), dim3(256), 0, s,
                         h->bitmap, h->m, b->cnt, cur, b->visited, levels, nf, nbits);
      break;
    }
  }
  launch_prep(s, h->csc_colptr, nf, h->n, 0, b->cnt + nxt, h->ptr, h->xbeg, h->split, h->scan_state,
              h->tile_mm, h->ctl, &sel, Ep{0, 0}, h->prep_samp);
  hipLaunchKernelGGL(k_bfs_signal, dim3(1), dim3(64), 0, s, b->cnt, nxt, h->ctl, h->sig_dev + kSigBfs);
}

}  // namespace

namespace spmspv {
// BFS buffers (visited and frontier bit arrays, frontier lists, the
// bottom-up queue): sized by m and n, so allocated once per handle.
int bfs_alloc(spmspv_handle_t h) {
  if (h->bfs) return SPMSPV_SUCCESS;
  auto* b = new BfsState();
  const size_t n = h->n, m = h->m, words = (std::max(n, m) + 31) / 32 + 2;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off = (off + bytes + 255) / 256 * 256;
    return o;
  };
  const size_t o_vis = take(words * 4), o_fb0 = take(words * 4), o_fb1 = take(words * 4);
  const size_t o_f0 = take(n * 4 + 4), o_f1 = take(n * 4 + 4), o_cnt = take(64), o_r = take(64);
  const size_t o_q = take(m * 4 + 4);
  if (hipMalloc(&b->alloc, off) != hipSuccess) {
    delete b;
    return set_error(h, SPMSPV_OUT_OF_MEMORY, "bfs buffers");
  }
  char* p = static_cast<char*>(b->alloc);
  b->visited = reinterpret_cast<unsigned*>(p + o_vis);
  b->fbits[0] = reinterpret_cast<unsigned*>(p + o_fb0);
  b->fbits[1] = reinterpret_cast<unsigned*>(p + o_fb1);
  b->front[0] = reinterpret_cast<int*>(p + o_f0);
  b->front[1] = reinterpret_cast<int*>(p + o_f1);
  b->cnt = reinterpret_cast<int*>(p + o_cnt);
  b->reached = reinterpret_cast<unsigned long long*>(p + o_r);
  b->queue = reinterpret_cast<int*>(p + o_q);
  if (!h->stream) {
    if (hipStreamCreateWithFlags(&b->own_stream, hipStreamNonBlocking) != hipSuccess) {
      (void)hipFree(b->alloc);
      delete b;
      return set_error(h, SPMSPV_DEVICE_ERROR, "bfs stream");
    }
  }
  h->bfs = b;
  return SPMSPV_SUCCESS;
}

// The captured level graphs hold the matrix, arena and row-tiling pointers
// and sizes they were captured with; every change of those drops them.
void bfs_drop_graphs(spmspv_handle_t h) {
  if (!h || !h->bfs) return;
  for (auto& g : h->bfs->graph)
    for (auto& e : g)
      if (e) {
        (void)hipGraphExecDestroy(e);
        e = nullptr;
      }
  h->bfs->levels_bound = nullptr;
}
}  // namespace spmspv

extern "C" {

int spmspv_bfs_release(spmspv_handle_t h) {
  if (!h || !h->bfs) return SPMSPV_SUCCESS;
  for (auto& g : h->bfs->graph)
    for (auto& e : g)
      if (e) (void)hipGraphExecDestroy(e);
  if (h->bfs->own_stream) (void)hipStreamDestroy(h->bfs->own_stream);
  if (h->bfs->alloc) (void)hipFree(h->bfs->alloc);
  delete h->bfs;
  h->bfs = nullptr;
  return SPMSPV_SUCCESS;
}

// The level loop; with max_levels > 0 it also records, per level (the
// frontier of level L is the one reached at level L-1), the device time of
// the level and the selector features of its frontier.
static int bfs_run(spmspv_handle_t h, int src, int* d_levels, int policy, int use_graphs,
                   double stats[4], int max_levels, float* level_ms, float* level_feat) {
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  if (src < 0 || src >= h->n || !d_levels || policy < -1 || policy >= SPMSPV_NUM_SLOTS)
    return set_error(h, SPMSPV_INVALID_ARGUMENT, "bfs: bad source, levels or policy");
  if (h->m != h->n) return set_error(h, SPMSPV_INVALID_ARGUMENT, "bfs needs a square matrix");
  if (!h->csr_row) return set_error(h, SPMSPV_MAT_FORMAT_ERR, "bfs needs CSR input");
  int rc;
  if ((rc = ensure_arena(h)) || (rc = need_csc(h))) return rc;
  if ((rc = ensure_row_tiles(h)) || (rc = ensure_mat_features(h)) || (rc = bfs_alloc(h))) return rc;
  BfsState* b = h->bfs;
  hipStream_t s = h->stream ? h->stream : b->own_stream;
  const int smode = policy >= 0 ? 2 : h->sel_mode, sfixed = policy >= 0 ? policy : h->sel_fixed;
  // the BFS levels' own weights when set (trained on BFS levels), else the
  // SpMSpV selector's
  const float* wsel = h->have_bfs_weights ? h->bfs_weights_dev : (h->have_weights ? h->weights_dev : nullptr);
  if (use_graphs && (b->levels_bound != d_levels || b->bound_mode != smode || b->bound_fixed != sfixed ||
                     b->bound_weights != wsel)) {
    bfs_drop_graphs(h);
    b->levels_bound = d_levels;
    b->bound_mode = smode;
    b->bound_fixed = sfixed;
    b->bound_weights = wsel;
  }
  // level L's kernels are enqueued on s, which the handle's own launches
  // (prep, sort) must share
  hipStream_t saved = h->stream;
  h->stream = s;
  SelectorArgs sel{wsel, h->mat, smode, sfixed};
  const auto t0 = std::chrono::steady_clock::now();
  hipLaunchKernelGGL(k_bfs_init, dim3(std::min(4096, (h->n + 255) / 256)), dim3(256), 0, s, d_levels,
                     h->n, b->visited, b->fbits[0], b->fbits[1], b->front[0], b->cnt, b->reached, src);
  launch_prep(s, h->csc_colptr, b->front[0], h->n, 0, b->cnt, h->ptr, h->xbeg, h->split,
              h->scan_state, h->tile_mm, h->ctl, &sel, next_ep(h, 1), h->prep_samp);
  // the level words: frontier size, choice and the level counter as tag
  long long* sw = h->sig + kSigBfs;
  __atomic_store_n(sw, kSigPending, __ATOMIC_RELEASE);
  hipLaunchKernelGGL(k_bfs_signal, dim3(1), dim3(64), 0, s, b->cnt, 0, h->ctl, h->sig_dev + kSigBfs);
  hipError_t e = hipGetLastError();
  const bool trace = max_levels > 0 && level_ms && level_feat;
  // waits for the word of level `tag` (polling; the trace's events need the
  // stream drained instead)
  long long word = 0;
  bool out_of_sequence = false;
  auto level_word = [&](int tag) -> hipError_t {
    if (trace) {
      hipError_t r = hipStreamSynchronize(s);
      if (r != hipSuccess) return r;
      word = __atomic_load_n(sw, __ATOMIC_ACQUIRE);
    } else if (wait_signal(h, kSigBfs, &word) != SPMSPV_SUCCESS) {
      return hipErrorUnknown;
    }
    if ((word >> 40) != tag) out_of_sequence = true;
    return out_of_sequence ? hipErrorUnknown : hipSuccess;
  };
  if (e == hipSuccess) e = level_word(0);
  int cur = 0, nlev = 1;
  int slot_hist[8] = {0};
  while (e == hipSuccess && (int)(word & 0xffffffff) > 0) {
    int slot = policy >= 0 ? policy : (int)((word >> 32) & 0xff);
    if (slot < 0 || slot >= SPMSPV_NUM_SLOTS) slot = SPMSPV_K_COL_LB_ATOMIC;
    slot_hist[slot]++;
    const int lev = nlev - 1;  // index of the level about to run
    if (trace && lev < max_levels) {
      std::memcpy(level_feat + 16 * lev, sw + 1, sizeof(float) * 16);
      (void)hipEventRecord(h->ev0, s);
    }
    __atomic_store_n(sw, kSigPending, __ATOMIC_RELEASE);
    if (use_graphs) {
      hipGraphExec_t& ge = b->graph[slot][cur];
      if (!ge) {
        hipGraph_t g = nullptr;
        e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) break;
        enqueue_level(h, s, slot, cur, d_levels, sel);
        e = hipStreamEndCapture(s, &g);
        if (e != hipSuccess) break;
        e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
          ge = nullptr;
          break;
        }
      }
      e = hipGraphLaunch(ge, s);
    } else {
      enqueue_level(h, s, slot, cur, d_levels, sel);
      e = hipGetLastError();
    }
    if (trace && lev < max_levels && e == hipSuccess) e = hipEventRecord(h->ev1, s);
    if (e == hipSuccess) e = level_word(nlev);
    if (trace && lev < max_levels && e == hipSuccess) {
      float ms = 0.f;
      e = hipEventElapsedTime(&ms, h->ev0, h->ev1);
      level_ms[lev] = ms;
    }
    cur ^= 1;
    ++nlev;
    if (nlev > h->n + 1) break;  // cannot happen on a correct run; bounds the loop
  }
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  // device epochs advance 16 per level; clear the look-back states long
  // before (ctl->epoch + offset) could wrap onto a stale word
  h->dev_levels += nlev;
  if (h->dev_levels > 200000) {
    launch_fill_i32(s, reinterpret_cast<int*>(h->sb.dstate), (long long)h->sort_words * 2, 0);
    launch_fill_i32(s, reinterpret_cast<int*>(h->sb.red_state), (long long)h->red_words * 2, 0);
    launch_fill_i32(s, reinterpret_cast<int*>(h->scan_state), (long long)h->scan_words * 2, 0);
    h->dev_levels = 0;
  }
  h->stream = saved;
  if (out_of_sequence) return set_error(h, SPMSPV_DEVICE_ERROR, "bfs: a level's word arrived out of sequence");
  if (e != hipSuccess) return set_error(h, SPMSPV_DEVICE_ERROR, std::string("bfs: ") + hipGetErrorString(e));
  hipLaunchKernelGGL(k_reached, dim3(std::min(1024, (h->n + 255) / 256)), dim3(256), 0, s, d_levels,
                     h->csc_colptr, h->n, b->reached);
  unsigned long long red[2] = {0, 0};
  e = hipMemcpyAsync(red, b->reached, sizeof(red), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return set_error(h, SPMSPV_DEVICE_ERROR, std::string("bfs: ") + hipGetErrorString(e));
  if (stats) {
    stats[0] = nlev - 1;             // levels with a nonempty frontier
    stats[1] = (double)red[1] / 2.0; // undirected edges in the reached component
    stats[2] = secs;
    stats[3] = (double)red[0];
  }
  return SPMSPV_SUCCESS;
}

int spmspv_bfs(spmspv_handle_t h, int src, int* d_levels, int policy, int use_graphs,
               double stats[4]) {
  return bfs_run(h, src, d_levels, policy, use_graphs, stats, 0, nullptr, nullptr);
}

int spmspv_bfs_trace(spmspv_handle_t h, int src, int* d_levels, int policy, int use_graphs,
                     double stats[4], int max_levels, float* level_ms, float* level_feat) {
  if (max_levels < 1 || !level_ms || !level_feat)
    return set_error(h, SPMSPV_INVALID_ARGUMENT, "bfs trace: need max_levels >= 1 and both arrays");
  return bfs_run(h, src, d_levels, policy, use_graphs, stats, max_levels, level_ms, level_feat);
}

}  // extern "C"
