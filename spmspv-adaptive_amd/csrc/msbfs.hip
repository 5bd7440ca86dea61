// Multi-source BFS: up to 64 level-synchronous BFSs advanced together, one
// bit per source in a 64-bit word per vertex (visited set and frontier), so
// one pass over the edges of a level serves every source. The reference has
// no BFS; its harnesses replay one frontier per SpMSpV call
// (la/test/test_application.cu:240-415). Semiring: boolean OR-AND on the
// 64-bit words, integer arithmetic only.
//
// Per level, from the current frontier (dense words front[cur], list F):
//   push (small frontiers): a team of lanes per frontier vertex u ORs
//     front[cur][u] & ~visit[v] into front[nxt][v] for every v in column u
//     (64-bit atomics); the first writer of a word lists v; a commit pass
//     over the listed v folds the new bits into visit[v], records the level
//     of each new (vertex, source) pair and lists v in the next frontier;
//   pull (large frontiers): every row v not yet reached by all sources ORs
//     front[cur][u] over its neighbours u (CSR row v), stopping once every
//     source has reached v; a lane per row tests its first 8 neighbours,
//     rows left open are queued for 16-lane teams;
// then the old frontier's words are cleared. The host picks the direction
// from the frontier's edge count (pull above nnz / 14) and stops when the
// frontier is empty.
//
// Levels are written vertex-major: levels[v * nsrc + i] is the level of v
// in the BFS from srcs[i] (-1 unreached).
#include <algorithm>
#include <chrono>
#include <vector>

#include "handle.hpp"

using namespace spmspv;

struct MsState {
  void* alloc = nullptr;
  unsigned long long* visit = nullptr;     // n
  unsigned long long* front[2] = {nullptr, nullptr};  // n each
  int* list[2] = {nullptr, nullptr};       // frontier vertex lists, n each
  int* touched = nullptr;                  // n: push targets / pull queue
  int* srcs = nullptr;                     // 64: the sources of a call
  unsigned long long* cnt = nullptr;       // [0],[1] list sizes, [2] touched, [3] frontier degree sum,
                                           // [4] reached pairs, [5] edges of reached pairs
};

namespace {

constexpr int kMsGrid = 2048;
constexpr int kMsHead = 8;    // neighbours a lane tests in the pull head pass
constexpr int kMsTeam = 16;   // lanes per vertex in the push and pull tail passes

__device__ __forceinline__ unsigned long long wave_append_ull(unsigned long long* counter, bool pred,
                                                             unsigned long long add, unsigned long long* addsum,
                                                             int* pos_out) {
  *pos_out = -1;
  const uint64_t b = __ballot(pred);
  if (!b) return 0;
  unsigned long long sum = pred ? add : 0ull;
  sum = wave_sum(sum);
  const int leader = __builtin_ctzll(b);
  unsigned long long base = 0;
  if (lane() == leader) {
    base = atomicAdd(counter, (unsigned long long)__popcll(b));
    if (addsum) atomicAdd(addsum, sum);
  }
  base = __shfl(base, leader, kWave);
  const uint64_t lt = lane() == 0 ? 0ull : (~0ull >> (kWave - lane()));
  *pos_out = pred ? (int)(base + __popcll(b & lt)) : -1;
  return b;
}

// Records the new (vertex, source) pairs of v and lists v in the next
// frontier (with its degree for the direction choice). Convergent.
__device__ __forceinline__ void discover(int v, unsigned long long nw, int nsrc, int level, int* levels,
                                         int* nlist, unsigned long long* ncount, unsigned long long* degsum,
                                         int deg) {
  int pos;
  wave_append_ull(ncount, nw != 0ull, (unsigned long long)deg, degsum, &pos);
  if (nw) {
    nlist[pos] = v;
    unsigned long long b = nw;
    int* lv = levels + (long long)v * nsrc;
    while (b) {
      const int i = __builtin_ctzll(b);
      b &= b - 1;
      lv[i] = level;
    }
  }
}

__global__ void k_ms_init(int n, int nsrc, const int* __restrict__ srcs, unsigned long long* visit,
                          unsigned long long* f0, unsigned long long* f1, int* levels) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
    visit[v] = 0ull;
    f0[v] = 0ull;
    f1[v] = 0ull;
  }
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < (long long)n * nsrc; k += stride)
    levels[k] = -1;
}

// Sources: their bits, level 0 and their degrees summed into *degsum for
// the first level's direction (after k_ms_init, same stream).
__global__ void k_ms_sources(int nsrc, const int* __restrict__ srcs, const int* __restrict__ colptr,
                             unsigned long long* visit, unsigned long long* f0, int* levels,
                             unsigned long long* degsum) {
  const int i = threadIdx.x;
  if (i >= nsrc) return;
  const int s = srcs[i];
  atomicOr(visit + s, 1ull << i);
  atomicOr(f0 + s, 1ull << i);
  levels[(long long)s * nsrc + i] = 0;
  atomicAdd(degsum, (unsigned long long)(colptr[s + 1] - colptr[s]));
}

__global__ __launch_bounds__(256) void k_ms_push(const int* __restrict__ colptr, const int* __restrict__ crow,
                                                 const int* __restrict__ flist, const unsigned long long* cnt,
                                                 int cur, const unsigned long long* __restrict__ front,
                                                 const unsigned long long* __restrict__ visit,
                                                 unsigned long long* nfront, int* touched,
                                                 unsigned long long* tcount) {
  const long long nf = (long long)cnt[cur];
  const int tl = threadIdx.x & (kMsTeam - 1);
  const long long team = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / kMsTeam;
  const long long nteams = (long long)gridDim.x * blockDim.x / kMsTeam;
  // whole waves iterate together (wave_append is convergent)
  const long long wave_teams = kWave / kMsTeam;
  const long long wbase = team - (team % wave_teams);
  for (long long t0 = wbase; t0 < nf; t0 += nteams) {
    const long long t = t0 + (team - wbase);
    const bool tin = t < nf;
    int u = 0, b = 0, e = 0;
    unsigned long long fu = 0ull;
    if (tin) {
      u = flist[t];
      fu = front[u];
      b = colptr[u];
      e = colptr[u + 1];
    }
    // the longest column of the wave's teams bounds the wave's trips
    int len = e - b;
#pragma unroll
    for (int o = kMsTeam; o < kWave; o <<= 1) len = max(len, __shfl_xor(len, o));
    for (int k0 = 0; k0 < len; k0 += kMsTeam) {
      const int k = b + k0 + tl;
      const bool go = tin && k < e;
      int v = 0;
      unsigned long long add = 0ull, old = 0ull;
      if (go) {
        v = crow[k];
        add = fu & ~visit[v];
        if (add) old = atomicOr(nfront + v, add);
      }
      int pos;
      wave_append_ull(tcount, go && add != 0ull && old == 0ull, 0ull, nullptr, &pos);
      if (pos >= 0) touched[pos] = v;
    }
  }
}

__global__ __launch_bounds__(256) void k_ms_commit(const int* __restrict__ touched, const unsigned long long* cnt,
                                                   const int* __restrict__ colptr, int nsrc, int level,
                                                   unsigned long long* visit,
                                                   const unsigned long long* __restrict__ nfront, int* levels,
                                                   int* nlist, unsigned long long* ncount,
                                                   unsigned long long* degsum) {
  const long long nt = (long long)cnt[2];
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long b0 = (long long)blockIdx.x * blockDim.x; b0 < nt; b0 += stride) {
    const long long k = b0 + threadIdx.x;
    int v = 0, deg = 0;
    unsigned long long nw = 0ull;
    if (k < nt) {
      v = touched[k];
      nw = nfront[v];  // every bit was masked with ~visit at push time
      visit[v] |= nw;
      deg = colptr[v + 1] - colptr[v];
    }
    discover(v, nw, nsrc, level, levels, nlist, ncount, degsum, deg);
  }
}

__global__ __launch_bounds__(256) void k_ms_pull_head(const int* __restrict__ row_ptr, const int* __restrict__ col,
                                                      int m, int nsrc, int level, unsigned long long full,
                                                      const unsigned long long* __restrict__ front,
                                                      unsigned long long* visit, unsigned long long* nfront,
                                                      int* levels, int* nlist, unsigned long long* ncount,
                                                      unsigned long long* degsum, int* queue,
                                                      unsigned long long* qcount) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long base = (long long)blockIdx.x * blockDim.x; base < m; base += stride) {
    const long long r = base + threadIdx.x;
    bool open = r < m;
    unsigned long long vis = 0ull;
    if (open) {
      vis = visit[r];
      open = vis != full;
    }
    int b = 0, e = 0;
    if (open) {
      b = row_ptr[r];
      e = row_ptr[r + 1];
    }
    const bool go = open && b < e;
    int u[kMsHead];
#pragma unroll
    for (int j = 0; j < kMsHead; ++j) u[j] = go ? col[min(b + j, e - 1)] : 0;
    unsigned long long acc = 0ull;
#pragma unroll
    for (int j = 0; j < kMsHead; ++j)
      if (go && b + j < e) acc |= front[u[j]];
    const bool more = go && e - b > kMsHead && (acc | vis) != full;
    unsigned long long nw = 0ull;
    if (go && !more) {
      nw = acc & ~vis;
      if (nw) {
        visit[r] = vis | nw;
        nfront[r] = nw;
      }
    }
    if (more) nfront[r] = acc;  // partial, finished by the tail pass
    discover((int)r, nw, nsrc, level, levels, nlist, ncount, degsum, e - b);
    int pos;
    wave_append_ull(qcount, more, 0ull, nullptr, &pos);
    if (more) queue[pos] = (int)r;
  }
}

__global__ __launch_bounds__(256) void k_ms_pull_tail(const int* __restrict__ row_ptr, const int* __restrict__ col,
                                                      int nsrc, int level, unsigned long long full,
                                                      const unsigned long long* __restrict__ front,
                                                      unsigned long long* visit, unsigned long long* nfront,
                                                      int* levels, int* nlist, unsigned long long* ncount,
                                                      unsigned long long* degsum, const int* __restrict__ queue,
                                                      const unsigned long long* qcount) {
  const long long nq = (long long)*qcount;
  const int tl = threadIdx.x & (kMsTeam - 1);
  const long long team = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / kMsTeam;
  const long long nteams = (long long)gridDim.x * blockDim.x / kMsTeam;
  const long long wave_teams = kWave / kMsTeam;
  const long long wbase = team - (team % wave_teams);
  for (long long t0 = wbase; t0 < nq; t0 += nteams) {
    const long long t = t0 + (team - wbase);
    const bool tin = t < nq;
    int r = 0, b = 0, e = 0;
    unsigned long long vis = full, acc = 0ull;
    if (tin) {
      r = queue[t];
      b = row_ptr[r] + kMsHead;
      e = row_ptr[r + 1];
      vis = visit[r];
      acc = nfront[r];
    }
    int len = e - b;
#pragma unroll
    for (int o = kMsTeam; o < kWave; o <<= 1) len = max(len, __shfl_xor(len, o));
    for (int k0 = 0; k0 < len; k0 += kMsTeam) {
      const int k = b + k0 + tl;
      if (tin && k < e && (acc | vis) != full) acc |= front[col[k]];
      // the team's OR, so that every lane sees when all sources reached r
#pragma unroll
      for (int o = 1; o < kMsTeam; o <<= 1) acc |= __shfl_xor(acc, o);
    }
    const bool lead = tin && tl == 0;
    unsigned long long nw = 0ull;
    if (lead) {
      nw = acc & ~vis;
      nfront[r] = nw;  // the partial word becomes the new bits (or 0)
      if (nw) visit[r] = vis | nw;
    }
    discover(r, nw, nsrc, level, levels, nlist, ncount, degsum, lead ? row_ptr[r + 1] - row_ptr[r] : 0);
  }
}

__global__ void k_ms_clear(const int* __restrict__ flist, const unsigned long long* cnt, int cur,
                           unsigned long long* front) {
  const long long nf = (long long)cnt[cur];
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < nf; k += (long long)gridDim.x * blockDim.x)
    front[flist[k]] = 0ull;
}

// Before a level: the next list's size, the touched / queue count and the
// frontier degree sum start at 0.
__global__ void k_ms_begin(unsigned long long* cnt, int nxt) {
  cnt[nxt] = 0ull;
  cnt[2] = 0ull;
  cnt[3] = 0ull;
}

// Reached (vertex, source) pairs and the edges they span: sum over v of
// popcount(visit[v]) and popcount(visit[v]) * deg(v).
__global__ void k_ms_reached(int n, const unsigned long long* __restrict__ visit, const int* __restrict__ colptr,
                             unsigned long long* out) {
  unsigned long long c = 0, d = 0;
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (long long)gridDim.x * blockDim.x) {
    const unsigned long long p = (unsigned long long)__popcll(visit[v]);
    c += p;
    d += p * (unsigned long long)(colptr[v + 1] - colptr[v]);
  }
  c = wave_sum(c);
  d = wave_sum(d);
  if (lane() == 0) {
    atomicAdd(out, c);
    atomicAdd(out + 1, d);
  }
}

}  // namespace

namespace spmspv {
int ms_alloc(spmspv_handle_t h) {
  if (h->ms) return SPMSPV_SUCCESS;
  auto* st = new MsState();
  const size_t n = h->n;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off = (off + bytes + 255) / 256 * 256;
    return o;
  };
  const size_t ov = take(n * 8), of0 = take(n * 8), of1 = take(n * 8), ol0 = take(n * 4 + 4),
               ol1 = take(n * 4 + 4), ot = take(n * 4 + 4), oc = take(64 * 8), os = take(64 * 4);
  if (hipMalloc(&st->alloc, off) != hipSuccess) {
    delete st;
    return set_error(h, SPMSPV_OUT_OF_MEMORY, "multi-source BFS buffers");
  }
  char* p = static_cast<char*>(st->alloc);
  st->visit = reinterpret_cast<unsigned long long*>(p + ov);
  st->front[0] = reinterpret_cast<unsigned long long*>(p + of0);
  st->front[1] = reinterpret_cast<unsigned long long*>(p + of1);
  st->list[0] = reinterpret_cast<int*>(p + ol0);
  st->list[1] = reinterpret_cast<int*>(p + ol1);
  st->touched = reinterpret_cast<int*>(p + ot);
  st->cnt = reinterpret_cast<unsigned long long*>(p + oc);
  st->srcs = reinterpret_cast<int*>(p + os);
  h->ms = st;
  return SPMSPV_SUCCESS;
}

void ms_release(spmspv_handle_t h) {
  if (!h || !h->ms) return;
  if (h->ms->alloc) (void)hipFree(h->ms->alloc);
  delete h->ms;
  h->ms = nullptr;
}
}  // namespace spmspv

extern "C" {

int spmspv_bfs_multi(spmspv_handle_t h, int nsrc, const int* srcs, int* d_levels, double stats[4]) {
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  if (nsrc < 1 || nsrc > 64 || !srcs || !d_levels)
    return set_error(h, SPMSPV_INVALID_ARGUMENT, "bfs_multi: 1..64 sources and a levels array are needed");
  if (h->m != h->n) return set_error(h, SPMSPV_INVALID_ARGUMENT, "bfs_multi needs a square matrix");
  if (!h->csr_row) return set_error(h, SPMSPV_MAT_FORMAT_ERR, "bfs_multi needs CSR input");
  for (int i = 0; i < nsrc; ++i)
    if (srcs[i] < 0 || srcs[i] >= h->n) return set_error(h, SPMSPV_INVALID_ARGUMENT, "bfs_multi: source out of range");
  int rc;
  if ((rc = ensure_arena(h)) || (rc = need_csc(h)) || (rc = ms_alloc(h))) return rc;
  MsState* st = h->ms;
  hipStream_t s = h->stream;
  const int n = h->n;
  const unsigned long long full = nsrc == 64 ? ~0ull : ((1ull << nsrc) - 1ull);
  // the frontier of level 0: the distinct sources, listed
  std::vector<int> uniq(srcs, srcs + nsrc);
  std::sort(uniq.begin(), uniq.end());
  uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
  unsigned long long* hc = reinterpret_cast<unsigned long long*>(h->h_pin);  // 8 words of the pinned area
  const auto t0 = std::chrono::steady_clock::now();
  hc[0] = (unsigned long long)uniq.size();
  hc[1] = hc[2] = hc[3] = 0ull;
  HIP_TRY(h, hipMemcpyAsync(st->cnt, hc, sizeof(unsigned long long) * 4, hipMemcpyHostToDevice, s));
  HIP_TRY(h, hipMemcpyAsync(st->srcs, srcs, sizeof(int) * nsrc, hipMemcpyHostToDevice, s));
  HIP_TRY(h, hipMemcpyAsync(st->list[0], uniq.data(), sizeof(int) * uniq.size(), hipMemcpyHostToDevice, s));
  const int grid = (int)std::max<long long>(1, std::min<long long>(kMsGrid, ((long long)n + 255) / 256));
  hipLaunchKernelGGL(k_ms_init, dim3(grid), dim3(256), 0, s, n, nsrc, st->srcs, st->visit, st->front[0],
                     st->front[1], d_levels);
  hipLaunchKernelGGL(k_ms_sources, dim3(1), dim3(64), 0, s, nsrc, st->srcs, h->csc_colptr, st->visit,
                     st->front[0], d_levels, st->cnt + 3);
  HIP_TRY(h, hipMemcpyAsync(hc, st->cnt, sizeof(unsigned long long) * 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(h, hipStreamSynchronize(s));
  // the frontier's edge count decides the direction (pull above nnz / 14)
  unsigned long long fdeg = hc[3];
  int cur = 0, level = 1, levels_run = 0;
  unsigned long long fcount = uniq.size();
  hipError_t e = hipSuccess;
  while (fcount > 0 && level <= n) {
    const int nxt = cur ^ 1;
    hipLaunchKernelGGL(k_ms_begin, dim3(1), dim3(1), 0, s, st->cnt, nxt);
    const bool pull = fdeg * 14ull > (unsigned long long)h->nnz;
    if (pull) {
      hipLaunchKernelGGL(k_ms_pull_head, dim3(grid), dim3(256), 0, s, h->csr_row, h->csr_col, n, nsrc, level, full,
                         st->front[cur], st->visit, st->front[nxt], d_levels, st->list[nxt], st->cnt + nxt,
                         st->cnt + 3, st->touched, st->cnt + 2);
      hipLaunchKernelGGL(k_ms_pull_tail, dim3(kMsGrid), dim3(256), 0, s, h->csr_row, h->csr_col, nsrc, level, full,
                         st->front[cur], st->visit, st->front[nxt], d_levels, st->list[nxt], st->cnt + nxt,
                         st->cnt + 3, st->touched, st->cnt + 2);
    } else {
      hipLaunchKernelGGL(k_ms_push, dim3(kMsGrid), dim3(256), 0, s, h->csc_colptr, h->csc_row, st->list[cur],
                         st->cnt, cur, st->front[cur], st->visit, st->front[nxt], st->touched, st->cnt + 2);
      hipLaunchKernelGGL(k_ms_commit, dim3(kMsGrid), dim3(256), 0, s, st->touched, st->cnt, h->csc_colptr, nsrc,
                         level, st->visit, st->front[nxt], d_levels, st->list[nxt], st->cnt + nxt, st->cnt + 3);
    }
    hipLaunchKernelGGL(k_ms_clear, dim3(kMsGrid), dim3(256), 0, s, st->list[cur], st->cnt, cur, st->front[cur]);
    e = hipMemcpyAsync(hc, st->cnt, sizeof(unsigned long long) * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) break;
    fcount = hc[nxt];
    fdeg = hc[3];
    cur = nxt;
    ++level;
    ++levels_run;
  }
  if (e != hipSuccess) return set_error(h, SPMSPV_DEVICE_ERROR, std::string("bfs_multi: ") + hipGetErrorString(e));
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  HIP_TRY(h, hipMemsetAsync(st->cnt + 4, 0, 16, s));
  hipLaunchKernelGGL(k_ms_reached, dim3(grid), dim3(256), 0, s, n, st->visit, h->csc_colptr, st->cnt + 4);
  HIP_TRY(h, hipMemcpyAsync(hc, st->cnt + 4, 16, hipMemcpyDeviceToHost, s));
  HIP_TRY(h, hipStreamSynchronize(s));
  if (stats) {
    stats[0] = levels_run;              // levels with a nonempty frontier
    stats[1] = (double)hc[1] / 2.0;     // undirected edges summed over the sources' components
    stats[2] = secs;
    stats[3] = (double)hc[0];           // reached (vertex, source) pairs
  }
  return SPMSPV_SUCCESS;
}

}  // extern "C"
