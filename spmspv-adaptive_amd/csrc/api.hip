// C-ABI of the engine: lifecycle, matrix and x input, conversions,
// features and the eight kernel slots. Every entry point cites the
// reference method it replaces in include/spmspv_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include <chrono>

#include "handle.hpp"
#include "selector.hpp"

using namespace spmspv;

namespace {
thread_local std::string g_err;

// Entry i of a k-key sub-vector of x: a pseudo-random position inside the
// i-th of k equal intervals of x, so the keys stay in order and the
// sub-vector's columns look like those of a random x of that size (evenly
// spaced positions of a dense x would be the columns whose low bits are
// zero, on a power-law matrix its heaviest ones).
__global__ void k_spread_pick(const int* __restrict__ xk, const float* __restrict__ xv, int xn, int k,
                              int* __restrict__ ok, float* __restrict__ ov) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k) return;
  const long long j0 = (long long)i * xn / k, j1 = (long long)(i + 1) * xn / k;
  unsigned hsh = (unsigned)i * 0x9E3779B1u;
  hsh ^= hsh >> 15;
  hsh *= 0x85EBCA77u;
  hsh ^= hsh >> 13;
  const long long j = j1 > j0 ? j0 + (long long)(hsh % (unsigned long long)(j1 - j0)) : j0;
  ok[i] = xk[j];
  ov[i] = xv[j];
}

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

struct Carver {
  char* base;
  size_t off = 0;
  template <typename T>
  T* take(size_t count) {
    off = align_up(off, 256);
    T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
    off += sizeof(T) * std::max<size_t>(count, 1);
    return p;
  }
};

int pow2_team(double avg) {
  int t = 1;
  while (t < avg && t < 64) t <<= 1;
  return t;
}
}  // namespace

namespace spmspv {

int set_error(spmspv_handle_t h, int code, const std::string& msg) {
  if (h) h->err = msg;
  g_err = msg;
  return code;
}


static int check_launch(spmspv_handle_t h, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return set_error(h, SPMSPV_DEVICE_ERROR, std::string(what) + ": " + hipGetErrorString(e));
  return SPMSPV_SUCCESS;
}

// Lay out (or measure, when base == nullptr) every buffer of the arena.
static size_t carve(spmspv_handle_t h, char* base) {
  Carver c{base};
  const size_t m = h->m, n = h->n, nnz = h->nnz;
  const size_t lb_tiles = (n + nnz + kLbTile - 1) / kLbTile + 2;
  const size_t prep_tiles = (n + kPrepTile - 1) / kPrepTile + 2;
  const size_t mwords = (m + 31) / 32;
  const size_t nwords = (n + 31) / 32;
  const size_t cmp_tiles = (std::max(mwords, nwords) + kCmpNT - 1) / kCmpNT + 2;
  const size_t sort_tiles = (nnz + kSortTile - 1) / kSortTile + 2;
  const size_t red_tiles = (nnz + kRedTile - 1) / kRedTile + 2;
  const size_t row_tiles = (m + nnz + kRowNT * kRowVTMin - 1) / (kRowNT * kRowVTMin) + 2;
  ;
  Ctl* ctl = c.take<Ctl>(1);
  MatFeat* mat = c.take<MatFeat>(1);
  double* mp = c.take<double>(256 * 5);
  int* ptr = c.take<int>(n + 1);
  int* xbeg = c.take<int>(n);
  int* split = c.take<int>(lb_tiles);
  int* tls = c.take<int>(lb_tiles);
  int* tmm = c.take<int>(2 * prep_tiles);
  const size_t bucket_tiles = m / 8192 + 258;
  unsigned long long* ss =
      c.take<unsigned long long>(std::max(std::max(prep_tiles, cmp_tiles), bucket_tiles));
  float* acc = c.take<float>(m);
  unsigned* bm = c.take<unsigned>(mwords + 2);
  unsigned char* mk = c.take<unsigned char>(mwords * 32 + 64);
  SortBufs sb;
  sb.keyA = c.take<unsigned>(nnz);
  sb.valA = c.take<float>(nnz);
  sb.keyB = c.take<unsigned>(nnz);
  sb.valB = c.take<float>(nnz);
  const size_t dscan_tiles = sort_tiles * 256 / kDScanTile + 2;
  sb.dcount = c.take<unsigned>(sort_tiles * 256);
  sb.dstate = c.take<unsigned long long>(dscan_tiles);
  sb.red_state = c.take<unsigned long long>(red_tiles);
  sb.carry_val = c.take<float>(red_tiles);
  sb.carry_idx = c.take<int>(red_tiles);
  int* ykt = c.take<int>(m);
  float* yvt = c.take<float>(m);
  float* xdi = c.take<float>(n);
  float* xen = c.take<float>(n);
  float* xen5a = c.take<float>(n);
  float* xen5b = c.take<float>(n);
  unsigned* bi = c.take<unsigned>(nwords + 2);
  int* xki = c.take<int>(n);
  float* xvi = c.take<float>(n);
  int* rsplit = c.take<int>(row_tiles + 1);
  float* rcv = c.take<float>(row_tiles);
  int* rcr = c.take<int>(row_tiles);
  int* rch = c.take<int>(row_tiles);
  long long* ndev = c.take<long long>(4);
  float* w = c.take<float>(SPMSPV_MLP_NWEIGHTS);
  float* wb = c.take<float>(SPMSPV_MLP_NWEIGHTS);
  long long* selp = c.take<long long>(5 * kSelMaxGrid);
  long long* psamp = c.take<long long>(2 * (kSampleKeys / kPrepNT));
  BinWs bws;
  bws.cursor = c.take<unsigned>((size_t)kBinMax * kBinCurStride);
  bws.base = c.take<int>(kBinMax + 1);
  bws.start = c.take<int>(kBinMax + 1);
  bws.tbl = c.take<int>(kBinTbl);
  if (base) {
    h->prep_samp = psamp;
    h->bin_ws = bws;
    h->bin_base_bits = -1;
    h->scan_words = std::max(std::max(prep_tiles, cmp_tiles), bucket_tiles);
    h->sort_words = dscan_tiles;
    h->red_words = red_tiles;
    h->ctl = ctl;
    h->sel_part = selp;
    h->mat = mat;
    h->mat_partials = mp;
    h->ptr = ptr;
    h->xbeg = xbeg;
    h->split = split;
    h->tl_starts = tls;
    h->tile_mm = tmm;
    h->scan_state = ss;
    h->acc = acc;
    h->bitmap = bm;
    h->marks = mk;
    h->sb = sb;
    h->yk_tmp = ykt;
    h->yv_tmp = yvt;
    h->x_dense_int = xdi;
    h->x_enc = xen;
    h->x_enc5b[0] = xen5a;
    h->x_enc5b[1] = xen5b;
    h->enc5_clean[0] = h->enc5_clean[1] = false;
    h->bits_int = bi;
    h->xk_int = xki;
    h->xv_int = xvi;
    h->row_split = rsplit;
    h->row_carry_val = rcv;
    h->row_carry_row = rcr;
    h->row_carry_hit = rch;
    h->n_dev = ndev;
    h->weights_dev = w;
    h->bfs_weights_dev = wb;
  }
  return c.off;
}

int ensure_arena(spmspv_handle_t h) {
  if (h->reserved) return SPMSPV_SUCCESS;
  const size_t bytes = carve(h, nullptr);
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess)
    return set_error(h, SPMSPV_OUT_OF_MEMORY,
                     "arena of " + std::to_string(bytes) + " bytes: " + hipGetErrorString(e));
  h->arena = p;
  h->arena_bytes = bytes;
  carve(h, static_cast<char*>(p));
  // Everything starts at zero: control block, look-back states, the
  // accumulator and bitmaps whose all-zero invariant the kernels keep.
  HIP_TRY(h, hipMemsetAsync(p, 0, bytes, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  h->reserved = true;
  return SPMSPV_SUCCESS;
}

int ensure_row_tiles(spmspv_handle_t h) {
  if (h->row_ready) return SPMSPV_SUCCESS;
  if (!h->csr_row) return set_error(h, SPMSPV_MAT_FORMAT_ERR, "row kernels need CSR input");
  int rc = ensure_arena(h);
  if (rc) return rc;
  const int tile = kRowNT * h->row_vt;
  h->n_row_tiles = (int)(((long long)h->m + h->nnz + tile - 1) / tile);
  launch_row_split(h->stream, h->csr_row, h->m, h->nnz, h->row_split, h->n_row_tiles, tile);
  rc = check_launch(h, "row split");
  if (rc) return rc;
  h->row_ready = true;
  return SPMSPV_SUCCESS;
}

int ensure_csr5(spmspv_handle_t h) {
  const int S = csr5_sigma_of(h->sigma, h->m, h->nnz);
  if (h->csr5_ready && h->csr5.S == S) return SPMSPV_SUCCESS;
  if (!h->csr_row) return set_error(h, SPMSPV_MAT_FORMAT_ERR, "CSR5 needs CSR input");
  int rc = ensure_arena(h);
  if (rc) return rc;
  const size_t bytes = csr5_bytes(h->m, h->n, h->nnz, S);
  if (h->csr5_bytes < bytes) {  // not reserved by alloc_pre_buffer
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    if (h->csr5_mem) HIP_TRY(h, hipFree(h->csr5_mem));
    h->csr5_mem = nullptr;
    h->csr5_bytes = 0;
    HIP_TRY(h, hipMalloc(&h->csr5_mem, bytes));
    h->csr5_bytes = bytes;
  }
  // columns ordered by nonzero count for the gathers' cache locality (the
  // sort buffers hold nnz items, so the order needs n <= nnz)
  const ColumnOrderArgs order{h->sb, (long long)h->nnz, h->ctl, h->n_dev + 3, next_ep(h, 12),
                              h->csc_stale ? nullptr : h->csc_colptr};
  HIP_TRY(h, csr5_build(h->stream, h->csr_row, h->csr_col, h->csr_val, h->m, h->n, h->nnz, S, h->csr5_mem,
                        h->h_pin, &order, &h->csr5));
  h->enc5_valid = false;
  h->launches += 12;
  h->csr5_ready = true;
  return SPMSPV_SUCCESS;
}

int ensure_mat_features(spmspv_handle_t h) {
  if (h->mat_ready) return SPMSPV_SUCCESS;
  int rc = ensure_arena(h);
  if (rc) return rc;
  // Row degrees come from CSR; with CSC-only input the column pointer is
  // used (identical for the symmetric matrices of the benchmarks).
  const int* rp = h->csr_row ? h->csr_row : h->csc_colptr;
  const int rows = h->csr_row ? h->m : h->n;
  const int cols = h->csr_row ? h->n : h->m;
  if (!rp) return set_error(h, SPMSPV_MAT_FORMAT_ERR, "no matrix input");
  // degree histogram needs max degree + 1 <= cols + 1 slots
  int* hist = cols <= h->n ? h->ptr : h->yk_tmp;
  launch_mat_features(h->stream, rp, rows, cols, h->nnz, hist, h->mat_partials, h->mat, h->ctl);
  rc = check_launch(h, "mat features");
  if (rc) return rc;
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  h->mat_ready = true;
  return SPMSPV_SUCCESS;
}

int need_csc(spmspv_handle_t h) {
  if (!h->csc_colptr || h->csc_stale) {
    if (!h->csr_row) return set_error(h, SPMSPV_MAT_FORMAT_ERR, "no matrix input");
    int rc = spmspv_to_csc(h);
    if (rc) return rc;
  }
  return SPMSPV_SUCCESS;
}

// Makes x available as sparse keys/values (from dense when needed).
static int need_sparse_x(spmspv_handle_t h) {
  if (h->sparse_valid) return SPMSPV_SUCCESS;
  if (h->vec_type == SPMSPV_VEC_DENSE || h->vec_type == SPMSPV_VEC_BITARRAY)
    return spmspv_dense2sparse(h);
  return set_error(h, SPMSPV_VEC_FORMAT_ERR, "x is not set");
}

static int need_dense_x(spmspv_handle_t h) {
  if (h->dense_valid) return SPMSPV_SUCCESS;
  if (!h->sparse_valid) return set_error(h, SPMSPV_VEC_FORMAT_ERR, "x is not set");
  return spmspv_sparse2dense(h);
}

static int need_bits(spmspv_handle_t h) {
  if (h->bits_valid) return SPMSPV_SUCCESS;
  if (h->sparse_valid) {
    unsigned* b = h->bits();
    launch_fill_i32(h->stream, reinterpret_cast<int*>(b), (h->n + 31) / 32, 0);
    launch_set_bits(h->stream, h->x_key, h->xnnz, nullptr, h->n, b);
    h->bits_valid = true;
    return check_launch(h, "set bits");
  }
  if (h->dense_valid) return spmspv_dense2bitarray(h);
  return set_error(h, SPMSPV_VEC_FORMAT_ERR, "x is not set");
}

// The x_enc5 buffer for scattering a new x: one known to be all +0.0 if
// there is one, otherwise the current one, filled first.
static float* enc5_for_scatter(spmspv_handle_t h) {
  int cur = h->enc5_cur;
  if (!h->enc5_clean[cur] && h->enc5_clean[1 - cur]) cur = 1 - cur;
  if (!h->enc5_clean[cur]) {
    launch_fill_f32(h->stream, h->x_enc5b[cur], h->n, 0.f);
    h->launches++;
  }
  h->enc5_cur = cur;
  h->enc5_clean[cur] = false;
  return h->x_enc5b[cur];
}

// Support encoding of x in the CSR5 copy's column order (x_enc itself when
// the copy keeps the columns' order).
static int need_enc_x(spmspv_handle_t h);
static int need_enc_x5(spmspv_handle_t h, const float** out) {
  if (!h->csr5.cperm) {
    *out = h->x_enc;
    return need_enc_x(h);
  }
  if (h->enc5_valid) {
    *out = h->x_enc5b[h->enc5_cur];
    return SPMSPV_SUCCESS;
  }
  if (h->sparse_valid) {
    float* e = enc5_for_scatter(h);
    launch_scatter(h->stream, h->x_key, h->x_val, h->xnnz, nullptr, h->n, e, true, Gate{},
                   h->csr5.cperm);
  } else {
    int rc;
    if ((rc = need_dense_x(h)) || (rc = need_bits(h))) return rc;
    // every entry is written: take a buffer that is not clean, if any
    int cur = h->enc5_cur;
    if (h->enc5_clean[cur] && !h->enc5_clean[1 - cur]) cur = 1 - cur;
    h->enc5_cur = cur;
    h->enc5_clean[cur] = false;
    launch_encode_x(h->stream, h->dense_x(), h->bits(), h->n, h->x_enc5b[cur], h->csr5.cperm);
  }
  *out = h->x_enc5b[h->enc5_cur];
  int rc = check_launch(h, "encode x (CSR5 order)");
  if (rc) return rc;
  h->enc5_valid = true;
  return SPMSPV_SUCCESS;
}

// Support encoding of x for the row-merge kernels: from the sparse form
// directly, else from the dense values and the bit array.
static int need_enc_x(spmspv_handle_t h) {
  if (h->enc_valid) return SPMSPV_SUCCESS;
  if (h->sparse_valid) {
    launch_fill_f32(h->stream, h->x_enc, h->n, 0.f);
    launch_scatter(h->stream, h->x_key, h->x_val, h->xnnz, nullptr, h->n, h->x_enc, true);
  } else {
    int rc;
    if ((rc = need_dense_x(h)) || (rc = need_bits(h))) return rc;
    launch_encode_x(h->stream, h->dense_x(), h->bits(), h->n, h->x_enc);
  }
  h->enc_valid = true;
  return check_launch(h, "encode x");
}

Ep next_ep(spmspv_handle_t h, unsigned k) {
  // host epochs live in [1, 2^23); before they wrap, every look-back state
  // array is cleared so that no stale word can carry a reused epoch
  if (h->epoch_host + k >= (1u << 23) - 1) {
    launch_fill_i32(h->stream, reinterpret_cast<int*>(h->sb.dstate), (long long)h->sort_words * 2, 0);
    launch_fill_i32(h->stream, reinterpret_cast<int*>(h->sb.red_state), (long long)h->red_words * 2, 0);
    launch_fill_i32(h->stream, reinterpret_cast<int*>(h->scan_state), (long long)h->scan_words * 2, 0);
    h->epoch_host = 1;
  }
  Ep e{h->epoch_host, -1};
  h->epoch_host += k;
  return e;
}

// The bucket slot's bins for the current matrix and bucket_bits, built on
// the host from the row pointer (copied back, or counted from the CSC's
// row indices for a CSC-only matrix) once per matrix and bin count.
int ensure_bin_layout(spmspv_handle_t h) {
  if (h->bin_base_bits == h->bucket_bits) return SPMSPV_SUCCESS;
  const int m = h->m;
  std::vector<int> rp(m + 1, 0);
  if (h->csr_row) {
    HIP_TRY(h, hipMemcpyAsync(rp.data(), h->csr_row, sizeof(int) * (m + 1), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
  } else {
    int rc;
    if ((rc = need_csc(h))) return rc;
    launch_fill_i32(h->stream, h->yk_tmp, m, 0);
    launch_row_counts(h->stream, h->csc_row, h->nnz, h->yk_tmp);
    HIP_TRY(h, hipMemcpyAsync(rp.data() + 1, h->yk_tmp, sizeof(int) * m, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    for (int i = 0; i < m; ++i) rp[i + 1] += rp[i];
  }
  BinLayout L;
  const int nb = bin_layout(rp, m, h->bucket_bits, &L);
  if (nb < 0)
    return set_error(h, SPMSPV_INVALID_ARGUMENT,
                     "bucket slot: more than " + std::to_string(kBinMax) + " row bins would be needed");
  HIP_TRY(h, hipMemcpyAsync(h->bin_ws.start, L.start.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice,
                            h->stream));
  HIP_TRY(h, hipMemcpyAsync(h->bin_ws.base, L.base.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice, h->stream));
  HIP_TRY(h, hipMemcpyAsync(h->bin_ws.tbl, L.tbl.data(), sizeof(int) * L.tbl.size(), hipMemcpyHostToDevice,
                            h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  h->bin_ws.max_rows = L.max_rows;
  h->bin_ws.ntbl = (int)L.tbl.size();
  h->bin_ws.tbl_shift = L.tbl_shift;
  h->bin_nbins = nb;
  h->bin_base_bits = h->bucket_bits;
  return SPMSPV_SUCCESS;
}

int ensure_binlen_host(spmspv_handle_t h) {
  if (h->binlen_host >= 0) return SPMSPV_SUCCESS;
  HIP_TRY(h, hipMemcpyAsync(h->h_pin, &h->ctl->bin_len, sizeof(long long), hipMemcpyDeviceToHost,
                            h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  long long bl;
  std::memcpy(&bl, h->h_pin, sizeof(bl));
  h->binlen_host = bl;
  return SPMSPV_SUCCESS;
}

int run_prep(spmspv_handle_t h, bool with_split, const SelectorArgs* sel) {
  launch_prep(h->stream, h->csc_colptr, h->x_key, h->n, h->xnnz, nullptr, h->ptr, h->xbeg,
              with_split ? h->split : nullptr, h->scan_state, h->tile_mm, h->ctl, sel, next_ep(h, 1),
              h->prep_samp);
  h->launches++;
  int rc = check_launch(h, "prep");
  if (rc) return rc;
  if (with_split) h->prep_valid = true;
  return SPMSPV_SUCCESS;
}

static long long lb_tiles_hint(spmspv_handle_t h) {
  // tiles of the merged (x entries + products) sequence; bin_len <= nnz
  const long long bl = h->binlen_host >= 0 ? h->binlen_host : -1;
  if (bl < 0) return -1;
  return (h->xnnz + bl + kLbTile - 1) / kLbTile;
}

int run_kernel(spmspv_handle_t h, int kid, float* y_dense, int* y_key, float* y_val,
               bool want_nnz) {
  int rc = ensure_arena(h);
  if (rc) return rc;
  h->main_timed = false;  // set again below when this slot's tile kernel is timed
  h->count_sent = false;
  hipStream_t s = h->stream;
  // signal word 3 for the result count, when the caller wants it: marked
  // pending once, before the first launch of the call that may write it
  bool sig3_marked = false;
  auto mark_sig3 = [&]() {
    if (!sig3_marked) __atomic_store_n(h->sig + 3, kSigPending, __ATOMIC_RELEASE);
    sig3_marked = true;
  };
  auto count_sig = [&]() -> long long* {
    if (!h->count_to_host) return nullptr;
    mark_sig3();
    h->count_sent = true;
    return h->sig_dev + 3;
  };
  const bool want_sparse = y_key || y_val || want_nnz;
  const int m = h->m;
  // A fixed column slot (not behind a speculative gate) on a small x with a
  // sparse result first tries the single-launch path. When the caller waits
  // for the count anyway, the host reads the path's word first: a finished
  // call returns at once, otherwise the slot's launches follow with bin_len
  // known. Without a count, the slot's launches are queued behind it, gated
  // on its not having finished the call.
  struct GateRestore {
    spmspv_handle_t h;
    bool on = false;
    ~GateRestore() {
      if (on) h->gate = Gate{};
    }
  } gate_restore{h};
  if (kid >= SPMSPV_K_COL_NAIVE_ATOMIC && kid <= SPMSPV_K_COL_LB_SORT && !h->gate.word && !y_dense &&
      h->small_x_path && h->fuse_cap > 0 && h->vec_type != SPMSPV_VEC_DENSE && h->sparse_valid &&
      h->xnnz <= kFuseKeys &&
      h->binlen_host <= h->fuse_cap) {  // (binlen_host -1: not known yet)
    if ((rc = need_csc(h))) return rc;
    if (h->count_to_host) mark_sig3();
    const bool wait = h->count_to_host;
    if (wait) __atomic_store_n(h->sig + 1, kSigPending, __ATOMIC_RELEASE);
    const SelectorArgs fixed{nullptr, nullptr, 2, kid, wait ? h->sig_dev + 1 : nullptr, 0};
    launch_col_fused(s, h->csc_colptr, h->csc_row, h->csc_val, h->x_key, h->x_val, h->n, h->xnnz, m, h->ptr,
                     h->xbeg, h->split, h->ctl, fixed, y_key ? y_key : h->yk_tmp, y_val ? y_val : h->yv_tmp,
                     h->count_to_host ? h->sig_dev + 3 : nullptr, h->fuse_cap);
    h->launches++;
    if ((rc = check_launch(h, "small-x column path"))) return rc;
    h->prep_valid = true;
    if (wait) {
      long long word = 0;
      if ((rc = wait_signal(h, 1, &word))) return rc;
      if (sel_word_done(word)) {
        h->binlen_host = -1;
        h->count_sent = true;  // the count is in signal word 3
        return SPMSPV_SUCCESS;
      }
      h->binlen_host = sel_word_binlen(word);
    } else {
      h->binlen_host = -1;
      h->gate = Gate{&h->ctl->kernel_choice, kid};
      gate_restore.on = true;
    }
  }
  switch (kid) {
    case SPMSPV_K_COL_NAIVE_ATOMIC:
    case SPMSPV_K_COL_LB_ATOMIC: {
      if ((rc = need_csc(h)) || (rc = need_sparse_x(h))) return rc;
      float* acc = y_dense ? y_dense : h->acc;
      if (y_dense) launch_fill_f32(s, y_dense, m, 0.f, h->gate);
      if (kid == SPMSPV_K_COL_NAIVE_ATOMIC) {
        launch_naive_col(s, 0, h->csc_colptr, h->csc_row, h->csc_val, h->x_key, h->x_val, nullptr,
                         h->n, h->xnnz, nullptr, h->team_col, acc, h->marks, nullptr, nullptr,
                         h->gate);
      } else {
        if (!h->prep_valid && (rc = run_prep(h, true, nullptr))) return rc;
        launch_lb_gather(s, 0, h->ptr, h->xbeg, h->x_val, h->split, h->csc_row, h->csc_val, h->ctl,
                         h->xnnz, nullptr, acc, h->marks, nullptr, nullptr, lb_tiles_hint(h), h->gate);
      }
      launch_compact(s, nullptr, h->marks, m, acc, y_dense ? 0 : 1, y_key, y_val, h->scan_state, h->ctl,
                     next_ep(h, 1), nullptr, h->gate, count_sig());
      h->launches += 2;
      return check_launch(h, "column atomic kernel");
    }
    case SPMSPV_K_COL_NAIVE_SORT:
    case SPMSPV_K_COL_LB_SORT: {
      if ((rc = need_csc(h)) || (rc = need_sparse_x(h))) return rc;
      const bool lb = kid == SPMSPV_K_COL_LB_SORT;
      if (!h->prep_valid && (rc = run_prep(h, true, nullptr))) return rc;
      int* bk = reinterpret_cast<int*>(h->sb.keyA);
      float* bv = h->sb.valA;
      if (lb)
        launch_lb_gather(s, 1, h->ptr, h->xbeg, h->x_val, h->split, h->csc_row, h->csc_val, h->ctl,
                         h->xnnz, nullptr, nullptr, nullptr, bk, bv, lb_tiles_hint(h), h->gate);
      else
        launch_naive_col(s, 1, h->csc_colptr, h->csc_row, h->csc_val, h->x_key, h->x_val, h->ptr,
                         h->n, h->xnnz, nullptr, h->team_col, nullptr, nullptr, bk, bv, h->gate);
      int* ok = y_key ? y_key : h->yk_tmp;
      float* ov = y_val ? y_val : h->yv_tmp;
      // queued before the selector's word arrives (speculation): every grid
      // is sized from the device-side count instead of the host's bin_len
      long long hint = -1;
      if (!h->gate.word) {
        if ((rc = ensure_binlen_host(h))) return rc;
        hint = h->binlen_host;
      }
      launch_sort_reduce(s, h->sb, m, h->ctl, ok, ov, hint, h->nnz, next_ep(h, 5), h->gate, h->small_sort_cap);
      h->launches += 6;
      if (y_dense) {
        launch_fill_f32(s, y_dense, m, 0.f, h->gate);
        launch_scatter(s, ok, ov, 0, &h->ctl->ynnz, m, y_dense, false, h->gate);
      }
      return check_launch(h, "column sort kernel");
    }
    case SPMSPV_K_COL_TWO_LAYER_ATOMIC:
    case SPMSPV_K_COL_TWO_LAYER_SORT: {
      if ((rc = need_csc(h)) || (rc = need_sparse_x(h))) return rc;
      if (!h->prep_valid && (rc = run_prep(h, true, nullptr))) return rc;
      const long long tiles = h->binlen_host >= 0 ? (h->binlen_host + kLbTile - 1) / kLbTile : -1;
      if (kid == SPMSPV_K_COL_TWO_LAYER_ATOMIC) {
        float* acc = y_dense ? y_dense : h->acc;
        if (y_dense) launch_fill_f32(s, y_dense, m, 0.f);
        launch_tl_gather(s, 0, h->ptr, h->xbeg, h->x_val, h->xnnz, h->tl_starts, h->csc_row, h->csc_val, h->ctl,
                         acc, h->marks, nullptr, nullptr, tiles);
        launch_compact(s, nullptr, h->marks, m, acc, y_dense ? 0 : 1, y_key, y_val, h->scan_state, h->ctl,
                       next_ep(h, 1), nullptr, Gate{}, count_sig());
        h->launches += 3;
        return check_launch(h, "two-layer atomic kernel");
      }
      launch_tl_gather(s, 1, h->ptr, h->xbeg, h->x_val, h->xnnz, h->tl_starts, h->csc_row, h->csc_val, h->ctl,
                       nullptr, nullptr, reinterpret_cast<int*>(h->sb.keyA), h->sb.valA, tiles);
      int* ok = y_key ? y_key : h->yk_tmp;
      float* ov = y_val ? y_val : h->yv_tmp;
      if ((rc = ensure_binlen_host(h))) return rc;
      launch_sort_reduce(s, h->sb, m, h->ctl, ok, ov, h->binlen_host, h->nnz, next_ep(h, 5), Gate{},
                         h->small_sort_cap);
      h->launches += 7;
      if (y_dense) {
        launch_fill_f32(s, y_dense, m, 0.f);
        launch_scatter(s, ok, ov, 0, &h->ctl->ynnz, m, y_dense);
      }
      return check_launch(h, "two-layer sort kernel");
    }
    case SPMSPV_K_COL_BUCKET: {
      if ((rc = need_csc(h)) || (rc = need_sparse_x(h)) || (rc = ensure_bin_layout(h))) return rc;
      if (!h->prep_valid && (rc = run_prep(h, true, nullptr))) return rc;
      int* ok = y_key ? y_key : h->yk_tmp;
      float* ov = y_val ? y_val : h->yv_tmp;
      // grids from the host's bin_len when known, else from the bound nnz
      launch_bucket(s, h->ptr, h->xbeg, h->x_val, h->split, h->csc_row, h->csc_val, h->ctl, h->xnnz, h->bin_ws,
                    h->bin_nbins, h->sb.keyB, h->sb.valB, h->acc, h->marks, lb_tiles_hint(h), h->binlen_host,
                    h->nnz, h->gate);
      launch_compact(s, nullptr, h->marks, m, h->acc, 1, ok, ov, h->scan_state, h->ctl, next_ep(h, 1), nullptr,
                     h->gate, y_dense ? nullptr : count_sig(), h->bin_ws.cursor, h->bin_nbins);
      h->launches += 3;
      if (y_dense) {
        launch_fill_f32(s, y_dense, m, 0.f, h->gate);
        launch_scatter(s, ok, ov, 0, &h->ctl->ynnz, m, y_dense, false, h->gate);
      }
      return check_launch(h, "bucket kernel");
    }
    case SPMSPV_K_ROW_NAIVE:
    case SPMSPV_K_ROW_LB:
    case SPMSPV_K_SPMV_NAIVE:
    case SPMSPV_K_SPMV_LB: {
      if (!h->csr_row) return set_error(h, SPMSPV_MAT_FORMAT_ERR, "row kernels need CSR input");
      const bool masked = kid == SPMSPV_K_ROW_NAIVE || kid == SPMSPV_K_ROW_LB;
      const bool lb = kid == SPMSPV_K_ROW_LB || kid == SPMSPV_K_SPMV_LB;
      float* yout = y_dense ? y_dense : h->acc;
      // writing into the internal accumulator always tracks hits so that
      // the compaction can restore its all-zero state
      const bool hits = want_sparse || !y_dense;
      if (lb && h->spmv_type == 1) {
        if ((rc = ensure_csr5(h))) return rc;
        const float* xin = nullptr;
        if (masked || hits || h->csr5.cperm) {
          // the support encoding in the copy's column order (also serves the
          // plain product: +0.0 outside the support)
          if ((rc = need_enc_x5(h, &xin))) return rc;
        } else {
          if ((rc = need_dense_x(h))) return rc;
          xin = h->dense_x();
        }
        float* vals = nullptr;
        const int* kmap = nullptr;
        int nk = 0;
        // the idle x_enc5 buffer is cleared by the tile kernel on the way
        const int idle = 1 - h->enc5_cur;
        float* clr = h->csr5.cperm && xin == h->x_enc5b[h->enc5_cur] && !h->enc5_clean[idle]
                         ? h->x_enc5b[idle]
                         : nullptr;
        launch_csr5(s, h->csr5, masked, hits, xin, nullptr, m, h->nnz, yout, yout != h->acc, h->marks,
                    &vals, &kmap, &nk, h->time_main ? h->evk0 : nullptr,
                    h->time_main ? h->evk1 : nullptr, h->gate, clr);
        h->main_timed = h->time_main;
        if (hits)
          launch_compact(s, nullptr, h->marks, nk, vals, vals == h->acc ? 1 : 0, y_key, y_val,
                         h->scan_state, h->ctl, next_ep(h, 1), kmap, h->gate, count_sig());
        h->launches += 3;
        if ((rc = check_launch(h, "CSR5 row kernel"))) return rc;
        if (clr) {
          // the buffer counts as clean once the launches went in; a gated
          // launch clears only if it runs, which the choice settles
          if (h->gate.word) h->enc5_clear_pending = idle;
          else h->enc5_clean[idle] = true;
        }
        return SPMSPV_SUCCESS;
      }
      if (lb) {
        const float* xin = nullptr;
        if (masked || hits) {
          if ((rc = need_enc_x(h))) return rc;
          xin = h->x_enc;
        } else {
          if ((rc = need_dense_x(h))) return rc;
          xin = h->dense_x();
        }
        if ((rc = ensure_row_tiles(h))) return rc;
        launch_row_merge(s, masked, hits, h->csr_row, h->csr_col, h->csr_val, xin, nullptr, m, h->nnz,
                         h->row_split, h->n_row_tiles, h->row_vt, yout, h->bitmap, h->row_carry_val,
                         h->row_carry_row, h->row_carry_hit, h->time_main ? h->evk0 : nullptr,
                         h->time_main ? h->evk1 : nullptr, h->gate);
        h->main_timed = h->time_main;
      } else {
        if ((rc = need_dense_x(h))) return rc;
        if ((masked || hits) && (rc = need_bits(h))) return rc;
        launch_row_naive(s, masked, hits, h->csr_row, h->csr_col, h->csr_val, h->dense_x(), h->bits(),
                         m, h->team_row, yout, h->bitmap);
      }
      if (hits)
        launch_compact(s, h->bitmap, nullptr, m, yout, yout == h->acc ? 1 : 0, y_key, y_val,
                       h->scan_state, h->ctl, next_ep(h, 1), nullptr, h->gate, count_sig());
      h->launches += 2;
      return check_launch(h, "row kernel");
    }
    default:
      return set_error(h, SPMSPV_INVALID_ARGUMENT, "kernel id " + std::to_string(kid));
  }
}

int wait_signal(spmspv_handle_t h, int idx, long long* out) {
  const auto t0 = std::chrono::steady_clock::now();
  for (int spin = 0;; ++spin) {
    const long long v = __atomic_load_n(h->sig + idx, __ATOMIC_ACQUIRE);
    if (v != kSigPending) {
      *out = v;
      return SPMSPV_SUCCESS;
    }
    if ((spin & 255) == 255 &&
        std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(2000))
      break;  // a long kernel: block on the stream instead of spinning
    __builtin_ia32_pause();
  }
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  *out = __atomic_load_n(h->sig + idx, __ATOMIC_ACQUIRE);
  if (*out == kSigPending) return set_error(h, SPMSPV_DEVICE_ERROR, "device result never arrived");
  return SPMSPV_SUCCESS;
}

int readback_ynnz(spmspv_handle_t h, int* y_nnz) {
  __atomic_store_n(h->sig + 3, kSigPending, __ATOMIC_RELEASE);
  launch_signal_int(h->stream, &h->ctl->ynnz, h->sig_dev + 3);
  int rc = check_launch(h, "result count");
  if (rc) return rc;
  long long v = 0;
  if ((rc = wait_signal(h, 3, &v))) return rc;
  *y_nnz = (int)v;
  return SPMSPV_SUCCESS;
}

}  // namespace spmspv

// =================================================================== C-ABI
extern "C" {

const char* spmspv_last_error(spmspv_handle_t h) { return h ? h->err.c_str() : g_err.c_str(); }

int spmspv_create(spmspv_handle_t* out, int m, int n, int nnz) {
  if (!out) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle pointer");
  if (m <= 0 || n <= 0 || nnz < 0)
    return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "bad dimensions");
  auto* h = new spmspv_handle_s();
  h->m = m;
  h->n = n;
  h->nnz = nnz;
  h->team_col = std::min(64, std::max(4, pow2_team((double)nnz / n)));
  h->team_row = std::min(64, std::max(2, pow2_team((double)nnz / m)));
  hipError_t e = hipHostMalloc((void**)&h->h_pin, 64 * sizeof(int), hipHostMallocDefault);
  if (e == hipSuccess)
    e = hipHostMalloc((void**)&h->sig, kSigWords * sizeof(long long), hipHostMallocCoherent | hipHostMallocMapped);
  if (e == hipSuccess) e = hipHostGetDevicePointer((void**)&h->sig_dev, h->sig, 0);
  if (e != hipSuccess) {
    if (h->h_pin) (void)hipHostFree(h->h_pin);
    if (h->sig) (void)hipHostFree(h->sig);
    delete h;
    return set_error(nullptr, SPMSPV_DEVICE_ERROR, std::string("pinned buffer: ") + hipGetErrorString(e));
  }
  (void)hipEventCreate(&h->ev0);
  (void)hipEventCreate(&h->ev1);
  (void)hipEventCreate(&h->evk0);
  (void)hipEventCreate(&h->evk1);
  *out = h;
  return SPMSPV_SUCCESS;
}

int spmspv_bfs_release(spmspv_handle_t h);

int spmspv_destroy(spmspv_handle_t h) {
  if (!h) return SPMSPV_SUCCESS;
  (void)hipStreamSynchronize(h->stream);
  spmspv_bfs_release(h);
  pr_release(h);
  ms_release(h);
  if (h->arena) (void)hipFree(h->arena);
  if (h->csc_alloc) (void)hipFree(h->csc_alloc);
  if (h->flush_buf) (void)hipFree(h->flush_buf);
  if (h->batch_stash) (void)hipFree(h->batch_stash);
  if (h->csr5_mem) (void)hipFree(h->csr5_mem);
  if (h->h_pin) (void)hipHostFree(h->h_pin);
  if (h->sig) (void)hipHostFree(h->sig);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->evk0) (void)hipEventDestroy(h->evk0);
  if (h->evk1) (void)hipEventDestroy(h->evk1);
  delete h;
  return SPMSPV_SUCCESS;
}

int spmspv_set_stream(spmspv_handle_t h, void* stream) {
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  // the old stream's work drains first: a call may return once its count
  // is in host memory while its last kernel still restores the handle's
  // scratch state (marks, accumulator), which the new stream's work reads
  if (h->stream != static_cast<hipStream_t>(stream)) HIP_TRY(h, hipStreamSynchronize(h->stream));
  h->stream = static_cast<hipStream_t>(stream);
  return SPMSPV_SUCCESS;
}

int spmspv_warmup(spmspv_handle_t h) {
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  HIP_TRY(h, hipFree(nullptr));
  return SPMSPV_SUCCESS;
}

int spmspv_input_csr(spmspv_handle_t h, int nnz, int* row_ptr, int* col, float* val) {
  if (!h || !row_ptr || (nnz > 0 && (!col || !val)))
    return set_error(h, SPMSPV_INVALID_ARGUMENT, "InputCSR: null argument");
  if (nnz != h->nnz) return set_error(h, SPMSPV_INVALID_ARGUMENT, "InputCSR: nnz differs from handle");
  (void)hipStreamSynchronize(h->stream);
  bfs_drop_graphs(h);
  h->csr_row = row_ptr;
  h->csr_col = col;
  h->csr_val = val;
  h->row_ready = false;
  h->csr5_ready = false;
  h->mat_ready = false;
  h->bin_base_bits = -1;
  h->bucket_lo = 0;  // the bucket rule's range belongs to the previous matrix
  h->bucket_hi = -1;
  // a CSC this handle transposed from the previous CSR no longer matches
  if (h->csc_owned) h->csc_stale = true;
  h->prep_valid = false;
  h->binlen_host = -1;
  h->enc5_valid = false;
  return SPMSPV_SUCCESS;
}

int spmspv_input_csc(spmspv_handle_t h, int nnz, int* row, int* col_ptr, float* val) {
  if (!h || !col_ptr || (nnz > 0 && (!row || !val)))
    return set_error(h, SPMSPV_INVALID_ARGUMENT, "InputCSC: null argument");
  if (nnz != h->nnz) return set_error(h, SPMSPV_INVALID_ARGUMENT, "InputCSC: nnz differs from handle");
  (void)hipStreamSynchronize(h->stream);
  bfs_drop_graphs(h);
  if (h->csc_alloc) {
    (void)hipFree(h->csc_alloc);
    h->csc_alloc = nullptr;
  }
  h->csc_row = row;
  h->csc_colptr = col_ptr;
  h->csc_val = val;
  h->csc_owned = false;
  h->csc_stale = false;
  h->mat_ready = false;
  h->bin_base_bits = -1;
  h->bucket_lo = 0;
  h->bucket_hi = -1;
  h->prep_valid = false;
  h->binlen_host = -1;
  return SPMSPV_SUCCESS;
}

int spmspv_to_csc(spmspv_handle_t h) {
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  if (!h->csr_row) return set_error(h, SPMSPV_MAT_FORMAT_ERR, "ToCSC needs CSR input");
  int rc = ensure_arena(h);
  if (rc) return rc;
  if (!h->csc_alloc) {
    const size_t a = align_up(sizeof(int) * (size_t)(h->n + 1), 256);
    const size_t b = align_up(sizeof(int) * (size_t)std::max(h->nnz, 1), 256);
    const size_t c = sizeof(float) * (size_t)std::max(h->nnz, 1);
    HIP_TRY(h, hipMalloc(&h->csc_alloc, a + b + c));
    char* p = static_cast<char*>(h->csc_alloc);
    h->csc_colptr = reinterpret_cast<int*>(p);
    h->csc_row = reinterpret_cast<int*>(p + a);
    h->csc_val = reinterpret_cast<float*>(p + a + b);
    h->csc_owned = true;
  }
  bfs_drop_graphs(h);
  h->prep_valid = false;
  h->binlen_host = -1;
  HIP_TRY(h, launch_csr_to_csc_sorted(h->stream, h->csr_row, h->csr_col, h->csr_val, h->m, h->n, h->nnz,
                           h->csc_colptr, h->csc_row, h->csc_val, h->sb, h->n_dev, h->ctl, next_ep(h, 4)));
  h->launches += 6;
  rc = check_launch(h, "ToCSC");
  if (rc) return rc;
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  h->csc_stale = false;
  return SPMSPV_SUCCESS;
}

int spmspv_alloc_pre_buffer(spmspv_handle_t h) {
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  int rc = ensure_arena(h);
  if (rc) return rc;
  // Everything a later call could allocate is reserved here: the CSR5 copy
  // at its largest width, the BFS, multi-source BFS and PageRank buffers and a batch stash of
  // kBatchReserve vectors.
  if (h->csr_row) {
    size_t most = 0;
    for (int S = 1; S <= 32; ++S) most = std::max(most, csr5_bytes(h->m, h->n, h->nnz, S));
    if (h->csr5_bytes < most) {
      HIP_TRY(h, hipStreamSynchronize(h->stream));
      if (h->csr5_mem) HIP_TRY(h, hipFree(h->csr5_mem));
      h->csr5_mem = nullptr;
      h->csr5_bytes = 0;
      h->csr5_ready = false;
      HIP_TRY(h, hipMalloc(&h->csr5_mem, most));
      h->csr5_bytes = most;
    }
  }
  if (h->m == h->n && ((rc = bfs_alloc(h)) || (rc = pr_alloc(h)) || (rc = ms_alloc(h)))) return rc;
  if (h->batch_cap < spmspv_handle_s::kBatchReserve) {
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    if (h->batch_stash) HIP_TRY(h, hipFree(h->batch_stash));
    h->batch_stash = nullptr;
    h->batch_cap = 0;
    HIP_TRY(h, hipMalloc(&h->batch_stash, sizeof(long long) * spmspv_handle_s::kBatchReserve));
    h->batch_cap = spmspv_handle_s::kBatchReserve;
  }
  if (h->csr_row && (rc = ensure_row_tiles(h))) return rc;
  if ((h->csr_row || h->csc_colptr) && (rc = ensure_mat_features(h))) return rc;
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return SPMSPV_SUCCESS;
}

int spmspv_dealloc_pre_buffer(spmspv_handle_t h) {
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  (void)hipStreamSynchronize(h->stream);
  bfs_drop_graphs(h);
  if (h->arena) (void)hipFree(h->arena);
  h->arena = nullptr;
  h->reserved = false;
  h->row_ready = false;
  h->mat_ready = false;
  h->bin_base_bits = -1;
  h->prep_valid = false;
  h->have_weights = false;
  h->have_bfs_weights = false;
  return SPMSPV_SUCCESS;
}

int spmspv_hola_preprocess(spmspv_handle_t h) {
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  if (!h->csr_row) return set_error(h, SPMSPV_FORMAT_ERR_HOLA_SPMV, "holaPreprocess needs CSR");
  int rc = ensure_row_tiles(h);
  if (rc) return rc;
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return SPMSPV_SUCCESS;
}
int spmspv_hola_postprocess(spmspv_handle_t h) { return h ? SPMSPV_SUCCESS : SPMSPV_INVALID_ARGUMENT; }
int spmspv_csr5_preprocess(spmspv_handle_t h) {
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  int rc = ensure_csr5(h);
  if (rc) return rc;
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return SPMSPV_SUCCESS;
}
int spmspv_csr5_postprocess(spmspv_handle_t h) {
  // frees the CSR5 copy (the caller's CSR was never modified)
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  if (h->csr5_mem) HIP_TRY(h, hipFree(h->csr5_mem));
  h->csr5_mem = nullptr;
  h->csr5_bytes = 0;
  h->csr5_ready = false;
  return SPMSPV_SUCCESS;
}
int spmspv_set_sigma(spmspv_handle_t h, int sigma) {
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  if (sigma > 32) return set_error(h, SPMSPV_INVALID_ARGUMENT, "sigma must be at most 32");
  h->sigma = sigma;  // the CSR5 copy is rebuilt on next use when its width changes
  return SPMSPV_SUCCESS;
}
int spmspv_set_spmv_type(spmspv_handle_t h, int code) {
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  if (code != 0 && code != 1) return set_error(h, SPMSPV_INVALID_ARGUMENT, "spmv type must be 0 or 1");
  h->spmv_type = code;
  return SPMSPV_SUCCESS;
}

int spmspv_set_vector_type(spmspv_handle_t h, int code) {
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  if (code < 0 || code > 2) return set_error(h, SPMSPV_INVALID_ARGUMENT, "vector type must be 0..2");
  h->vec_type = code;
  return SPMSPV_SUCCESS;
}

int spmspv_set_sparse_x(spmspv_handle_t h, int* key, float* val, int xnnz) {
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  if (xnnz < 0 || xnnz > h->n) return set_error(h, SPMSPV_INVALID_ARGUMENT, "xnnz out of range");
  if (xnnz > 0 && (!key || !val)) return set_error(h, SPMSPV_INVALID_ARGUMENT, "null x arrays");
  h->x_key = key;
  h->x_val = val;
  h->xnnz = xnnz;
  h->vec_type = SPMSPV_VEC_SPARSE;
  h->sparse_valid = true;
  h->dense_valid = false;
  h->bits_valid = false;
  h->enc_valid = h->enc5_valid = false;
  h->prep_valid = false;
  h->binlen_host = -1;
  return SPMSPV_SUCCESS;
}

int spmspv_set_x(spmspv_handle_t h, float* x) {
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  h->x_dense_user = x;
  // a dense x given on its own becomes the current x (vector type 1)
  if (h->vec_type == SPMSPV_VEC_DENSE) {
    h->dense_valid = x != nullptr;
    h->sparse_valid = false;
    h->bits_valid = false;
    h->enc_valid = h->enc5_valid = false;
    h->prep_valid = false;
    h->binlen_host = -1;
  } else {
    h->dense_valid = false;
    h->enc_valid = h->enc5_valid = false;
  }
  return SPMSPV_SUCCESS;
}

int spmspv_set_bitvector(spmspv_handle_t h, int* bits, int num_ints) {
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  if (bits && num_ints < (h->n + 31) / 32)
    return set_error(h, SPMSPV_INVALID_ARGUMENT, "bit array shorter than ceil(n/32) words");
  h->bits_user = reinterpret_cast<unsigned*>(bits);
  h->num_ints_user = num_ints;
  h->bits_valid = false;
  h->enc_valid = h->enc5_valid = false;
  return SPMSPV_SUCCESS;
}

int spmspv_set_sparse_x_dense_x(spmspv_handle_t h, int* key, float* val, float* x, int xnnz) {
  int rc = spmspv_set_sparse_x(h, key, val, xnnz);
  if (rc) return rc;
  h->x_dense_user = x;
  h->dense_valid = false;  // the caller's dense copy is filled on demand
  h->enc_valid = h->enc5_valid = false;
  return SPMSPV_SUCCESS;
}

int spmspv_sparse2dense(spmspv_handle_t h) {
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  if (!h->sparse_valid) return set_error(h, SPMSPV_VEC_FORMAT_ERR, "sparse2dense: x is not sparse");
  int rc = ensure_arena(h);
  if (rc) return rc;
  float* d = h->dense_x();
  launch_fill_f32(h->stream, d, h->n, 0.f);
  launch_scatter(h->stream, h->x_key, h->x_val, h->xnnz, nullptr, h->n, d);
  h->dense_valid = true;
  return check_launch(h, "sparse2dense");
}

int spmspv_sparse2bitarray(spmspv_handle_t h) {
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  if (!h->sparse_valid) return set_error(h, SPMSPV_VEC_FORMAT_ERR, "sparse2bitarray: x is not sparse");
  int rc = spmspv_sparse2dense(h);
  if (rc) return rc;
  h->bits_valid = false;
  h->enc_valid = h->enc5_valid = false;
  rc = need_bits(h);
  if (rc) return rc;
  h->vec_type = SPMSPV_VEC_BITARRAY;
  return SPMSPV_SUCCESS;
}

int spmspv_dense2bitarray(spmspv_handle_t h) {
  // bits where the dense x is nonzero (the reference's dense2bitarray_ker
  // tested `!val`, spmspv.h:1354-1369, i.e. marked the zeros)
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  if (!h->dense_valid) return set_error(h, SPMSPV_VEC_FORMAT_ERR, "dense2bitarray: no dense x");
  int rc = ensure_arena(h);
  if (rc) return rc;
  launch_bitmap_from_dense(h->stream, h->dense_x(), h->n, h->bits());
  h->bits_valid = true;
  h->vec_type = SPMSPV_VEC_BITARRAY;
  return check_launch(h, "dense2bitarray");
}

int spmspv_dense2sparse(spmspv_handle_t h) {
  // keeps entries != 0 (spmspv.h:1176-1223); writes the handle's own key/val
  // buffers unless the caller set sparse buffers with set_sparse_x
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  if (!h->dense_valid) return set_error(h, SPMSPV_VEC_FORMAT_ERR, "dense2sparse: no dense x");
  int rc = ensure_arena(h);
  if (rc) return rc;
  int* k = h->x_key ? h->x_key : h->xk_int;
  float* v = h->x_val ? h->x_val : h->xv_int;
  // the handle's own bit array as scratch (the compaction clears it): a
  // bit array the caller lent with set_bitvector keeps its contents
  unsigned* b = h->bits_int;
  launch_bitmap_from_dense(h->stream, h->dense_x(), h->n, b);
  launch_compact(h->stream, b, nullptr, h->n, h->dense_x(), 0, k, v, h->scan_state, h->ctl,
                 next_ep(h, 1));
  h->launches++;
  rc = check_launch(h, "dense2sparse");
  if (rc) return rc;
  // compaction cleared the bit array; read the count back
  int cnt = 0;
  rc = readback_ynnz(h, &cnt);
  if (rc) return rc;
  h->x_key = k;
  h->x_val = v;
  h->xnnz = cnt;
  h->sparse_valid = true;
  if (!h->bits_user) h->bits_valid = false;  // the scratch above was the internal array
  h->enc_valid = h->enc5_valid = false;
  h->prep_valid = false;
  h->binlen_host = -1;
  h->vec_type = SPMSPV_VEC_SPARSE;
  return SPMSPV_SUCCESS;
}

int spmspv_bitarray2sparse(spmspv_handle_t h) {
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  if (h->vec_type != SPMSPV_VEC_BITARRAY)
    return set_error(h, SPMSPV_VEC_FORMAT_ERR, "bitarray2sparse: input is not a bit array");
  return spmspv_dense2sparse(h);
}

int spmspv_dense2sparse_outer(spmspv_handle_t h, int m, float* d_dense, int* ynnz, int* d_key,
                              float* d_val) {
  if (!h || !d_dense || m <= 0 || m > std::max(h->m, h->n))
    return set_error(h, SPMSPV_INVALID_ARGUMENT, "dense2sparse_outer: bad arguments");
  int rc = ensure_arena(h);
  if (rc) return rc;
  // use the m-sized bitmap region when it fits, else the n-sized one
  unsigned* b = m <= h->m ? h->bitmap : h->bits_int;
  launch_bitmap_from_dense(h->stream, d_dense, m, b);
  launch_compact(h->stream, b, nullptr, m, d_dense, 0, d_key, d_val, h->scan_state, h->ctl,
                 next_ep(h, 1));
  h->launches++;
  rc = check_launch(h, "dense2sparse_outer");
  if (rc) return rc;
  if (ynnz) return readback_ynnz(h, ynnz);
  return SPMSPV_SUCCESS;
}

int spmspv_sparse2dense_outer(spmspv_handle_t h, int ynnz, int* d_key, float* d_val, int m,
                              float* d_dense) {
  if (!h || !d_dense || m <= 0 || ynnz < 0 || ynnz > m || (ynnz > 0 && (!d_key || !d_val)))
    return set_error(h, SPMSPV_INVALID_ARGUMENT, "sparse2dense_outer: bad arguments");
  launch_fill_f32(h->stream, d_dense, m, 0.f);
  launch_scatter(h->stream, d_key, d_val, ynnz, nullptr, m, d_dense);
  return check_launch(h, "sparse2dense_outer");
}

// ------------------------------------------------------------- features
int spmspv_vec_features(spmspv_handle_t h, long long* bin_len, int* max_len, int* min_len) {
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  int rc;
  if ((rc = ensure_arena(h)) || (rc = need_csc(h)) || (rc = need_sparse_x(h))) return rc;
  if (!h->prep_valid && (rc = run_prep(h, true, nullptr))) return rc;
  HIP_TRY(h, hipMemcpyAsync(h->h_pin, &h->ctl->bin_len, 16, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  long long bl;
  std::memcpy(&bl, h->h_pin, sizeof(bl));
  h->binlen_host = bl;
  if (bin_len) *bin_len = bl;
  if (max_len) *max_len = h->h_pin[2];
  if (min_len) *min_len = h->h_pin[3];
  return SPMSPV_SUCCESS;
}

int spmspv_compute_binlen(spmspv_handle_t h, long long* bin_len) {
  return spmspv_vec_features(h, bin_len, nullptr, nullptr);
}

int spmspv_mat_features(spmspv_handle_t h, float out[7]) {
  if (!h || !out) return set_error(h, SPMSPV_INVALID_ARGUMENT, "null argument");
  int rc = ensure_mat_features(h);
  if (rc) return rc;
  MatFeat mf;
  HIP_TRY(h, hipMemcpy(&mf, h->mat, sizeof(mf), hipMemcpyDeviceToHost));
  out[0] = (float)mf.max_row;
  out[1] = (float)mf.min_row;
  out[2] = (float)mf.avg_row;
  out[3] = (float)mf.range;
  out[4] = (float)mf.std_row;
  out[5] = (float)mf.equality;
  out[6] = (float)mf.gini;
  return SPMSPV_SUCCESS;
}

int spmspv_extract_mat_feature(spmspv_handle_t h, float* std_row) {
  if (!h || !std_row) return set_error(h, SPMSPV_INVALID_ARGUMENT, "null argument");
  int rc = ensure_mat_features(h);
  if (rc) return rc;
  MatFeat mf;
  HIP_TRY(h, hipMemcpy(&mf, h->mat, sizeof(mf), hipMemcpyDeviceToHost));
  *std_row = (float)mf.std_trunc;
  return SPMSPV_SUCCESS;
}

int spmspv_compute_gm(spmspv_handle_t h, double out[3]) {
  if (!h || !out) return set_error(h, SPMSPV_INVALID_ARGUMENT, "null argument");
  long long bl = h->binlen_host;
  if (bl < 0) {
    int rc = spmspv_compute_binlen(h, &bl);
    if (rc) return rc;
  }
  const double m = h->m, nnz = h->nnz, x = h->xnnz, b = (double)bl;
  out[0] = 4.0 * (3.0 * x + 3.0 * b);
  out[1] = 4.0 * ((m + 1) + 2.0 * b + nnz + m);
  out[2] = 4.0 * ((m + 1) + 3.0 * nnz + m);
  return SPMSPV_SUCCESS;
}

// ------------------------------------------------------------- kernels
int spmspv_run(spmspv_handle_t h, int kid, float alpha, float* d_y_dense, int* d_y_key,
               float* d_y_val, int* y_nnz) {
  (void)alpha;
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  h->count_to_host = y_nnz != nullptr;
  int rc = run_kernel(h, kid, d_y_dense, d_y_key, d_y_val, y_nnz != nullptr);
  h->count_to_host = false;
  if (rc) return rc;
  if (!y_nnz) return SPMSPV_SUCCESS;
  if (!h->count_sent) return readback_ynnz(h, y_nnz);
  long long v = 0;
  if ((rc = wait_signal(h, 3, &v))) return rc;
  *y_nnz = (int)v;
  return SPMSPV_SUCCESS;
}

int* spmspv_ynnz_device(spmspv_handle_t h) {
  if (!h || !h->reserved) return nullptr;
  return &h->ctl->ynnz;
}

int spmspv_csc_nosort_naive(spmspv_handle_t h, float alpha, int* ynnz, float* d_y) {
  return spmspv_run(h, SPMSPV_K_COL_NAIVE_ATOMIC, alpha, d_y, nullptr, nullptr, ynnz);
}
int spmspv_csc_nosort_naive_keyval(spmspv_handle_t h, float alpha, int* ynnz, float* d_y, int* k,
                                   float* v) {
  return spmspv_run(h, SPMSPV_K_COL_NAIVE_ATOMIC, alpha, d_y, k, v, ynnz);
}
int spmspv_csc_sort_naive(spmspv_handle_t h, float alpha, int* ynnz, int* k, float* v, int method) {
  (void)method;
  return spmspv_run(h, SPMSPV_K_COL_NAIVE_SORT, alpha, nullptr, k, v, ynnz);
}
int spmspv_csc_nosort_merge(spmspv_handle_t h, int is_binlen_computed, float alpha, int* ynnz,
                            float* d_y, int* k, float* v) {
  (void)is_binlen_computed;
  return spmspv_run(h, SPMSPV_K_COL_LB_ATOMIC, alpha, d_y, k, v, ynnz);
}
int spmspv_csc_nosort_merge_keyval(spmspv_handle_t h, int is_binlen_computed, float alpha,
                                   int* ynnz, float* d_y, int* k, float* v) {
  (void)is_binlen_computed;
  return spmspv_run(h, SPMSPV_K_COL_LB_ATOMIC, alpha, d_y, k, v, ynnz);
}
int spmspv_csc_sort_merge(spmspv_handle_t h, int is_binlen_computed, float alpha, int* ynnz,
                          float* d_y, int* k, float* v) {
  (void)is_binlen_computed;
  return spmspv_run(h, SPMSPV_K_COL_LB_SORT, alpha, d_y, k, v, ynnz);
}
int spmspv_csc_bucket(spmspv_handle_t h, float alpha, int* ynnz, int* k, float* v, int buckets) {
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  if (buckets != 0) {
    int bits = 1;
    while (bits < 10 && (1 << bits) < buckets) ++bits;
    if (buckets < 2 || buckets > 1024 || (1 << bits) != buckets)
      return set_error(h, SPMSPV_INVALID_ARGUMENT, "buckets must be 0 or a power of two in [2, 1024]");
    h->bucket_bits = bits;
  }
  return spmspv_run(h, SPMSPV_K_COL_BUCKET, alpha, nullptr, k, v, ynnz);
}
int spmspv_csc_my_nosort(spmspv_handle_t h, int is_binlen_computed, float alpha, int* ynnz, float* d_y,
                         int* k, float* v) {
  (void)is_binlen_computed;
  return spmspv_run(h, SPMSPV_K_COL_TWO_LAYER_ATOMIC, alpha, d_y, k, v, ynnz);
}
int spmspv_csc_my_nosort_keyval(spmspv_handle_t h, int is_binlen_computed, float alpha, int* ynnz, float* d_y,
                                int* k, float* v) {
  (void)is_binlen_computed;
  return spmspv_run(h, SPMSPV_K_COL_TWO_LAYER_ATOMIC, alpha, d_y, k, v, ynnz);
}
int spmspv_csc_sort_my(spmspv_handle_t h, int is_binlen_computed, float alpha, int* ynnz, float* d_y, int* k,
                       float* v) {
  (void)is_binlen_computed;
  return spmspv_run(h, SPMSPV_K_COL_TWO_LAYER_SORT, alpha, d_y, k, v, ynnz);
}
int spmspv_naivespmspv(spmspv_handle_t h, float alpha, float* d_y) {
  return spmspv_run(h, SPMSPV_K_ROW_NAIVE, alpha, d_y, nullptr, nullptr, nullptr);
}
// holaspmspv / csr5spmspv run their format whatever set_spmv_type says
static int run_with_type(spmspv_handle_t h, int type, int kid, float alpha, float* d_y) {
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  const int saved = h->spmv_type;
  h->spmv_type = type;
  const int rc = spmspv_run(h, kid, alpha, d_y, nullptr, nullptr, nullptr);
  h->spmv_type = saved;
  return rc;
}
int spmspv_holaspmspv(spmspv_handle_t h, float alpha, float* d_y) {
  return run_with_type(h, 0, SPMSPV_K_ROW_LB, alpha, d_y);
}
int spmspv_csr5spmspv(spmspv_handle_t h, float alpha, float* d_y) {
  return run_with_type(h, 1, SPMSPV_K_ROW_LB, alpha, d_y);
}
int spmspv_spmspv(spmspv_handle_t h, float alpha, float* d_y) {
  return spmspv_run(h, SPMSPV_K_ROW_LB, alpha, d_y, nullptr, nullptr, nullptr);
}
int spmspv_naivespmv(spmspv_handle_t h, float alpha, float* d_y) {
  return spmspv_run(h, SPMSPV_K_SPMV_NAIVE, alpha, d_y, nullptr, nullptr, nullptr);
}
int spmspv_holaspmv(spmspv_handle_t h, float alpha, float* d_y) {
  return run_with_type(h, 0, SPMSPV_K_SPMV_LB, alpha, d_y);
}
int spmspv_csr5spmv(spmspv_handle_t h, float alpha, float* d_y) {
  return run_with_type(h, 1, SPMSPV_K_SPMV_LB, alpha, d_y);
}
int spmspv_spmv(spmspv_handle_t h, float alpha, float* d_y) {
  return spmspv_run(h, SPMSPV_K_SPMV_LB, alpha, d_y, nullptr, nullptr, nullptr);
}

// ------------------------------------------------------------- selector
int spmspv_selector_set_weights(spmspv_handle_t h, const float* w, int count) {
  if (!h || !w || count != SPMSPV_MLP_NWEIGHTS)
    return set_error(h, SPMSPV_INVALID_ARGUMENT, "selector weights: expected SPMSPV_MLP_NWEIGHTS floats");
  int rc = ensure_arena(h);
  if (rc) return rc;
  HIP_TRY(h, hipMemcpy(h->weights_dev, w, sizeof(float) * count, hipMemcpyHostToDevice));
  h->have_weights = true;
  h->sel_mode = 0;
  return SPMSPV_SUCCESS;
}

int spmspv_bfs_selector_set_weights(spmspv_handle_t h, const float* w, int count) {
  if (!h || !w || count != SPMSPV_MLP_NWEIGHTS)
    return set_error(h, SPMSPV_INVALID_ARGUMENT, "bfs selector weights: expected SPMSPV_MLP_NWEIGHTS floats");
  int rc = ensure_arena(h);
  if (rc) return rc;
  HIP_TRY(h, hipMemcpy(h->bfs_weights_dev, w, sizeof(float) * count, hipMemcpyHostToDevice));
  h->have_bfs_weights = true;
  return SPMSPV_SUCCESS;
}

int spmspv_selector_mode(spmspv_handle_t h, int mode, int fixed_kernel) {
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  if (mode < 0 || mode > 2 || fixed_kernel < 0 || fixed_kernel >= SPMSPV_NUM_SLOTS)
    return set_error(h, SPMSPV_INVALID_ARGUMENT, "selector mode");
  if (mode == 0 && !h->have_weights)
    return set_error(h, SPMSPV_INVALID_ARGUMENT, "selector mode 0 needs weights");
  h->sel_mode = mode;
  h->sel_fixed = fixed_kernel;
  return SPMSPV_SUCCESS;
}

int spmspv_select(spmspv_handle_t h, int* kernel_id) {
  if (!h || !kernel_id) return set_error(h, SPMSPV_INVALID_ARGUMENT, "null argument");
  int rc;
  if ((rc = ensure_arena(h)) || (rc = need_csc(h)) || (rc = ensure_mat_features(h)) ||
      (rc = need_sparse_x(h)))
    return rc;
  // the selecting workgroup writes the selector word straight to a host word
  SelectorArgs sel{h->have_weights ? h->weights_dev : nullptr, h->mat, h->sel_mode, h->sel_fixed,
                   h->sig_dev + 1};
  __atomic_store_n(h->sig + 1, kSigPending, __ATOMIC_RELEASE);
  if ((rc = run_prep(h, true, &sel))) return rc;
  long long word = 0;
  if ((rc = wait_signal(h, 1, &word))) return rc;
  h->binlen_host = sel_word_binlen(word);
  *kernel_id = sel_word_choice(word);
  return SPMSPV_SUCCESS;
}

int spmspv_features(spmspv_handle_t h, float out[SPMSPV_NUM_FEATURES]) {
  if (!h || !out) return set_error(h, SPMSPV_INVALID_ARGUMENT, "null argument");
  int kid;
  int rc = spmspv_select(h, &kid);
  if (rc) return rc;
  HIP_TRY(h, hipMemcpy(out, h->ctl->feat, sizeof(float) * SPMSPV_NUM_FEATURES, hipMemcpyDeviceToHost));
  return SPMSPV_SUCCESS;
}

// Speculative dispatch. The selector's choice reaches the host one round
// trip after the selector kernel; instead of leaving the device idle for
// it, the slot this handle's selector chose last time for an x of the same
// size class (xnnz within a power of two) is queued right behind the
// selector, every launch carrying a gate on the device-side choice. A right
// guess costs nothing more; a wrong one costs the gated launches (each exits
// at once) and the chosen slot is queued as usual. Only slots whose launches
// need nothing from the host about x are guessed: the atomic and sort column
// slots (the sort then sizes its grids from the device-side bin_len), the
// bucket slot and the load-balanced row slots.
// slots whose launches read the column scan of x (ptr / xbeg / split)
static bool needs_scan(int kid) {
  return kid == SPMSPV_K_COL_NAIVE_SORT || kid == SPMSPV_K_COL_LB_ATOMIC ||
         kid == SPMSPV_K_COL_LB_SORT || kid == SPMSPV_K_COL_BUCKET;
}

static bool speculable(spmspv_handle_t h, int kid) {
  if (kid == SPMSPV_K_COL_NAIVE_ATOMIC || kid == SPMSPV_K_COL_LB_ATOMIC ||
      kid == SPMSPV_K_COL_NAIVE_SORT || kid == SPMSPV_K_COL_LB_SORT)
    return true;
  // the bucket slot once its bin layout exists (building it synchronises)
  if (kid == SPMSPV_K_COL_BUCKET) return h->bin_base_bits == h->bucket_bits;
  if (kid == SPMSPV_K_ROW_LB || kid == SPMSPV_K_SPMV_LB)
    return h->csr_row && (h->spmv_type == 1 ? h->csr5_ready : h->row_ready);
  return false;
}

static int size_class(int xnnz) { return xnnz > 0 ? 32 - __builtin_clz((unsigned)xnnz) : 0; }

int spmspv_run_auto(spmspv_handle_t h, float alpha, float* d_y_dense, int* d_y_key, float* d_y_val,
                    int* y_nnz, int* kernel_used) {
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  int rc;
  if ((rc = ensure_arena(h)) || (rc = need_csc(h)) || (rc = ensure_mat_features(h)) ||
      (rc = need_sparse_x(h)))
    return rc;
  // a large x's features-only launch may return the sampled estimate of
  // bin_len (only row slots, which need none, skip the exact column scan)
  SelectorArgs sel{h->have_weights ? h->weights_dev : nullptr, h->mat, h->sel_mode, h->sel_fixed,
                   h->sig_dev + 1, 1, h->prev_choice, h->switch_penalty, h->bucket_lo, h->bucket_hi};
  h->enc5_clear_pending = -1;
  __atomic_store_n(h->sig + 1, kSigPending, __ATOMIC_RELEASE);
  const int cls = size_class(h->xnnz);
  const int guess = h->spec_last[cls] - 1;
  // a small x with a sparse result goes through the single-launch path,
  // which also selects when the products do not fit; when the last call of
  // this size class fitted, nothing is queued behind it
  // (a selector pinned to a row slot keeps it: the path computes a column slot's result)
  const bool fused = h->small_x_path && h->fuse_cap > 0 && !d_y_dense && h->xnnz <= kFuseKeys &&
                     !(h->sel_mode == 2 && (h->sel_fixed < 0 || h->sel_fixed > SPMSPV_K_COL_LB_SORT));
  const bool spec = h->speculate && guess >= 0 && speculable(h, guess) && !(fused && h->spec_fused[cls]);
  if (fused) {
    launch_col_fused(h->stream, h->csc_colptr, h->csc_row, h->csc_val, h->x_key, h->x_val, h->n, h->xnnz, h->m,
                     h->ptr, h->xbeg, h->split, h->ctl, sel, d_y_key ? d_y_key : h->yk_tmp,
                     d_y_val ? d_y_val : h->yv_tmp, nullptr, h->fuse_cap);
    h->launches++;
    if ((rc = check_launch(h, "small-x column path"))) return rc;
    h->prep_valid = true;
  } else if (spec && !needs_scan(guess)) {
    // the guessed slot needs none of the column scan: features only, with
    // x's support encoding (which the row slots read) scattered on the way
    const bool row = guess >= SPMSPV_K_ROW_NAIVE && guess <= SPMSPV_K_SPMV_LB;
    // the CSR5 slots read x in the copy's column order
    const bool order5 = row && h->spmv_type == 1 && h->csr5_ready && h->csr5.cperm &&
                        (guess == SPMSPV_K_ROW_LB || guess == SPMSPV_K_SPMV_LB);
    float* enc = nullptr;
    if (row && !(order5 ? h->enc5_valid : h->enc_valid)) {
      if (order5) {
        enc = enc5_for_scatter(h);
      } else {
        enc = h->x_enc;
        launch_fill_f32(h->stream, enc, h->n, 0.f);
        h->launches++;
      }
    }
    launch_select_feat(h->stream, h->csc_colptr, h->x_key, h->x_val, h->n, h->xnnz, h->sel_part, h->ctl,
                       sel, enc, order5 ? h->csr5.cperm : nullptr);
    h->launches++;
    if ((rc = check_launch(h, "selector"))) return rc;
    if (enc) (order5 ? h->enc5_valid : h->enc_valid) = true;
  } else if ((rc = run_prep(h, true, &sel))) {
    return rc;
  }
  if (spec) {
    struct Ungate {
      spmspv_handle_t h;
      ~Ungate() { h->gate = Gate{}; }
    } ungate{h};
    h->gate = Gate{&h->ctl->kernel_choice, guess};
    h->count_to_host = y_nnz != nullptr;
    rc = run_kernel(h, guess, d_y_dense, d_y_key, d_y_val, y_nnz != nullptr);
    h->count_to_host = false;
    if (rc) return rc;
    if (y_nnz && !h->count_sent) {
      __atomic_store_n(h->sig + 3, kSigPending, __ATOMIC_RELEASE);
      launch_signal_int(h->stream, &h->ctl->ynnz, h->sig_dev + 3, h->gate);
      if ((rc = check_launch(h, "result count"))) return rc;
    }
  }
  long long word = 0;
  const int clear_pending = h->enc5_clear_pending;
  h->enc5_clear_pending = -1;
  if ((rc = wait_signal(h, 1, &word))) return rc;
  const int kid = sel_word_choice(word);
  if (kid < 0 || kid > SPMSPV_K_COL_BUCKET)  // the eight classes or the bucket rule's slot
    return set_error(h, SPMSPV_DEVICE_ERROR, "selector returned no slot");
  if (fused) h->spec_fused[cls] = sel_word_done(word);
  h->prev_choice = kid;
  if (sel_word_done(word)) {
    // the single launch finished the call: its word carries y_nnz
    h->binlen_host = -1;
    if (kernel_used) *kernel_used = kid;
    if (y_nnz) *y_nnz = (int)sel_word_binlen(word);
    return SPMSPV_SUCCESS;
  }
  h->binlen_host = sel_word_estimated(word) ? -1 : sel_word_binlen(word);
  h->spec_last[cls] = (signed char)(kid + 1);
  if (kernel_used) *kernel_used = kid;
  if (spec && kid == guess) {
    h->spec_hits++;
    // the gated tile kernel ran, and with it the clearing of the idle buffer
    if (clear_pending >= 0) h->enc5_clean[clear_pending] = true;
    if (y_nnz) {
      long long v = 0;
      if ((rc = wait_signal(h, 3, &v))) return rc;
      *y_nnz = (int)v;
    }
    return SPMSPV_SUCCESS;
  }
  if (spec) h->spec_misses++;
  return spmspv_run(h, kid, alpha, d_y_dense, d_y_key, d_y_val, y_nnz);
}

int spmspv_autotune_report(spmspv_handle_t h, int out[10]) {
  if (!h || !out) return set_error(h, SPMSPV_INVALID_ARGUMENT, "null argument");
  out[0] = h->team_col;
  out[1] = h->team_row;
  out[2] = h->row_vt;
  out[3] = h->spmv_type;
  out[4] = h->spmv_type == 1 ? csr5_sigma_of(h->sigma, h->m, h->nnz) : 0;
  out[5] = 1 << h->bucket_bits;
  out[6] = h->small_x_path ? h->fuse_cap : 0;
  out[7] = h->small_sort_cap;
  // the bucket rule's bin_len range (0, -1: off; INT_MAX as the upper end: none)
  out[8] = (int)std::min<long long>(h->bucket_lo, INT_MAX);
  out[9] = h->bucket_hi < h->bucket_lo ? -1 : (int)std::min<long long>(h->bucket_hi, INT_MAX);
  return SPMSPV_SUCCESS;
}

int spmspv_set_bucket_range(spmspv_handle_t h, long long lo, long long hi) {
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  if (hi >= lo && (h->csr_row || h->csc_colptr)) {
    // a rule the slot cannot serve on this matrix (too many row bins) would
    // fail every auto call in its range: refused, the rule left off
    int rc = ensure_arena(h);
    if (!rc) rc = ensure_bin_layout(h);
    if (rc) {
      h->bucket_lo = 0;
      h->bucket_hi = -1;
      return rc;
    }
  }
  h->bucket_lo = lo;
  h->bucket_hi = hi;
  return SPMSPV_SUCCESS;
}

int spmspv_set_switch_penalty(spmspv_handle_t h, float penalty) {
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  if (!(penalty >= 0.f)) return set_error(h, SPMSPV_INVALID_ARGUMENT, "penalty must be >= 0");
  h->switch_penalty = penalty;
  return SPMSPV_SUCCESS;
}

int spmspv_set_small_x_path(spmspv_handle_t h, int on) {
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  h->small_x_path = on != 0;
  return SPMSPV_SUCCESS;
}

int spmspv_set_speculation(spmspv_handle_t h, int on) {
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  h->speculate = on != 0;
  return SPMSPV_SUCCESS;
}

int spmspv_speculation_stats(spmspv_handle_t h, long long* hits, long long* misses) {
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  if (hits) *hits = h->spec_hits;
  if (misses) *misses = h->spec_misses;
  return SPMSPV_SUCCESS;
}

int spmspv_time_main_kernel(spmspv_handle_t h, int on) {
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  h->time_main = on != 0;
  h->main_timed = false;
  return SPMSPV_SUCCESS;
}

int spmspv_main_kernel_ms(spmspv_handle_t h, float* ms) {
  if (!h || !ms) return set_error(h, SPMSPV_INVALID_ARGUMENT, "null argument");
  if (!h->main_timed)
    return set_error(h, SPMSPV_INVALID_ARGUMENT, "no timed main kernel (row load-balanced slots only)");
  HIP_TRY(h, hipEventSynchronize(h->evk1));
  HIP_TRY(h, hipEventElapsedTime(ms, h->evk0, h->evk1));
  return SPMSPV_SUCCESS;
}

// ------------------------------------------------------------- autotuner
int spmspv_autotune(spmspv_handle_t h, int reps, int* report) {
  return spmspv_autotune2(h, reps, 1, report);
}

int spmspv_autotune2(spmspv_handle_t h, int reps, int csr5, int* report) {
  if (!h || reps <= 0) return set_error(h, SPMSPV_INVALID_ARGUMENT, "autotune: bad arguments");
  int rc;
  if ((rc = ensure_arena(h)) || (rc = need_csc(h)) || (rc = need_sparse_x(h))) return rc;
  auto time_slot = [&](int kid, float* out) -> int {
    std::vector<float> ms;
    for (int r = 0; r <= reps; ++r) {
      HIP_TRY(h, hipEventRecord(h->ev0, h->stream));
      int cnt = 0;  // as a caller runs it: the count comes back to the host
      int e = spmspv_run(h, kid, 1.f, nullptr, h->yk_tmp, h->yv_tmp, &cnt);
      if (e) return e;
      HIP_TRY(h, hipEventRecord(h->ev1, h->stream));
      HIP_TRY(h, hipEventSynchronize(h->ev1));
      float t = 0.f;
      HIP_TRY(h, hipEventElapsedTime(&t, h->ev0, h->ev1));
      if (r) ms.push_back(t);  // the first run warms up
    }
    std::sort(ms.begin(), ms.end());
    *out = ms[ms.size() / 2];
    return SPMSPV_SUCCESS;
  };
  // a later candidate replaces the current pick only when it is more than
  // 2 % faster, so that run-to-run noise between near-equal candidates does
  // not flip the choice
  auto better = [](float t, float best) { return t < 0.98f * best; };
  const int teams[] = {2, 4, 8, 16, 32, 64};
  float best = 1e30f, t = 0.f;
  int pick = h->team_col;
  // the team and tile candidates are timed on their own launches
  const bool saved_small = h->small_x_path;
  h->small_x_path = false;
  for (int tc : teams) {
    h->team_col = tc;
    if ((rc = time_slot(SPMSPV_K_COL_NAIVE_ATOMIC, &t))) return rc;
    if (better(t, best)) {
      best = t;
      pick = tc;
    }
  }
  h->team_col = pick;
  if (h->csr_row) {
    best = 1e30f;
    pick = h->team_row;
    for (int tr : teams) {
      h->team_row = tr;
      if ((rc = time_slot(SPMSPV_K_SPMV_NAIVE, &t))) return rc;
      if (better(t, best)) {
        best = t;
        pick = tr;
      }
    }
    h->team_row = pick;
    best = 1e30f;
    pick = h->row_vt;
    const int saved_type = h->spmv_type;
    h->spmv_type = 0;
    for (int vt : {5, 7, 11}) {
      h->row_vt = vt;
      h->row_ready = false;
      if ((rc = ensure_row_tiles(h)) || (rc = time_slot(SPMSPV_K_SPMV_LB, &t))) return rc;
      if (better(t, best)) {
        best = t;
        pick = vt;
      }
    }
    h->row_vt = pick;
    h->row_ready = false;
    if ((rc = ensure_row_tiles(h))) return rc;
    // the CSR5 tiles compete with the merge-path tiles for K5/K7
    int type = 0, sigma = h->sigma;
    if (csr5) {
      h->spmv_type = 1;
      for (int sg : {8, 16, 24, 32}) {
        h->sigma = sg;
        if ((rc = time_slot(SPMSPV_K_SPMV_LB, &t))) return rc;
        if (better(t, best)) {
          best = t;
          type = 1;
          sigma = sg;
        }
      }
    } else {
      type = saved_type;
    }
    h->spmv_type = type;
    h->sigma = sigma;
    if (type == 1 && (rc = ensure_csr5(h))) return rc;
  }
  h->small_x_path = saved_small;
  // the bucket slot's number of row bins, among the counts whose layout
  // fits kBinMax bins (a matrix of very many rows may fit none: the slot and
  // its rule then stay off)
  bool bucket_ok = false;
  {
    float bbest = 1e30f;
    int bpick = h->bucket_bits;
    for (int bb = 8; bb <= 10; ++bb) {
      h->bucket_bits = bb;
      if (ensure_bin_layout(h) != SPMSPV_SUCCESS) continue;
      if ((rc = time_slot(SPMSPV_K_COL_BUCKET, &t))) return rc;
      bucket_ok = true;
      if (better(t, bbest)) {
        bbest = t;
        bpick = bb;
      }
    }
    h->bucket_bits = bpick;
    if (!bucket_ok) h->err.clear();
  }
  // Thresholds of the single-workgroup paths, timed on sub-vectors of x
  // whose bin_len lies just under each candidate: the largest candidate at
  // which the single workgroup still beats the multi-launch path is kept.
  // A sub-vector takes k keys spread over x, one from each of k equal
  // intervals at a pseudo-random place in it (written to the handle's
  // internal x buffers), so its columns are spread like those of a sparse x
  // of that size; the first k keys of a dense x would be k adjacent columns,
  // whose contiguous reads flatter the single workgroup. When x itself lives
  // in the internal buffers, prefixes of x are used.
  int* xk = h->x_key;
  float* xv = h->x_val;
  const int xn = h->xnnz;
  const bool spread = xk != h->xk_int && xv != h->xv_int;
  auto subset = [&](int k) -> int {
    if (!spread) return spmspv_set_sparse_x(h, xk, xv, k);
    if (k > 0)
      hipLaunchKernelGGL(k_spread_pick, dim3((k + 255) / 256), dim3(256), 0, h->stream, xk, xv, xn, k, h->xk_int,
                         h->xv_int);
    int r = check_launch(h, "autotune sub-vector");
    return r ? r : spmspv_set_sparse_x(h, h->xk_int, h->xv_int, k);
  };
  auto prefix = [&](long long target, int kmax, int* kout, long long* blout) -> int {
    int lo = 0, hi = kmax, r;  // bin_len(sub-vector of lo keys) <= target
    long long bl = 0;
    while (lo < hi) {
      const int mid = (lo + hi + 1) / 2;
      if ((r = subset(mid)) || (r = spmspv_vec_features(h, &bl, nullptr, nullptr))) return r;
      if (bl <= target) lo = mid;
      else hi = mid - 1;
    }
    if ((r = subset(lo)) || (r = spmspv_vec_features(h, &bl, nullptr, nullptr))) return r;
    *kout = lo;
    *blout = bl;
    return SPMSPV_SUCCESS;
  };
  {
    int cap = 0;
    for (int c : {2048, 4096, 8192, kFuseCap}) {
      int k;
      long long bl;
      if ((rc = prefix(c, std::min(xn, kFuseKeys), &k, &bl))) return rc;
      if (bl < c / 2) break;  // x holds no prefix near this size
      float tf = 0.f, t2 = 0.f, t3 = 0.f;
      h->small_x_path = true;
      h->fuse_cap = kFuseCap;
      if ((rc = time_slot(SPMSPV_K_COL_LB_SORT, &tf))) return rc;
      h->small_x_path = false;
      if ((rc = time_slot(SPMSPV_K_COL_LB_ATOMIC, &t2)) || (rc = time_slot(SPMSPV_K_COL_LB_SORT, &t3))) return rc;
      if (!(tf < std::min(t2, t3))) break;
      cap = c;
    }
    h->small_x_path = true;
    h->fuse_cap = cap > 0 ? cap : h->fuse_cap;
  }
  {
    const bool saved_path = h->small_x_path;
    h->small_x_path = false;
    int cap = 0;
    for (int c : {1024, 2048, kSmallSortCap}) {
      int k;
      long long bl;
      if ((rc = prefix(c, xn, &k, &bl))) return rc;
      if (bl < c / 2) break;
      float t1 = 0.f, tm = 0.f;
      h->small_sort_cap = kSmallSortCap;
      if ((rc = time_slot(SPMSPV_K_COL_LB_SORT, &t1))) return rc;
      h->small_sort_cap = 0;
      if ((rc = time_slot(SPMSPV_K_COL_LB_SORT, &tm))) return rc;
      if (!(t1 < tm)) break;
      cap = c;
    }
    h->small_x_path = saved_path;
    h->small_sort_cap = cap > 0 ? cap : kSmallSortCap;
  }
  // The bin_len range of the bucket rule: on sub-vectors of x whose bin_len
  // lies just under nnz / 4^k (k = 5..0), the bucket slot against the
  // load-balanced column and row slots (2, 3, 5, 7), which the rule
  // replaces; the range spans the sizes where it is more than 2 % faster
  // than all four. Each end is then narrowed by two bisection steps
  // (geometric) between the last size where the bucket slot loses and the
  // first where it wins, and set three quarters of the way to the losing one.
  if (bucket_ok) {
    h->bucket_lo = 0;
    h->bucket_hi = -1;
    // wins at a sub-vector whose bin_len is just under target: 1 yes, 0 no,
    // -1 no such sub-vector; *blout its bin_len
    auto wins = [&](long long target, int* out, long long* blout) -> int {
      int kk;
      long long bl;
      int r;
      if ((r = prefix(target, xn, &kk, &bl))) return r;
      *blout = bl;
      if (bl < target / 2 || bl <= 0) {
        *out = -1;
        return SPMSPV_SUCCESS;
      }
      float tb = 0.f, best = 1e30f;
      if ((r = time_slot(SPMSPV_K_COL_BUCKET, &tb))) return r;
      for (int sl : {SPMSPV_K_COL_LB_ATOMIC, SPMSPV_K_COL_LB_SORT, SPMSPV_K_ROW_LB, SPMSPV_K_SPMV_LB}) {
        float t = 0.f;
        if ((r = time_slot(sl, &t))) return r;
        best = std::min(best, t);
      }
      *out = better(tb, best) ? 1 : 0;
      return SPMSPV_SUCCESS;
    };
    // the edge between a losing size a and a winning size b (either order)
    auto edge = [&](long long a, long long b, long long* e) -> int {
      for (int step = 0; step < 2; ++step) {
        const long long mid = (long long)std::sqrt((double)a * (double)b);
        int w;
        long long bl;
        int r;
        if ((r = wins(mid, &w, &bl))) return r;
        if (w < 0) break;
        if (w) b = bl;
        else a = bl;
      }
      // three quarters of the way (geometrically) towards the losing size:
      // just past the crossover the bucket slot loses by little, while the
      // selector sees a sampled bin_len (a few tens of percent off) for a
      // large x
      *e = (long long)std::exp(0.75 * std::log((double)a) + 0.25 * std::log((double)b));
      return SPMSPV_SUCCESS;
    };
    long long prev_lose = -1, lo = -1, hi = -1, first_win = -1, last_win = -1;
    bool done = false;
    for (int k = 5; k >= 0 && !done; --k) {
      const long long c = std::max<long long>(1, (long long)h->nnz >> (2 * k));
      int w;
      long long bl;
      if ((rc = wins(c, &w, &bl))) return rc;
      if (w < 0) continue;  // x holds no sub-vector near this size
      if (w) {
        if (first_win < 0) {
          first_win = bl;
          if (prev_lose < 0) lo = 1;
          else if ((rc = edge(prev_lose, bl, &lo))) return rc;
        }
        last_win = bl;
      } else {
        if (first_win >= 0) {
          if ((rc = edge(bl, last_win, &hi))) return rc;
          done = true;  // the range ends here
        }
        prev_lose = bl;
      }
    }
    if (first_win >= 0) {
      h->bucket_lo = lo;
      h->bucket_hi = done ? hi : LLONG_MAX;  // still ahead at the largest size: no upper end
    }
  } else {
    h->bucket_lo = 0;
    h->bucket_hi = -1;
  }
  if ((rc = spmspv_set_sparse_x(h, xk, xv, xn))) return rc;
  // captured BFS levels hold the old tiling and thresholds
  bfs_drop_graphs(h);
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  if (report) {
    report[0] = h->team_col;
    report[1] = h->team_row;
    report[2] = h->row_vt;
    report[3] = h->spmv_type;
    report[4] = h->spmv_type == 1 ? csr5_sigma_of(h->sigma, h->m, h->nnz) : 0;
  }
  return SPMSPV_SUCCESS;
}

// ------------------------------------------------------------- utilities
int spmspv_device_count(int* count) {
  hipError_t e = hipGetDeviceCount(count);
  if (e != hipSuccess) {
    *count = 0;
    return set_error(nullptr, SPMSPV_DEVICE_ERROR, hipGetErrorString(e));
  }
  return SPMSPV_SUCCESS;
}
int spmspv_malloc(void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 1));
  return e == hipSuccess ? SPMSPV_SUCCESS : set_error(nullptr, SPMSPV_OUT_OF_MEMORY, hipGetErrorString(e));
}
int spmspv_free(void* p) {
  hipError_t e = hipFree(p);
  return e == hipSuccess ? SPMSPV_SUCCESS : set_error(nullptr, SPMSPV_DEVICE_ERROR, hipGetErrorString(e));
}
int spmspv_memcpy_h2d(void* d, const void* s, size_t bytes) {
  hipError_t e = hipMemcpy(d, s, bytes, hipMemcpyHostToDevice);
  return e == hipSuccess ? SPMSPV_SUCCESS : set_error(nullptr, SPMSPV_DEVICE_ERROR, hipGetErrorString(e));
}
int spmspv_memcpy_d2h(void* d, const void* s, size_t bytes) {
  hipError_t e = hipMemcpy(d, s, bytes, hipMemcpyDeviceToHost);
  return e == hipSuccess ? SPMSPV_SUCCESS : set_error(nullptr, SPMSPV_DEVICE_ERROR, hipGetErrorString(e));
}
int spmspv_memset(void* d, int v, size_t bytes) {
  hipError_t e = hipMemset(d, v, bytes);
  return e == hipSuccess ? SPMSPV_SUCCESS : set_error(nullptr, SPMSPV_DEVICE_ERROR, hipGetErrorString(e));
}
int spmspv_memcpy_d2d(void* d, const void* s, size_t bytes) {
  hipError_t e = hipMemcpy(d, s, bytes, hipMemcpyDeviceToDevice);
  return e == hipSuccess ? SPMSPV_SUCCESS : set_error(nullptr, SPMSPV_DEVICE_ERROR, hipGetErrorString(e));
}
int spmspv_synchronize(void) {
  hipError_t e = hipDeviceSynchronize();
  return e == hipSuccess ? SPMSPV_SUCCESS : set_error(nullptr, SPMSPV_DEVICE_ERROR, hipGetErrorString(e));
}
int spmspv_timer_start(spmspv_handle_t h) {
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  HIP_TRY(h, hipEventRecord(h->ev0, h->stream));
  return SPMSPV_SUCCESS;
}
int spmspv_timer_stop(spmspv_handle_t h, float* ms) {
  if (!h || !ms) return set_error(h, SPMSPV_INVALID_ARGUMENT, "null argument");
  HIP_TRY(h, hipEventRecord(h->ev1, h->stream));
  HIP_TRY(h, hipEventSynchronize(h->ev1));
  HIP_TRY(h, hipEventElapsedTime(ms, h->ev0, h->ev1));
  return SPMSPV_SUCCESS;
}
int spmspv_flush_cache(spmspv_handle_t h, size_t bytes) {
  if (!h) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "null handle");
  if (h->flush_bytes < bytes) {
    if (h->flush_buf) (void)hipFree(h->flush_buf);
    h->flush_buf = nullptr;
    HIP_TRY(h, hipMalloc(&h->flush_buf, bytes));
    h->flush_bytes = bytes;
  }
  launch_fill_f32(h->stream, h->flush_buf, (long long)(bytes / 4), 1.0f);
  return check_launch(h, "flush");
}

}  // extern "C"
