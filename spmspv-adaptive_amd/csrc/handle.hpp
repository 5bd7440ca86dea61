// The handle behind the C-ABI: the matrix (borrowed or owned), the current
// x, and one device arena carved at set-up time so that no call after it
// allocates (the reference's allocPreBuffer arena, spmspv.h:1722-1763, plus
// the hidden per-call allocations it made, here sized up front).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/spmspv_amd.h"
#include "kernels.hpp"

struct BfsState;
struct PrState;
struct MsState;

struct spmspv_handle_s {
  int m = 0, n = 0, nnz = 0;
  hipStream_t stream = nullptr;
  std::string err;

  // matrix
  int* csr_row = nullptr;
  int* csr_col = nullptr;
  float* csr_val = nullptr;
  int* csc_colptr = nullptr;
  int* csc_row = nullptr;
  float* csc_val = nullptr;
  bool csc_owned = false;
  bool csc_stale = false;  // handle-owned CSC transposed from an earlier CSR
  void* csc_alloc = nullptr;

  // x
  int* x_key = nullptr;
  float* x_val = nullptr;
  int xnnz = 0;
  float* x_dense_user = nullptr;
  unsigned* bits_user = nullptr;
  int num_ints_user = 0;
  int vec_type = -1;
  bool sparse_valid = false;  // x_key/x_val describe the current x
  bool dense_valid = false;   // dense x holds the current x
  bool bits_valid = false;    // bit array holds the current x's support
  bool prep_valid = false;    // ptr/xbeg/split/bin_len computed for the current x
  bool enc_valid = false;     // x_enc holds the support encoding of the current x
  bool enc5_valid = false;    // x_enc5b[enc5_cur] holds it in the CSR5 column order
  long long binlen_host = -1; // bin_len read back for the current x (or -1)
  int spmv_type = 0;          // 0 merge (Hola role), 1 CSR5
  int sigma = -1;

  // arena
  bool reserved = false;
  void* arena = nullptr;
  size_t arena_bytes = 0;
  spmspv::Ctl* ctl = nullptr;
  long long* sel_part = nullptr;  // partials of the features-only selector launch
  spmspv::MatFeat* mat = nullptr;
  double* mat_partials = nullptr;
  int* ptr = nullptr;        // n+1 (x scan); doubles as the degree histogram
  int* xbeg = nullptr;       // n
  int* split = nullptr;      // LB tiles + 2
  int* tl_starts = nullptr;  // LB tiles + 2: first x entry of each two-layer tile
  int* tile_mm = nullptr;    // 2 per prep tile
  long long* prep_samp = nullptr;  // the first prep tiles' sampled selector partials
  unsigned long long* scan_state = nullptr;
  size_t scan_words = 0, sort_words = 0, red_words = 0;  // look-back state sizes
  float* acc = nullptr;      // m, all-zero between calls
  unsigned* bitmap = nullptr;  // ceil(m/32), all-zero between calls
  unsigned char* marks = nullptr;  // one byte per row (padded to 32), all-zero between calls
  spmspv::SortBufs sb{};
  spmspv::BinWs bin_ws{};  // bucket slot's per-bin cursors, bases and done counter
  int bin_base_bits = -1;  // bucket_bits the bin layout was built for (-1: none)
  int bin_nbins = 0;       // bins of that layout
  int* yk_tmp = nullptr;     // m
  float* yv_tmp = nullptr;   // m
  float* x_dense_int = nullptr;  // n
  float* x_enc = nullptr;        // n: +0 outside x's support, nonzero bits inside
  // n each: the same in the CSR5 copy's column order, two buffers. The
  // tile kernel reading one clears the other (its x is no longer needed),
  // so the next x is scattered into a buffer that is already all +0.0
  float* x_enc5b[2] = {nullptr, nullptr};
  int enc5_cur = 0;                      // the buffer of the current x
  bool enc5_clean[2] = {false, false};   // buffer known to be all +0.0
  int enc5_clear_pending = -1;           // cleared by a gated tile kernel if it runs
  unsigned* bits_int = nullptr;  // ceil(n/32)
  int* xk_int = nullptr;     // n (x converted from dense)
  float* xv_int = nullptr;   // n
  int* row_split = nullptr;
  int n_row_tiles = 0;
  float* row_carry_val = nullptr;
  int* row_carry_row = nullptr;
  int* row_carry_hit = nullptr;
  long long* n_dev = nullptr;
  float* weights_dev = nullptr;
  float* bfs_weights_dev = nullptr;  // the BFS levels' own selector weights
  float* flush_buf = nullptr;
  size_t flush_bytes = 0;
  bool mat_ready = false;
  bool row_ready = false;

  // selector
  bool have_weights = false;
  bool have_bfs_weights = false;
  int sel_mode = 1;
  int sel_fixed = 0;
  int team_col = 16;
  int team_row = 8;
  int row_vt = 7;  // row-merge items per thread (autotuned)
  int bucket_bits = 9;  // bucket slot: about 2^bucket_bits row bins of equal nonzeros (autotuned, or set by csc_bucket)
  // bin_len range in which auto mode runs the bucket slot in place of a
  // load-balanced column or row slot (autotuned; off while hi < lo)
  long long bucket_lo = 0, bucket_hi = -1;

  // host mirrors
  int* h_pin = nullptr;       // small pinned readback area (64 ints)
  // Host-coherent words the device writes results into directly (the host
  // marks a word kSigPending before the launch and polls it): no copy
  // kernel and no stream wait between the selector and the chosen kernel.
  long long* sig = nullptr;
  long long* sig_dev = nullptr;  // the same words, device address
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // events around the slot's main kernel alone (spmspv_time_main_kernel)
  hipEvent_t evk0 = nullptr, evk1 = nullptr;
  bool time_main = false;
  bool main_timed = false;
  // set by the caller around run_kernel: the slot's final compaction, where
  // one ends the call, writes the result count to signal word 3 itself
  // (count_sent then tells that it did)
  bool count_to_host = false;
  bool count_sent = false;
  long long launches = 0;
  unsigned epoch_host = 1;   // next host-side look-back epoch (bit 23 clear)
  long long dev_levels = 0;  // device-epoch sequences run (captured BFS levels)

  BfsState* bfs = nullptr;
  PrState* pr = nullptr;  // PageRank driver buffers
  MsState* ms = nullptr;  // multi-source BFS buffers
  void* csr5_mem = nullptr;          // CSR5 copy of the CSR (spmv_type 1)
  size_t csr5_bytes = 0;
  bool csr5_ready = false;
  spmspv::Csr5 csr5{};
  long long* batch_stash = nullptr;  // selector word (sel_word) per vector of a batch
  // speculative dispatch (spmspv_run_auto): the gate launches carry while
  // the selector's choice is still on its way, and per x-size class
  // (log2 of xnnz) the last choice + 1 (0: none yet)
  spmspv::Gate gate{};
  bool speculate = true;
  signed char spec_last[33] = {};
  int prev_choice = -1;        // slot of the last selector-dispatched call
  float switch_penalty = 0.f;  // switching-aware choice (0: off)
  bool spec_fused[33] = {};  // the last call of the size class ended in the single-launch path
  bool small_x_path = true;  // column slots on a small x try the single-launch path first
  int fuse_cap = spmspv::kFuseCap;             // largest bin_len of that path (autotuned)
  int small_sort_cap = spmspv::kSmallSortCap;  // largest single-workgroup sort (autotuned)
  long long spec_hits = 0, spec_misses = 0;
  int batch_cap = 0;
  // batch stash reserved by alloc_pre_buffer; a larger batch grows it
  static constexpr int kBatchReserve = 4096;

  float* dense_x() { return x_dense_user ? x_dense_user : x_dense_int; }
  unsigned* bits() { return bits_user ? bits_user : bits_int; }
};

// Returns SPMSPV_DEVICE_ERROR (with the failing call recorded) when a HIP
// runtime call fails.
#define HIP_TRY(h, expr)                                                                     \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return set_error(h, SPMSPV_DEVICE_ERROR, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr long long kSigPending = -0x5A5A5A5A5A5ALL;
// host-coherent signal words: 1 selector word, 3 result count, 16.. a BFS
// level's frontier size and choice (16) and its selector features (17..24)
constexpr int kSigWords = 32, kSigBfs = 16;

namespace spmspv {
int set_error(spmspv_handle_t h, int code, const std::string& msg);
int ensure_arena(spmspv_handle_t h);
int ensure_row_tiles(spmspv_handle_t h);
int ensure_mat_features(spmspv_handle_t h);
// Builds the CSR5 copy (sigma from set_sigma or nonzeros per row).
int ensure_csr5(spmspv_handle_t h);
// Builds the handle-owned CSC from the CSR when no CSC was given or the
// CSR changed since it was built.
int need_csc(spmspv_handle_t h);
// Allocates the BFS buffers once per handle (square matrices).
int bfs_alloc(spmspv_handle_t h);
// Allocates / frees the PageRank driver's buffers (square matrices).
int pr_alloc(spmspv_handle_t h);
void pr_release(spmspv_handle_t h);
// Allocates / frees the multi-source BFS buffers (square matrices).
int ms_alloc(spmspv_handle_t h);
void ms_release(spmspv_handle_t h);
// Drops the captured BFS level graphs (they hold matrix and arena pointers).
void bfs_drop_graphs(spmspv_handle_t h);
int run_prep(spmspv_handle_t h, bool with_split, const SelectorArgs* sel);
int run_kernel(spmspv_handle_t h, int kid, float* y_dense, int* y_key, float* y_val, bool want_nnz);
int readback_ynnz(spmspv_handle_t h, int* y_nnz);
// Waits for the device to write signal word idx (polls, then blocks on the
// stream for long kernels).
int wait_signal(spmspv_handle_t h, int idx, long long* out);
// Reserves k consecutive host epochs for the look-back kernels of one call.
Ep next_ep(spmspv_handle_t h, unsigned k);
// Reads bin_len of the current x back to the host (one synchronisation).
int ensure_binlen_host(spmspv_handle_t h);
// Builds the bucket slot's bin layout for the current matrix and bucket_bits.
int ensure_bin_layout(spmspv_handle_t h);
}  // namespace spmspv
