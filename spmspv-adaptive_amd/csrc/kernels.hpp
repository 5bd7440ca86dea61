// Launch interface of the gfx950 kernels. All launches go to the given
// stream and never allocate; sizes that only the device knows (bin_len,
// tile counts) are read by the kernels from the control block.
#pragma once

#include <vector>

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace spmspv {

// Tile geometry (one place, so the arena sizing in the handle agrees).
constexpr int kPrepNT = 256, kPrepIPT = 8, kPrepTile = kPrepNT * kPrepIPT;
// odd items per thread: a lane walking VT consecutive LDS words then hits
// a different bank than its neighbours (stride 7 is coprime with 32 banks)
constexpr int kLbNT = 256, kLbVT = 7, kLbTile = kLbNT * kLbVT;
constexpr int kCmpNT = 256;  // bit-array words per compaction tile (8,192 rows, 8 per thread)
constexpr int kSortNT = 256, kSortIPT = 12, kSortTile = kSortNT * kSortIPT;
constexpr int kDScanNT = 256, kDScanIPT = 8, kDScanTile = kDScanNT * kDScanIPT;
constexpr int kRedNT = 256, kRedIPT = 7, kRedTile = kRedNT * kRedIPT;
constexpr int kRowNT = 256, kRowVT = 7, kRowTile = kRowNT * kRowVT;
// row-merge items per thread the autotuner may pick (odd: conflict-free)
constexpr int kRowVTMin = 5;
constexpr int kSmallSortCap = kSortTile;  // single-workgroup LDS sort + reduce
constexpr int kPersistGrid = 1024;        // workgroups of a persistent launch

// Ticket slots in Ctl (one per persistent kernel kind).
enum Slot { kSlotPrep = 0, kSlotCompact, kSlotHist, kSlotSort, kSlotReduce, kSlotMat, kSlotBfs,
            kSlotBatch, kSlotBucket, kSlotSelect,
            // done[10..11] as one 64-bit word: the compaction's arrivals
            // (high half) and result count (low half)
            kSlotCount = 10 };
constexpr int kSelMaxGrid = 512;  // workgroups of the features-only selector launch

struct MatFeat {  // device-resident matrix features (computeMatFeture_serial)
  double max_row, min_row, avg_row, range, std_row, equality, gini, std_trunc;
  double m, n, nnz;
};

// Device-side switch for work queued before the host knows whether it is
// wanted: a launch whose gate has a word does nothing unless *word == want
// (read as the kernel starts; an earlier launch on the stream wrote it).
struct Gate {
  const int* word = nullptr;
  int want = 0;
};
__device__ __forceinline__ bool gate_shut(const Gate& g) {
  return g.word != nullptr && *g.word != g.want;
}

// --- column path ---------------------------------------------------------
// Column-length gather + exclusive scan + (bin_len, max, min) + merge-path
// splits for the load-balanced gather; when `mlp` is non-null the last
// workgroup also builds the selector features and runs the MLP.
struct SelectorArgs {
  const float* weights;  // NWEIGHTS floats or null
  const MatFeat* mat;
  int mode;              // 0 mlp, 1 tree, 2 fixed
  int fixed;
  long long* stash;      // when set: the selector word of this x (sel_word) is stored here
  // features-only launch for a large x: the selector word may carry the
  // sampled estimate of bin_len (flagged kSelEstimated) instead of the sum
  int estimate;
  // switching-aware choice: with penalty > 0 and prev >= 0, a slot of the
  // previous call's family (column 0-3, row SpMSpV 4-5, SpMV 6-7) whose
  // logit is within penalty of the top one is kept instead of a switch
  int prev;
  float penalty;
  // tuned size rule behind the network: a load-balanced column or row slot
  // (2, 3, 5, 7) becomes the bucket slot (8) when bin_lo <= bin_len <= bin_hi
  // (off by default: bin_hi < bin_lo)
  long long bin_lo = 0, bin_hi = -1;
};
__host__ __device__ inline int slot_family(int k) { return k < 4 ? 0 : (k < 6 ? 1 : 2); }
// The selector's x inputs for a large x (xnnz > kSampleMin) come from
// kSampleKeys evenly spaced keys (positions floor(i * xnnz / kSampleKeys)):
// bin_len is their column lengths' sum scaled by xnnz / kSampleKeys, and
// max_len their longest column. Every launch that runs the selector uses
// this definition, so training and dispatch see the same features.
constexpr int kSampleMin = 1 << 16, kSampleKeys = 1024;
// {bin_len, kernel choice} in one 64-bit word, so that a single store
// publishes both (no ordering fence); never negative, choice 0xff = none.
// Flag kSelEstimated in the choice byte: bin_len is the sampled estimate.
constexpr int kSelEstimated = 0x40;
// Flag kSelDone: the launch that selected also computed the result (the
// single-launch small-x column path); the bin_len field holds y_nnz. Set
// in ctl->kernel_choice too, where it shuts every gated launch.
constexpr int kSelDone = 0x80;
__host__ __device__ inline long long sel_word(long long bin_len, int choice) {
  return (bin_len << 8) | (long long)(choice & 0xff);
}
__host__ __device__ inline long long sel_word_binlen(long long w) { return w >> 8; }
__host__ __device__ inline int sel_word_choice(long long w) {
  const int c = (int)(w & 0xff);
  return c == 0xff ? -1 : (c & 0x3f);
}
__host__ __device__ inline bool sel_word_done(long long w) {
  const int c = (int)(w & 0xff);
  return c != 0xff && (c & kSelDone) != 0;
}
__host__ __device__ inline bool sel_word_estimated(long long w) {
  const int c = (int)(w & 0xff);
  return c != 0xff && (c & kSelEstimated) != 0;
}
// samp (2 * kSampleKeys / kPrepNT long longs, or null): the first tiles'
// partials of a large x's sampled selector inputs, so that the sampling
// overlaps the scan instead of following the last tile's look-back.
void launch_prep(hipStream_t s, const int* colptr, const int* xkey, int ncols, int xnnz,
                 const int* xnnz_dev, int* ptr, int* xbeg, int* split, unsigned long long* state,
                 int* tile_mm, Ctl* ctl, const SelectorArgs* sel, Ep ep, long long* samp = nullptr);
// Features-only selector: ctl->{bin_len, max_len, min_len, xnnz} of x and,
// with sel.mat set, the selector's choice (and the selector word in
// sel.stash). part: 5 * kSelMaxGrid partials. No column scan. With enc set
// (already all +0.0), x's support encoding is scattered into it as well.
void launch_select_feat(hipStream_t s, const int* colptr, const int* xkey, const float* xval,
                        int ncols, int xnnz, long long* part, Ctl* ctl, const SelectorArgs& sel,
                        float* enc, const int* enc_perm = nullptr);
// mode 0: FP32 atomics into acc + marks; mode 1: (row, a*x) at the product
// index; mode 2: the row keys alone (boolean, cval and xval unread).
void launch_naive_col(hipStream_t s, int mode, const int* colptr, const int* crow, const float* cval,
                      const int* xkey, const float* xval, const int* ptr, int ncols, int xnnz,
                      const int* xnnz_dev, int team, float* acc, unsigned char* marks, int* bkey,
                      float* bval, Gate g = {});
// modes as launch_naive_col.
void launch_lb_gather(hipStream_t s, int mode, const int* ptr, const int* xbeg, const float* xval,
                      const int* split, const int* crow, const float* cval, const Ctl* ctl,
                      int xnnz, const int* xnnz_dev, float* acc, unsigned char* marks, int* bkey,
                      float* bval, long long tiles_hint, Gate g = {});
// Two-layer load balance: tiles of exactly kLbTile products; starts holds
// ceil(bin_len / kLbTile) + 1 ints. mode 0: atomics into acc + marks,
// mode 1: bins at the product index.
void launch_tl_gather(hipStream_t s, int mode, const int* ptr, const int* xbeg, const float* xval, int xnnz,
                      int* starts, const int* crow, const float* cval, const Ctl* ctl, float* acc,
                      unsigned char* marks, int* bkey, float* bval, long long tiles_hint, Gate g = {});
// hits (bit array, or mark bytes when marks != null) -> sorted (key, val);
// clears the hits (and acc entries when reset_acc) so the accumulator
// returns to all-zero. With keymap, hit k is reported as key keymap[k]
// (keymap ascending keeps the keys sorted). With sig (may be host-mapped),
// the result count is also written to *sig once every workgroup's stores
// are visible, as launch_signal_int after this launch would.
// clr (nullable): nclr words at a stride of kBinCurStride set to 0 (the
// bucket slot's cursors, once its accumulate kernel has read them).
void launch_compact(hipStream_t s, unsigned* bitmap, unsigned char* marks, int m, float* acc,
                    int reset_acc, int* ykey, float* yval, unsigned long long* state, Ctl* ctl,
                    Ep ep, const int* keymap = nullptr, Gate g = {}, long long* sig = nullptr,
                    unsigned* clr = nullptr, int nclr = 0);
// bitmap of (v != 0) over a dense vector (whole words, no atomics)
void launch_bitmap_from_dense(hipStream_t s, const float* v, int n, unsigned* bits);
void launch_fill_f32(hipStream_t s, float* p, long long n, float v, Gate g = {});
// *dst = *src once every earlier launch on s has finished (dst may be host-mapped)
void launch_signal_int(hipStream_t s, const int* src, long long* dst, Gate g = {});
void launch_fill_i32(hipStream_t s, int* p, long long n, int v);
// dense[key[i]] = val[i]; with `encode`, a zero value is stored as -0.0
// (the support encoding: +0.0 outside the support, nonzero bits inside).
// With perm, entry k lands at dense[perm[k]].
void launch_scatter(hipStream_t s, const int* key, const float* val, int nnz, const int* nnz_dev,
                    int len, float* dense, bool encode = false, Gate g = {}, const int* perm = nullptr);
// enc[c] = bit c of bits ? (x[c] == 0 ? -0.0 : x[c]) : +0.0 (at enc[perm[c]]
// with perm)
void launch_encode_x(hipStream_t s, const float* x, const unsigned* bits, int n, float* enc,
                     const int* perm = nullptr);
void launch_set_bits(hipStream_t s, const int* key, int nnz, const int* nnz_dev, int len,
                     unsigned* bits);
void launch_clear_keys(hipStream_t s, const int* key, int nnz, float* dense, unsigned* bits);

// Single-launch small-x column path: x of at most kFuseKeys keys whose
// products (bin_len) fit kFuseCap run scan, gather, LDS sort, reduce and
// the sorted output in one workgroup (k_col_fused, sort.hip). It also
// writes the column scan (ptr, xbeg, split, ctl) like launch_prep, so a
// call that does not fit goes on with the usual launches (gated on
// ctl->kernel_choice, which gets kSelDone when the launch finished the
// call). In auto mode (sel.mode != 2) a call that fits is reported as
// kFusedSlot, the sort-based slot whose result it computes. Returns false
// (nothing launched) when xnnz > kFuseKeys.
constexpr int kFuseNT = 1024, kFuseKeys = 4096, kFuseCap = kFuseNT * 11;
constexpr int kFusedSlot = 1;  // SPMSPV_K_COL_NAIVE_SORT
// cap (<= kFuseCap): largest bin_len that takes the single launch (the
// autotuner may lower it).
bool launch_col_fused(hipStream_t s, const int* colptr, const int* crow, const float* cval, const int* xkey,
                      const float* xval, int ncols, int xnnz, int m, int* ptr, int* xbeg, int* split, Ctl* ctl,
                      const SelectorArgs& sel, int* ykey, float* yval, long long* count_sig, int cap);

// --- sort path -----------------------------------------------------------
// Sorts (bkey, bval)[bin_len] by key (< m) with an LSD radix sort
// (onesweep), then reduces equal keys into (ykey, yval) and ctl->ynnz.
// bin_len <= kSmallSortCap takes a single-workgroup LDS path.
struct SortBufs {
  unsigned* keyA;
  float* valA;
  unsigned* keyB;
  float* valB;
  unsigned* dcount;     // 256 per sort tile: per-tile digit counts, then their scan
  unsigned long long* dstate;     // look-back words of the digit-count scan
  unsigned long long* red_state;  // per reduce tile
  float* carry_val;
  int* carry_idx;
};
// n_hint: item count when the host knows it (exact grids), else -1 and the
// grids cover n_cap. ep: first of the 5 epochs the sort + reduce use.
// small_cap (<= kSmallSortCap): largest item count that takes the
// single-workgroup path (the autotuner may lower it).
void launch_sort_reduce(hipStream_t s, const SortBufs& b, int m, Ctl* ctl, int* ykey, float* yval,
                        long long binlen_hint, long long n_cap, Ep ep, Gate g = {},
                        int small_cap = kSmallSortCap);
// Sorts (keyA, valA)[*n_dev] by key < range; returns 0 if the result is in
// A, 1 if in B. With valA and valB null it sorts the keys alone.
int launch_radix_sort(hipStream_t s, const SortBufs& b, int range, const long long* n_dev, Ctl* ctl,
                      long long n_hint, long long n_cap, bool skip_small, Ep ep, Gate g = {},
                      int small_cap = kSmallSortCap);

// Bucket slot: row bins, ranges of consecutive rows with about equal
// numbers of the matrix's nonzeros (about nnz / 2^bucket_bits each) and at
// most 2^kBinRowBits rows, at most kBinMax of them. Bin b owns the product
// slots [base[b], base[b+1]) of bkey/bval (nnz capacity): the nonzeros of
// its rows, an upper bound on its products. bin_layout computes the layout
// on the host from the row pointer (returns the bin count, or -1 when more
// than kBinMax bins would be needed). Per call the products are scattered
// into their bins and summed per bin in LDS windows into acc (marks set);
// the caller's compaction gives the sorted result and clears acc, marks
// and the cursors (one per 128-byte line; its clr argument). Needs k_prep's
// scan and splits.
constexpr int kBinRowBits = 12, kBinMax = 6144, kBinChunk = 16384, kBinCurStride = 32, kBinTbl = 2048;
struct BinLayout {
  std::vector<int> start;  // first row of each bin, then m
  std::vector<int> base;   // first product slot of each bin, then nnz
  std::vector<int> tbl;    // bin of the first row of each block of 2^tbl_shift rows
  int tbl_shift = 0;
  int max_rows = 1;        // rows of the largest bin
};
struct BinWs {
  unsigned* cursor;  // kBinMax * kBinCurStride: products reserved per bin (word b * kBinCurStride)
  int* base;         // kBinMax + 1
  int* start;        // kBinMax + 1
  int* tbl;          // kBinTbl
  int ntbl = 1, tbl_shift = 0;
  int max_rows = 1;
};
int bin_layout(const std::vector<int>& row_ptr, int m, int bucket_bits, BinLayout* out);
// cnt[crow[i]] += 1 for every nonzero (nonzeros per row of a CSC matrix)
void launch_row_counts(hipStream_t s, const int* crow, long long nnz, int* cnt);
void launch_bucket(hipStream_t s, const int* ptr, const int* xbeg, const float* xval, const int* split,
                   const int* crow, const float* cval, const Ctl* ctl, int xnnz, const BinWs& ws, int nbins,
                   unsigned* bkey, float* bval, float* acc, unsigned char* marks, long long tiles_hint,
                   long long binlen_hint, long long binlen_cap, Gate g = {});

// --- row path ------------------------------------------------------------
// --- CSR5 for wave64 (spmv_type 1) ----------------------------------------
struct Csr5 {
  int S = 0;            // items per lane (sigma)
  int n = 0;            // columns
  // new index of every column (most nonzeros first: the gathers of the hot
  // columns then share cache lines), or null when the columns keep their
  // order; ccol holds the new indices and x must be given in that order
  int* cperm = nullptr;
  const int* row_ptr = nullptr;  // the CSR it was built from
  int ntiles = 0;       // tiles of 64 * S items
  int n_ne = 0, n_empty = 0;
  int* ccol = nullptr;  // transposed column indices, ntiles * 64 * S
  float* cval = nullptr;
  unsigned* flags = nullptr;  // one row-start marker word per lane and tile
  unsigned* tinfo = nullptr;  // first non-empty row index | continuation bit
  int* rows_ne = nullptr;     // non-empty rows, ascending
  int* rows_empty = nullptr;  // empty rows, ascending
  int* rank = nullptr;        // rank[r] = non-empty rows before r
  float* y_ne = nullptr;      // per non-empty row results when some rows are empty
  float* carry_val = nullptr;
  int* carry_row = nullptr;
  int* carry_hit = nullptr;
};
// sigma (<= 0: from nonzeros per row) rounded up to an instantiated width
int csr5_sigma_of(int sigma, int m, long long nnz);
size_t csr5_bytes(int m, int n, long long nnz, int S);
// What the column ordering of a CSR5 build may use: the sort buffers (n <=
// their capacity), the control block, a device count word, fresh epochs,
// and the CSC column pointer when there is one (else columns are counted).
struct ColumnOrderArgs {
  SortBufs sb;
  long long sb_cap;
  Ctl* ctl;
  long long* n_dev;
  Ep ep;  // 12 consecutive epochs
  const int* csc_colptr;
};
// cperm[c] = rank of column c by nonzero count (descending, ties by index).
// Returns false (no order) when the sort buffers cannot hold n keys.
bool launch_column_order(hipStream_t s, const ColumnOrderArgs& a, const int* col, int nnz, int n,
                         int* cperm);
// Builds the CSR5 copy of a CSR into `mem` (csr5_bytes); with `order`, the
// columns are renumbered by nonzero count and every row's items put in that
// order (the caller's CSR is untouched); synchronises s.
hipError_t csr5_build(hipStream_t s, const int* row_ptr, const int* col, const float* val, int m,
                      int n, int nnz, int S, void* mem, int* h_pin, const ColumnOrderArgs* order,
                      Csr5* out);
// Same gather modes as launch_row_merge (x == nullptr: bit array only, for
// a copy without a column order). x is indexed by the copy's column order.
// Without empty rows the results go to y (one mark byte per row with hits).
// With empty rows they go to the non-empty-row space: vals_out = y_ne,
// marks indexed by non-empty row, keymap_out = rows_ne (for the
// compaction), and y is filled only when y_dense_rows (a dense output).
// n_out = entries of that space (m or n_ne).
// ev_begin / ev_end (nullable) are recorded around the tile kernel alone.
// clear (nullable, c.n floats, 16-byte aligned, not x) is set to +0.0 by the
// tile kernel as it runs.
void launch_csr5(hipStream_t s, const Csr5& c, bool masked, bool hits, const float* x,
                 const unsigned* xbits, int m, int nnz, float* y, bool y_dense_rows,
                 unsigned char* marks, float** vals_out, const int** keymap_out, int* n_out,
                 hipEvent_t ev_begin = nullptr, hipEvent_t ev_end = nullptr, Gate g = {},
                 float* clear = nullptr);

// Workgroups of `block` threads of `kernel` resident on the whole device
// at once (occupancy x CUs), cached per kernel.
int resident_blocks(const void* kernel, int block);
void launch_row_split(hipStream_t s, const int* row_ptr, int m, int nnz, int* split, int ntiles,
                      int tile);
// vt in {5, 7, 11}: merged items per thread (tile = 256 * vt).
// x == nullptr: the boolean form, x given by its bit array xbits alone:
// hit bits into ybits, no values read (val may be null), y not written (may
// be null). Otherwise, with masked or hits, x must be
// the support-encoded vector (launch_encode_x / launch_scatter with
// encode); without either it is plain dense x.
void launch_row_merge(hipStream_t s, bool masked, bool hits, const int* row_ptr, const int* col,
                      const float* val, const float* x, const unsigned* xbits, int m, int nnz,
                      const int* split, int ntiles, int vt, float* y, unsigned* ybits,
                      float* carry_val, int* carry_row, int* carry_hit,
                      hipEvent_t ev_begin = nullptr, hipEvent_t ev_end = nullptr, Gate g = {});
// val == nullptr and x == nullptr: the boolean form (hit bits into ybits
// only).
void launch_row_naive(hipStream_t s, bool masked, bool hits, const int* row_ptr, const int* col,
                      const float* val, const float* x, const unsigned* xbits, int m, int team,
                      float* y, unsigned* ybits);

// --- CSC construction ----------------------------------------------------
hipError_t launch_csr_to_csc_sorted(hipStream_t s, const int* row_ptr, const int* col, const float* val,
                              int m, int n, int nnz, int* colptr, int* crow, float* cval,
                              const SortBufs& b, long long* n_dev, Ctl* ctl, Ep ep);

// --- features ------------------------------------------------------------
void launch_mat_features(hipStream_t s, const int* row_ptr, int m, int n, int nnz, int* hist,
                         double* partials, MatFeat* out, Ctl* ctl);

}  // namespace spmspv
