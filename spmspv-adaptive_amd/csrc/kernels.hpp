// Launch interface of the gfx950 kernels. All launches go to the given
// stream and never allocate; sizes that only the device knows (bin_len,
// tile counts) are read by the kernels from the control block.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace spmspv {

// Tile geometry (one place, so the arena sizing in the handle agrees).
constexpr int kPrepNT = 256, kPrepIPT = 8, kPrepTile = kPrepNT * kPrepIPT;
constexpr int kLbNT = 256, kLbVT = 8, kLbTile = kLbNT * kLbVT;
constexpr int kCmpNT = 256, kCmpWPT = 8, kCmpTile = kCmpNT * kCmpWPT;  // bitmap words
constexpr int kSortNT = 256, kSortIPT = 12, kSortTile = kSortNT * kSortIPT;
constexpr int kRedNT = 256, kRedIPT = 8, kRedTile = kRedNT * kRedIPT;
constexpr int kRowNT = 256, kRowVT = 8, kRowTile = kRowNT * kRowVT;
constexpr int kSmallSortCap = kSortTile;  // single-workgroup LDS sort + reduce
constexpr int kPersistGrid = 1024;        // workgroups of a persistent launch

// Ticket slots in Ctl (one per persistent kernel kind).
enum Slot { kSlotPrep = 0, kSlotCompact, kSlotHist, kSlotSort, kSlotReduce, kSlotMat, kSlotBfs,
            kSlotBatch, kSlotBucket };

struct MatFeat {  // device-resident matrix features (computeMatFeture_serial)
  double max_row, min_row, avg_row, range, std_row, equality, gini, std_trunc;
  double m, n, nnz;
};

// --- column path ---------------------------------------------------------
// Column-length gather + exclusive scan + (bin_len, max, min) + merge-path
// splits for the load-balanced gather; when `mlp` is non-null the last
// workgroup also builds the selector features and runs the MLP.
struct SelectorArgs {
  const float* weights;  // NWEIGHTS floats or null
  const MatFeat* mat;
  int mode;              // 0 mlp, 1 tree, 2 fixed
  int fixed;
};
void launch_prep(hipStream_t s, const int* colptr, const int* xkey, int ncols, int xnnz,
                 const int* xnnz_dev,
                 int* ptr, int* xbeg, int* split, unsigned long long* state, int* tile_mm, Ctl* ctl,
                 const SelectorArgs* sel, int grid_hint);
void launch_naive_col(hipStream_t s, int mode, const int* colptr, const int* crow, const float* cval,
                      const int* xkey, const float* xval, const int* ptr, int ncols, int xnnz,
                      const int* xnnz_dev, int team, float* acc, unsigned* bitmap, int* bkey,
                      float* bval);
void launch_lb_gather(hipStream_t s, int mode, const int* ptr, const int* xbeg, const float* xval,
                      const int* split, const int* crow, const float* cval, const Ctl* ctl,
                      int xnnz, const int* xnnz_dev, float* acc, unsigned* bitmap, int* bkey,
                      float* bval, long long tiles_hint);
// bitmap -> sorted (key, val); resets the bitmap words (and acc entries when
// reset_acc) so the accumulator returns to all-zero.
void launch_compact(hipStream_t s, unsigned* bitmap, int m, float* acc, int reset_acc, int* ykey,
                    float* yval, unsigned long long* state, Ctl* ctl);
// bitmap of (v != 0) over a dense vector (whole words, no atomics)
void launch_bitmap_from_dense(hipStream_t s, const float* v, int n, unsigned* bits);
void launch_fill_f32(hipStream_t s, float* p, long long n, float v);
void launch_fill_i32(hipStream_t s, int* p, long long n, int v);
void launch_scatter(hipStream_t s, const int* key, const float* val, int nnz, const int* nnz_dev,
                    int len, float* dense);
void launch_set_bits(hipStream_t s, const int* key, int nnz, const int* nnz_dev, int len,
                     unsigned* bits);
void launch_clear_keys(hipStream_t s, const int* key, int nnz, float* dense, unsigned* bits);

// --- sort path -----------------------------------------------------------
// Sorts (bkey, bval)[bin_len] by key (< m) with an LSD radix sort
// (onesweep), then reduces equal keys into (ykey, yval) and ctl->ynnz.
// bin_len <= kSmallSortCap takes a single-workgroup LDS path.
struct SortBufs {
  unsigned* keyA;
  float* valA;
  unsigned* keyB;
  float* valB;
  unsigned* hist;       // 4*256 (zero between calls)
  unsigned* digoff;     // 4*256
  unsigned long long* state;      // per (tile, digit)
  unsigned long long* red_state;  // per reduce tile
  float* carry_val;
  int* carry_idx;
};
void launch_sort_reduce(hipStream_t s, const SortBufs& b, int m, Ctl* ctl, int* ykey, float* yval,
                        long long binlen_hint);
// Sorts (keyA, valA)[*n_dev] by key < range; returns 0 if the result is in
// A, 1 if in B.
int launch_radix_sort(hipStream_t s, const SortBufs& b, int range, const long long* n_dev, Ctl* ctl,
                      long long n_hint, bool skip_small);

// --- row path ------------------------------------------------------------
void launch_row_split(hipStream_t s, const int* row_ptr, int m, int nnz, int* split, int ntiles);
void launch_row_merge(hipStream_t s, bool masked, bool hits, const int* row_ptr, const int* col,
                      const float* val, const float* x, const unsigned* xbits, int m, int nnz,
                      const int* split, int ntiles, float* y, unsigned* ybits, float* carry_val,
                      int* carry_row, int* carry_hit);
void launch_row_naive(hipStream_t s, bool masked, bool hits, const int* row_ptr, const int* col,
                      const float* val, const float* x, const unsigned* xbits, int m, int team,
                      float* y, unsigned* ybits);

// --- CSC construction ----------------------------------------------------
void launch_csr_to_csc_sorted(hipStream_t s, const int* row_ptr, const int* col, const float* val,
                              int m, int n, int nnz, int* colptr, int* crow, float* cval,
                              const SortBufs& b, long long* n_dev, Ctl* ctl);

// --- features ------------------------------------------------------------
void launch_mat_features(hipStream_t s, const int* row_ptr, int m, int n, int nnz, int* hist,
                         double* partials, MatFeat* out, Ctl* ctl);

}  // namespace spmspv
