/*
 * CPU oracle for the adaptive SpMV/SpMSpV hot path.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or the timed CPU baseline. The product library never links it.
 *
 * Every function restates the reference's CPU algorithm from its text; each
 * cites the reference file:line it follows. Paths are relative to the
 * reference root; `la/` abbreviates `hice-spmspv/hice/la/`.
 *
 * Conventions: indices int32, values FP32 on input. The parity oracle
 * accumulates in FP64 and reports, per output row, the bound
 * sum_j |a_ij * x_j| so that FP32 results can be checked against
 * |y - y64| <= 1e-5 * bound.
 */
#ifndef SPMSPV_ORACLE_H_
#define SPMSPV_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* CSR -> CSC transpose with rows ascending inside every column (the
 * semantics of cusparseScsr2csc called by SpmspvHandle::ToCSC,
 * la/include/spmspv/csc-spmspv/spmspv.h:383-419). Stable counting sort. */
void oracle_csr_to_csc(int m, int n, int nnz,
                       const int* row_ptr, const int* col, const float* val,
                       int* col_ptr, int* row_idx, float* cval);

/* Reference-faithful FP32 row-ordered SpMV: serialspmv,
 * la/test/test_application.cu:107-116 (also generate_model_data.cu:401-410). */
void oracle_serialspmv_f32(int m, const int* row_ptr, const int* col,
                           const float* val, const float* x_dense,
                           float alpha, float* y);

/* Reference-faithful FP32 column-driven SpMSpV: SpmSpvCSCHost,
 * la/test/test_sort_spmspv.cu:329-377, with the `int x_val` truncation of
 * :340 removed. Accumulates into y_dense (size m, overwritten) and emits the
 * rows whose |y| >= FLT_EPSILON (the reference's filter, :356-372) in
 * ascending order. Returns y_nnz. */
int oracle_spmspv_csc_host_f32(int m, const int* col_ptr, const int* row_idx,
                               const float* cval, int xnnz, const int* x_key,
                               const float* x_val, float* y_dense,
                               int* y_key, float* y_val);

/* The build's all-core CPU port, NOT the reference (whose CPU path is
 * single-threaded): the SpmSpvCSCHost loop above with the rows split into
 * one contiguous range per OpenMP thread (each thread walks every column of
 * x's support and adds its own rows' products in x's order, so every row's
 * FP32 sum equals the serial one bit for bit), then the same filter.
 * Returns y_nnz. */
int oracle_spmspv_csc_host_f32_omp(int m, const int* col_ptr, const int* row_idx,
                                   const float* cval, int xnnz, const int* x_key,
                                   const float* x_val, float* y_dense,
                                   int* y_key, float* y_val, int threads);

/* FP64 SpMSpV the results are checked against, column-driven (the algorithm of SpmSpvCSCHost,
 * la/test/test_sort_spmspv.cu:342-351) with FP64 accumulation and the
 * structural support: a row is in the output iff some selected column has a
 * stored entry in it, whatever the value (also x values of 0).
 * y64/bound are dense scratch of size m (overwritten). Emits sorted keys,
 * FP64 values and per-row bounds; returns y_nnz. */
int oracle_spmspv_f64(int m, const int* col_ptr, const int* row_idx,
                      const float* cval, int xnnz, const int* x_key,
                      const float* x_val, double* y64, double* bound,
                      int* y_key, double* y_val, double* y_bound);

/* FP64 SpMV with dense x the results are checked against (row-driven, the loop of serialspmv,
 * la/test/test_application.cu:107-116, in FP64). Writes y[m], bound[m]. */
void oracle_spmv_f64(int m, const int* row_ptr, const int* col,
                     const float* val, const float* x_dense,
                     double* y, double* bound);

/* Level-synchronous BFS (the reference has none: it replays recorded frontiers). Textbook
 * queue BFS over the CSC pattern: the neighbours of u are the rows of column
 * u. levels[n] gets -1 for unreached vertices. Returns the number of levels
 * (max level + 1). queue is scratch of size n. */
int oracle_bfs(int n, const int* col_ptr, const int* row_idx, int src,
               int* levels, int* queue);

/* Per-matrix row-degree features, computeMatFeture_serial,
 * la/include/spmspv/csc-spmspv/spmspv.h:2274-2334 (same formulas as
 * la/test/generate_feature.cu:209-290). FP64 throughout (the reference mixes
 * float and double); the gini keeps the reference's index ranges, in which
 * S2 omits the largest degree (:2318-2324). out[7] = {max, min, avg,
 * range=(max-min)/n, std, edge_equality, gini}. deg_scratch size m. */
void oracle_mat_features(int m, int n, int nnz, const int* row_ptr,
                         int* deg_scratch, double out[7]);

/* stdRow as extractMatFeature computes it: the mean is truncated to int
 * before the deviations are summed (spmspv.h:929, :429). */
double oracle_std_row_trunc(int m, int nnz, const int* row_ptr);

/* computeVecFeture_serial, spmspv.h:2337-2361: bin_len = sum of column
 * lengths over x's keys, max and min column length. out = {bin_len, max, min}.
 * With xnnz == 0 it returns {0, 0, INT_MAX} like the reference. */
void oracle_vec_features(const int* col_ptr, int xnnz, const int* x_key,
                         int64_t out[3]);

/* Byte models GM1/GM2/GM3, spmspv.h:2724-2753 (sizeof(int)=sizeof(float)=4).
 * out[3] = {GM1, GM2, GM3}. */
void oracle_gm(int64_t m, int64_t nnz, int64_t xnnz, int64_t bin_len,
               double out[3]);

/* Column-length gather + exclusive scan (ScanFuseDriver,
 * la/include/spmspv/csc-spmspv/detail/device/csc_sort_scan_fuse.cuh:8-80):
 * ptr[0]=0, ptr[i+1]=ptr[i]+len(x_key[i]). Returns the total (bin_len). */
int64_t oracle_scan_col_lengths(const int* col_ptr, int xnnz, const int* x_key,
                                int64_t* ptr);

/* Residual-push PageRank in FP64 (the reference has no PageRank; its
 * harnesses replay frontiers recorded from one, la/test/test_application.cu:
 * 206-210, 252-256). Column u of the CSC holds u's out-edges with weights
 * a_vu; w_u = sum_v a_vu. p = 0, r = (1-d)/n, every vertex active; per
 * iteration, for the active set S in ascending order: p[u] += r[u],
 * x_u = r[u]/w_u (0 when w_u = 0), r[u] = 0; y = A x column by column;
 * r[v] += d*y[v] for every row v that a column of S reaches, and v joins
 * the next S (ascending) if r[v] > eps*(1-d)/n. At the end p += r.
 * rank[n] out; sizes (nullable, max_iter ints) = |S| per iteration.
 * Returns the number of iterations run. */
int oracle_pagerank_f64(int n, const int* col_ptr, const int* row_idx, const float* cval, double damping,
                        double eps, int max_iter, double* rank, int* sizes);

/* FP64 MLP 16->H1->H2->C with ReLU hidden layers and softmax cross-entropy
 * (the selector; the reference's selectors are sklearn trees,
 * model/1-test_dtree.py:170-196, so this has no reference counterpart).
 * Parameters packed as W1[H1*F] b1[H1] W2[H2*H1] b2[H2] W3[C*H2] b3[C]
 * (row-major, out x in). Returns the mean loss over the batch and, when
 * grad != NULL, writes d(mean loss)/d(params) in the same packing. Labels
 * may be given as hard labels (soft == NULL) or soft targets soft[B*C]. */
double oracle_mlp_loss_grad(int F, int H1, int H2, int C, int B,
                            const double* params, const double* X,
                            const int* labels, const double* soft,
                            double* grad);

/* Forward pass only: logits[B*C]. */
void oracle_mlp_forward(int F, int H1, int H2, int C, int B,
                        const double* params, const double* X, double* logits);

#ifdef __cplusplus
}
#endif
#endif
