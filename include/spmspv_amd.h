/*
 * spmspv_amd.h -- C-ABI of the MI355X (gfx950) adaptive SpMV/SpMSpV engine.
 *
 * Drop-in boundary for the reference's operator surface, the class
 * SpmspvHandle<int, unsigned int, float>
 * (hice-spmspv/hice/la/include/spmspv/csc-spmspv/spmspv.h:88-344; below
 * `spmspv.h` abbreviates that path). Every entry point names the reference
 * method it replaces. Plain pointers and sizes only: device pointers are
 * marked `d_`, host pointers are not.
 *
 * Contract (the reference's behaviour, with its defects fixed):
 *  - The Input and set_ calls borrow the caller's device arrays; the handle only
 *    frees what it allocated itself (spmspv.h:2863-2876 freed borrowed CSC).
 *  - Every kernel sets *y_nnz when y_nnz is non-NULL (the reference's
 *    _keyval kernels never did, spmspv.h:1881).
 *  - Sparse outputs are sorted ascending and carry the STRUCTURAL support:
 *    row i is present iff some column of x's support has a stored entry in
 *    row i, even when the sum is exactly 0 (the reference's dense2sparse_inner
 *    dropped exact zeros, spmspv.h:1882-1928, while its sort kernels kept
 *    them). All kernels therefore return identical index sets.
 *  - alpha is accepted and ignored, as in the reference (csr5_spmspv.h:34).
 *  - No allocation after spmspv_alloc_pre_buffer(): it reserves the arena,
 *    the CSR5 copy at its largest width, the BFS and PageRank buffers (square
 *    matrices)
 *    and the selector stash of a 4,096-vector batch. The exceptions are
 *    stated where they apply: spmspv_to_csc (handle-owned CSC, once per
 *    matrix), a spmspv_run_batched call of more than 4,096 vectors (grows
 *    the stash once), spmspv_flush_cache and spmspv_measure_ceilings
 *    (measurement helpers). Without
 *    alloc_pre_buffer, the first call that needs a buffer allocates it.
 *  - No host exit(); errors come back as SPMSPV_* codes and
 *    spmspv_last_error() text.
 *  - All work is queued on the handle's stream (spmspv_set_stream). Calls
 *    that return a host scalar (y_nnz, bin_len, features) wait for it, as
 *    the reference's blocking D2H reads did; the results it describes are
 *    in place by then (y_nnz may arrive while the call's last kernel still
 *    restores the handle's scratch state, so other streams read the results
 *    after an event or a synchronisation).
 */
#ifndef SPMSPV_AMD_H_
#define SPMSPV_AMD_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes: the reference's SPMSPV_* values
 * (la/include/spmspv/csc-spmspv/detail/common.h:11-30) plus new ones. */
#define SPMSPV_SUCCESS 0
#define SPMSPV_UNKOWN_FORMAT -1
#define SPMSPV_UNSUPPORTED_CSR_SPMV -2
#define SPMSPV_UNSUPPORTED_VALUE_TYPE -3
#define SPMSPV_UNSUPPORTED_CSR_SPMSPV -4
#define SPMSPV_UNSUPPORTED_COO_SPMSPV -5
#define SPMSPV_FORMAT_ERR_HOLA_SPMV -6
#define SPMSPV_MAT_FORMAT_ERR -13
#define SPMSPV_VEC_FORMAT_ERR -14
#define SPMSPV_INVALID_ARGUMENT -20
#define SPMSPV_NOT_RESERVED -21
#define SPMSPV_DEVICE_ERROR -22
#define SPMSPV_OUT_OF_MEMORY -23
#define SPMSPV_NOT_IMPLEMENTED -24

/* Kernel slots: the reference's 3-factor dispatch
 * (la/test/new_test_app.cu:404-465; solution x workload x write-back). */
#define SPMSPV_K_COL_NAIVE_ATOMIC 0 /* CscBasedNoSortNaiveSpmspv[_keyval]  spmspv.h:2117-2169 */
#define SPMSPV_K_COL_NAIVE_SORT 1   /* CscBasedSortNaiveSpmspv            spmspv.h:2171-2201 */
#define SPMSPV_K_COL_LB_ATOMIC 2    /* CscBasedNoSortMergeSpmspv[_keyval]  spmspv.h:2035-2088 */
#define SPMSPV_K_COL_LB_SORT 3      /* CscBasedSortMergeSpmspv            spmspv.h:2090-2115 */
#define SPMSPV_K_ROW_NAIVE 4        /* naivespmspv                        spmspv.h:1700-1706 */
#define SPMSPV_K_ROW_LB 5           /* spmspv -> holaspmspv / csr5spmspv   spmspv.h:1812-1821 */
#define SPMSPV_K_SPMV_NAIVE 6       /* naivespmv                          spmspv.h:1676-1683 */
#define SPMSPV_K_SPMV_LB 7          /* spmv -> holaspmv / csr5spmv         spmspv.h:1789-1798 */
#define SPMSPV_K_COL_BUCKET 8       /* CscBasedBucketSpmspv                csc_bucket_spmspv.cuh:766-1086 */
/* two-layer load balance (outside the selector's slots): CscBasedMyNoSortSpmspv[_keyval],
 * CscBasedSortMySpmspv, spmspv.h:1823-1878, 2000-2029 */
#define SPMSPV_K_COL_TWO_LAYER_ATOMIC 9
#define SPMSPV_K_COL_TWO_LAYER_SORT 10
#define SPMSPV_NUM_SLOTS 8          /* slots the selector chooses among */
#define SPMSPV_NUM_KERNELS 11

/* Vector types, set_vector_type (spmspv.h:977-982): 0 sparse, 1 dense,
 * 2 bit array. */
#define SPMSPV_VEC_SPARSE 0
#define SPMSPV_VEC_DENSE 1
#define SPMSPV_VEC_BITARRAY 2

#define SPMSPV_NUM_FEATURES 16

typedef struct spmspv_handle_s* spmspv_handle_t;

/* ---- lifecycle ---------------------------------------------------------- */
/* SpmspvHandle(m, n, nnz), spmspv.h:90-141. */
int spmspv_create(spmspv_handle_t* h, int m, int n, int nnz);
/* Destroy, spmspv.h:2863-2884 (frees only handle-owned memory). */
int spmspv_destroy(spmspv_handle_t h);
/* No reference counterpart: the reference used the default stream only. */
int spmspv_set_stream(spmspv_handle_t h, void* hip_stream);
/* warmup, spmspv.h:347-351. */
int spmspv_warmup(spmspv_handle_t h);
/* Text of the last error on this handle (thread-local if h is NULL). */
const char* spmspv_last_error(spmspv_handle_t h);

/* ---- matrix input ------------------------------------------------------- */
/* InputCSR, spmspv.h:353-364 (borrows d_row_ptr[m+1], d_col[nnz], d_val[nnz]). */
int spmspv_input_csr(spmspv_handle_t h, int nnz, int* d_row_ptr, int* d_col, float* d_val);
/* InputCSC, spmspv.h:366-377 (borrows d_row[nnz], d_col_ptr[n+1], d_val[nnz]). */
int spmspv_input_csc(spmspv_handle_t h, int nnz, int* d_row, int* d_col_ptr, float* d_val);
/* ToCSC, spmspv.h:383-419: device CSR->CSC transpose into handle-owned
 * arrays (replaces cusparseScsr2csc); rows ascending inside each column. */
int spmspv_to_csc(spmspv_handle_t h);
/* allocPreBuffer, spmspv.h:1722-1763: every per-call scratch buffer, sized
 * from m, n, nnz; also the nnz-balanced row tiling (holaPreprocess,
 * spmspv.h:1653-1665) when CSR is present. */
int spmspv_alloc_pre_buffer(spmspv_handle_t h);
/* deallocPreBuffer, spmspv.h:2849-2861. */
int spmspv_dealloc_pre_buffer(spmspv_handle_t h);
/* holaPreprocess / holaPostprocess, spmspv.h:1653-1674 (row tiling). */
int spmspv_hola_preprocess(spmspv_handle_t h);
int spmspv_hola_postprocess(spmspv_handle_t h);
/* CSR5Preprocess / CSR5Postprocess / setSigma, spmspv.h:1395-1582. */
int spmspv_csr5_preprocess(spmspv_handle_t h);
int spmspv_csr5_postprocess(spmspv_handle_t h);
int spmspv_set_sigma(spmspv_handle_t h, int sigma);
/* set_spmv_type, spmspv.h:1766-1771: 0 Hola-style, 1 CSR5 for K5/K7. */
int spmspv_set_spmv_type(spmspv_handle_t h, int code);

/* ---- x input ------------------------------------------------------------ */
/* set_vector_type, spmspv.h:977-982. */
int spmspv_set_vector_type(spmspv_handle_t h, int code);
/* set_sparse_x, spmspv.h:991-1002 (borrows d_key[xnnz], d_val[xnnz];
 * keys distinct, in [0,n), any order). */
int spmspv_set_sparse_x(spmspv_handle_t h, int* d_key, float* d_val, int xnnz);
/* set_x, spmspv.h:984-989 (borrows d_x[n]). */
int spmspv_set_x(spmspv_handle_t h, float* d_x);
/* set_bitvector, spmspv.h:1004-1011 (borrows d_bits[num_ints]). */
int spmspv_set_bitvector(spmspv_handle_t h, int* d_bits, int num_ints);
/* set_sparse_x_dense_x, spmspv.h:1013-1025. */
int spmspv_set_sparse_x_dense_x(spmspv_handle_t h, int* d_key, float* d_val, float* d_x, int xnnz);

/* ---- vector format conversions (spmspv.h:1121-1390, 1880-1998) ---------- */
int spmspv_sparse2dense(spmspv_handle_t h);     /* sparse2dense    :1239-1268 */
int spmspv_sparse2bitarray(spmspv_handle_t h);  /* sparse2bitarray :1270-1307 */
int spmspv_dense2sparse(spmspv_handle_t h);     /* dense2sparse    :1176-1223 */
int spmspv_dense2bitarray(spmspv_handle_t h);   /* dense2bitarray  :1371-1389 */
int spmspv_bitarray2sparse(spmspv_handle_t h);  /* bitarray2sparse :1310-1352 */
/* dense2sparse_outer, spmspv.h:1950-1998 (keeps entries != 0; *ynnz set). */
int spmspv_dense2sparse_outer(spmspv_handle_t h, int m, float* d_dense, int* ynnz, int* d_key,
                              float* d_val);
/* sparse2dense_outer, spmspv.h:1930-1948. */
int spmspv_sparse2dense_outer(spmspv_handle_t h, int ynnz, int* d_key, float* d_val, int m,
                              float* d_dense);

/* ---- features (spmspv.h:2259-2753) -------------------------------------- */
/* computeBinlenVer2 / computeBinlen, spmspv.h:2363-2402, 2702-2721. */
int spmspv_compute_binlen(spmspv_handle_t h, long long* bin_len);
/* computeVecFeture_serial, spmspv.h:2337-2361, on the device. */
int spmspv_vec_features(spmspv_handle_t h, long long* bin_len, int* max_len, int* min_len);
/* computeMatFeture_serial, spmspv.h:2274-2334, on the device, from the CSR
 * row pointer: out[7] = {max, min, avg, range, std, edge_equality, gini}. */
int spmspv_mat_features(spmspv_handle_t h, float out[7]);
/* extractMatFeatureParallel, spmspv.h:920-976: stdRow with int mean. */
int spmspv_extract_mat_feature(spmspv_handle_t h, float* std_row);
/* computeGM1/2/3, spmspv.h:2724-2753 (needs bin_len). out[3]. */
int spmspv_compute_gm(spmspv_handle_t h, double out[3]);
/* The selector's 16-entry feature vector for the current x (no reference
 * counterpart; built from the features above). */
int spmspv_features(spmspv_handle_t h, float out[SPMSPV_NUM_FEATURES]);

/* ---- kernels ------------------------------------------------------------ */
/* Runs kernel slot `kernel_id` on the current x:
 *   d_y_dense (m floats, nullable) receives the dense result;
 *   d_y_key/d_y_val (capacity m, nullable) receive the sorted sparse result;
 *   y_nnz (host, nullable) receives the support size (synchronises).
 * The needed x conversions (sparse->dense, sparse->bit array) are done first
 * when x was set as sparse, exactly as the reference harness did before the
 * row kernels (la/test/new_test_app.cu:318-330). */
int spmspv_run(spmspv_handle_t h, int kernel_id, float alpha, float* d_y_dense, int* d_y_key,
               float* d_y_val, int* y_nnz);
/* Device address of the last result's support size (int), for callers that
 * do not want the host synchronisation of y_nnz. */
int* spmspv_ynnz_device(spmspv_handle_t h);

/* Reference method names, each = spmspv_run with the matching slot. */
int spmspv_csc_nosort_naive(spmspv_handle_t h, float alpha, int* ynnz, float* d_y);
int spmspv_csc_nosort_naive_keyval(spmspv_handle_t h, float alpha, int* ynnz, float* d_y,
                                   int* d_y_key, float* d_y_val);
int spmspv_csc_sort_naive(spmspv_handle_t h, float alpha, int* ynnz, int* d_y_key, float* d_y_val,
                          int method);
int spmspv_csc_nosort_merge(spmspv_handle_t h, int is_binlen_computed, float alpha, int* ynnz,
                            float* d_y, int* d_y_key, float* d_y_val);
int spmspv_csc_nosort_merge_keyval(spmspv_handle_t h, int is_binlen_computed, float alpha,
                                   int* ynnz, float* d_y, int* d_y_key, float* d_y_val);
int spmspv_csc_sort_merge(spmspv_handle_t h, int is_binlen_computed, float alpha, int* ynnz,
                          float* d_y, int* d_y_key, float* d_y_val);
/* CscBasedBucketSpmspv (csc_bucket_spmspv.cuh:766-1086): products written
 * into row bins (ranges of rows with about equal numbers of nonzeros, each
 * owning as many product slots as its rows hold nonzeros, so no count pass),
 * each bin summed per row in an LDS window (64-bit fixed point scaled to the
 * row's largest product; FP32 error at most that of one rounding of the
 * exact sum plus 2^-32 of the largest product). buckets = the number of
 * equal-nonzero bins aimed at, a power of two in [2, 1024] (bins also hold
 * at most 4,096 rows, so large m gets more); 0 keeps the handle's current
 * count (512, or the autotuner's pick). The first call on a matrix (and on
 * a new count) builds the layout on the host from the row pointer (one
 * copy back and one synchronisation). Other values: INVALID_ARGUMENT; a
 * matrix needing more than 6,144 bins (m above about 16.7 M rows):
 * INVALID_ARGUMENT. */
int spmspv_csc_bucket(spmspv_handle_t h, float alpha, int* ynnz, int* d_y_key, float* d_y_val,
                      int buckets);
/* CscBasedMyNoSortSpmspv[_keyval] / CscBasedSortMySpmspv (spmspv.h:1823-1878,
 * 2000-2029): the two-layer load balance, tiles of exactly 1,792 products
 * (the reference's write-back without the block offset,
 * csc_two_layer_load_balance_fuse_reduction_modify.cuh:146-149, fixed). */
int spmspv_csc_my_nosort(spmspv_handle_t h, int is_binlen_computed, float alpha, int* ynnz, float* d_y,
                         int* d_y_key, float* d_y_val);
int spmspv_csc_my_nosort_keyval(spmspv_handle_t h, int is_binlen_computed, float alpha, int* ynnz,
                                float* d_y, int* d_y_key, float* d_y_val);
int spmspv_csc_sort_my(spmspv_handle_t h, int is_binlen_computed, float alpha, int* ynnz, float* d_y,
                       int* d_y_key, float* d_y_val);
int spmspv_naivespmspv(spmspv_handle_t h, float alpha, float* d_y);
int spmspv_holaspmspv(spmspv_handle_t h, float alpha, float* d_y);
int spmspv_csr5spmspv(spmspv_handle_t h, float alpha, float* d_y);
int spmspv_spmspv(spmspv_handle_t h, float alpha, float* d_y);
int spmspv_naivespmv(spmspv_handle_t h, float alpha, float* d_y);
int spmspv_holaspmv(spmspv_handle_t h, float alpha, float* d_y);
int spmspv_csr5spmv(spmspv_handle_t h, float alpha, float* d_y);
int spmspv_spmv(spmspv_handle_t h, float alpha, float* d_y);

/* ---- autotuner (no reference counterpart: the reference's tile sizes and
 * thread counts are compile-time macros, la/include/spmspv/config.h:8-30,
 * scan_config.h:4-26) ------------------------------------------------------ */
/* On the current x, times every candidate team size of the naive column
 * and row slots (2..64 lanes), every row-merge tile width (5, 7 or 11
 * items per lane) and every CSR5 width (sigma 8, 16, 24, 32) with the
 * handle's stream events, median of `reps`, and keeps the fastest for this
 * matrix: the K5/K7 format (set_spmv_type) included. report (nullable,
 * 5 ints) gets {team_col, team_row, row_vt, spmv_type, sigma (0 if the
 * merge-path tiles won)}. */
int spmspv_autotune(spmspv_handle_t h, int reps, int* report);
/* As spmspv_autotune; csr5 = 0 keeps the current K5/K7 format and only
 * tunes its parameters' merge-path side. Both also pick the bucket slot's
 * number of row bins (about 256, 512 or 1,024 of equal nonzeros), the two
 * size thresholds of the single-workgroup paths (the small-x single launch,
 * up to 11,264 products; the one-workgroup sort + reduce of the sort slots,
 * up to 3,072 items) and the bucket rule's bin_len range
 * (spmspv_set_bucket_range), each timed on sub-vectors of x (one key from
 * each of k equal intervals) whose bin_len lies just under a candidate. */
int spmspv_autotune2(spmspv_handle_t h, int reps, int csr5, int* report);
/* No counterpart in the reference: every tuned parameter, out[10] =
 * {team_col, team_row, row_vt, spmv_type, sigma (0 unless CSR5), row
 * buckets, small-x product threshold (0: path off), single-workgroup sort
 * threshold, bucket rule's lowest and highest bin_len (0 and -1: rule off;
 * INT_MAX: no upper end)}. */
int spmspv_autotune_report(spmspv_handle_t h, int out[10]);
/* No counterpart in the reference: the bucket rule behind the selector.
 * In spmspv_run_auto and spmspv_run_batched, a choice of slot 2, 3, 5 or 7
 * becomes the bucket slot (8) when lo <= bin_len <= hi (bin_len sampled for
 * a large x); hi < lo turns it off (the default). spmspv_autotune sets the
 * range where the bucket slot beat those four on sub-vectors of x (off when
 * the matrix needs more row bins than the slot holds); a new matrix
 * (InputCSR / InputCSC) turns it off again. Setting a range (hi >= lo) on a
 * matrix builds the slot's row bins (one host round trip the first time)
 * and is refused with SPMSPV_INVALID_ARGUMENT, the rule left off, when the
 * matrix needs more than 6,144 bins. */
int spmspv_set_bucket_range(spmspv_handle_t h, long long lo, long long hi);

/* ---- learned selector (no reference counterpart in C++: the reference's
 * predictors are stubs returning 0, la/test/new_test_app.cu:121-131) ------ */
/* MLP 16 -> 64 -> 64 -> 8, FP32. Parameters packed W1[64*16] b1[64]
 * W2[64*64] b2[64] W3[8*64] b3[8] (row-major, out x in), followed by the
 * 16 feature means and 16 inverse standard deviations used to normalise
 * the input. */
#define SPMSPV_MLP_F 16
#define SPMSPV_MLP_H1 64
#define SPMSPV_MLP_H2 64
#define SPMSPV_MLP_C 8
#define SPMSPV_MLP_NPARAM (64 * 16 + 64 + 64 * 64 + 64 + 8 * 64 + 8)
#define SPMSPV_MLP_NWEIGHTS (SPMSPV_MLP_NPARAM + 2 * SPMSPV_MLP_F)
int spmspv_selector_set_weights(spmspv_handle_t h, const float* w, int count);
/* mode: 0 = MLP (needs weights), 1 = decision-tree fallback, 2 = fixed slot
 * `fixed_kernel`. */
int spmspv_selector_mode(spmspv_handle_t h, int mode, int fixed_kernel);
/* Features + MLP inference in one device launch; returns the chosen slot. */
int spmspv_select(spmspv_handle_t h, int* kernel_id);
/* spmspv_select then spmspv_run of the chosen slot (the reference's
 * predict-then-run flow, la/test/new_test_app.cu:121-131, with its stub
 * predictors replaced by the selector). With speculation on (the
 * default), the slot the selector chose last time for an x of the same
 * size class (xnnz within a power of two) is queued behind the selector
 * before its choice reaches the host; its launches do nothing on the device
 * unless the choice matches. Results are the same either way. */
int spmspv_run_auto(spmspv_handle_t h, float alpha, float* d_y_dense, int* d_y_key,
                    float* d_y_val, int* y_nnz, int* kernel_used);
/* No counterpart in the reference: switches the speculative queueing of
 * spmspv_run_auto on (1) or off (0), and reports its right and wrong
 * guesses so far (either pointer may be null). */
int spmspv_set_speculation(spmspv_handle_t h, int on);
/* Switching-aware selection (the reference's "real" accounting charges the
 * x format conversion whenever consecutive calls change kernel family,
 * model/1-test_dtree.py:266-287): with penalty > 0, spmspv_run_auto keeps
 * the previous call's family (column 0-3, row SpMSpV 4-5, SpMV 6-7) when
 * that family's best slot has a logit within `penalty` of the network's top
 * choice. The network is trained on soft targets exp(-(t/t_best - 1)/tau),
 * so penalty = switching cost / (tau * t_best). 0 (the default) is off. */
int spmspv_set_switch_penalty(spmspv_handle_t h, float penalty);
/* No counterpart in the reference: on (1, the default) a column slot on an
 * x of at most 4,096 keys whose products fit one workgroup (11,264) runs
 * as one launch: scan, gather, LDS sort, reduce and the sorted output (the
 * sort-based slots' result); spmspv_run_auto then decides such calls by
 * this size rule before the network and reports slot 1. Off (0): every
 * slot always runs its own launches. Results are the same either way. */
int spmspv_set_small_x_path(spmspv_handle_t h, int on);
int spmspv_speculation_stats(spmspv_handle_t h, long long* hits, long long* misses);

/* Training on the GPU with the library's own forward/backward kernels.
 * X[B*16] raw features (host), targets[B*8] soft labels (host; rows sum to
 * 1). Trains in place on params[NPARAM] (host in/out) with Adam for
 * `epochs` passes of minibatches of `batch`; returns the final mean loss in
 * *loss. The normaliser (means, inverse stds) is computed from X and
 * appended, so the output array holds SPMSPV_MLP_NWEIGHTS floats. */
int spmspv_mlp_train(const float* X, const float* targets, int B, int epochs, int batch,
                     float lr, unsigned seed, float* weights_out, float* loss);
/* One forward+backward on the GPU for the gradient check: X[B*16] already
 * normalised, targets[B*8]; writes loss and grad[NPARAM]. */
int spmspv_mlp_loss_grad(const float* params, const float* X, const float* targets, int B,
                         float* loss, float* grad);
/* Inference on the GPU: logits[B*8] for normalised inputs X[B*16]. */
int spmspv_mlp_forward(const float* params, const float* X, int B, float* logits);

/* ---- BFS driver (the reference only replays recorded frontiers,
 * la/test/test_application.cu:207-261; this is the loop it lacked) ------- */
/* Level-synchronous BFS over the CSC pattern from `src`. d_levels[n]
 * receives int32 levels (-1 unreached). policy: -1 = learned per-level
 * selection, 0..7 = fixed slot. Levels are captured as hipGraphs when
 * use_graphs != 0. stats (nullable) gets {levels, edges_traversed,
 * seconds, reached}. */
int spmspv_bfs(spmspv_handle_t h, int src, int* d_levels, int policy, int use_graphs,
               double stats[4]);
/* No counterpart in the reference: up to 64 BFSs advanced together over the
 * same pattern as spmspv_bfs, one bit per source in a 64-bit word per
 * vertex, so one pass over a level's edges serves every source (top-down
 * push for small frontiers, bottom-up pull with early exit for large ones,
 * the boolean semiring in integer words). srcs[nsrc] (host, 1 <= nsrc <= 64,
 * repeats allowed); d_levels[n * nsrc] (device) receives the levels
 * vertex-major: d_levels[v * nsrc + i] is the level of v from srcs[i]
 * (-1 unreached), the same as spmspv_bfs from srcs[i]. stats (nullable) =
 * {levels, undirected edges summed over the sources' components, seconds,
 * reached (vertex, source) pairs}. */
int spmspv_bfs_multi(spmspv_handle_t h, int nsrc, const int* srcs, int* d_levels, double stats[4]);
/* No counterpart in the reference: spmspv_bfs that also records, for each of
 * the first max_levels levels, the level's device time in ms (level_ms,
 * host) and the 16 selector features of its frontier (level_feat, host,
 * 16 floats per level) -- the labels of the BFS selector's training set.
 * Uses the handle's timing events. */
int spmspv_bfs_trace(spmspv_handle_t h, int src, int* d_levels, int policy, int use_graphs,
                     double stats[4], int max_levels, float* level_ms, float* level_feat);
/* No counterpart in the reference: selector weights for the BFS levels
 * (trained on BFS level timings, whose boolean kernels rank differently from
 * the SpMSpV slots); without them the BFS driver uses the SpMSpV selector's
 * weights. */
int spmspv_bfs_selector_set_weights(spmspv_handle_t h, const float* w, int count);

/* ---- PageRank-style driver (the reference replays frontiers recorded from
 * a PageRank run, la/test/test_application.cu:206-210, 252-256; this runs
 * the iteration that produces them) ---------------------------------------- */
/* Residual-push PageRank over the CSC (column u = u's out-edges, weights
 * a_vu; the walk leaves u with probability a_vu / sum_v a_vu): every
 * iteration's active set is a sparse x for one selector-dispatched SpMSpV
 * (policy -1) or a fixed slot (0..7). damping d in (0,1); a vertex stays
 * active while its residual exceeds eps * (1 - d) / n; at most max_iter
 * iterations. d_rank[n] (device) receives the ranks, the solution of
 * (I - d A W^-1) p = (1 - d)/n within sum|p - p*| <= eps * d once converged.
 * stats (nullable) = {iterations, seconds, sum of x sizes, vertices still
 * active}; sizes (host, nullable, max_iter ints) = x size per iteration. */
int spmspv_pagerank(spmspv_handle_t h, float damping, float eps, int max_iter, int policy,
                    float* d_rank, double stats[4], int* sizes);

/* ---- batched multi-vector SpMSpV (no reference counterpart) ------------- */
/* nvec sparse vectors in CSR-of-vectors form: d_vptr[nvec+1] (device),
 * d_vkey, d_vval. Vector v's sorted result goes to d_ykey/d_yval at the
 * caller's device offsets d_yoff[v] .. d_yoff[v+1]-1 (capacity must be at
 * least min(m, bin_len of v), else SPMSPV_INVALID_ARGUMENT before anything
 * runs); its count to d_ynnz[v]. The selector picks each vector's slot in
 * one pass over the batch; kernel_used (host, nullable) gets the slots.
 * Two host synchronisations per batch (one sizing readback, one at the
 * end) instead of two per vector. The handle's x is unset afterwards. */
int spmspv_run_batched(spmspv_handle_t h, int nvec, const int* d_vptr, const int* d_vkey,
                       const float* d_vval, const long long* d_yoff, int* d_ykey, float* d_yval,
                       int* d_ynnz, int* kernel_used);

/* ---- no-allocation entry point ------------------------------------------ */
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc). The load-balanced atomic column SpMSpV (the kernels
 * of spmspv_run with SPMSPV_K_COL_LB_ATOMIC) on a caller's CSC
 * (d_col_ptr[n+1], d_row[nnz], d_val[nnz]) and sparse x (d_x_key[xnnz]
 * distinct keys in [0,n), d_x_val[xnnz]), without a handle: every scratch
 * buffer comes from one 32 MiB array in device memory that each GPU holds
 * until the process ends (enough for m <= 2^21 rows with xnnz <= 2^20 at any
 * nnz; larger sizes are SPMSPV_INVALID_ARGUMENT with the bytes needed). The
 * sorted result goes to d_y_key/d_y_val (capacity m) and its size to the
 * device int *d_y_nnz, all queued on `stream` (NULL: the default stream);
 * nothing is allocated and the host does not wait. Calls that overlap on
 * one GPU share its scratch array and race. Errors: spmspv_last_error(NULL). */
int spmspv_csc_spmspv_noalloc(int m, int n, long long nnz, const int* d_col_ptr, const int* d_row,
                              const float* d_val, const int* d_x_key, const float* d_x_val, int xnnz,
                              int* d_y_key, float* d_y_val, int* d_y_nnz, void* stream);

/* ---- utilities (the reference harness called the CUDA runtime directly) - */
int spmspv_device_count(int* count);
int spmspv_malloc(void** d_ptr, size_t bytes);
int spmspv_free(void* d_ptr);
int spmspv_memcpy_h2d(void* d_dst, const void* src, size_t bytes);
int spmspv_memcpy_d2h(void* dst, const void* d_src, size_t bytes);
int spmspv_memset(void* d_ptr, int value, size_t bytes);
int spmspv_memcpy_d2d(void* d_dst, const void* d_src, size_t bytes);
int spmspv_synchronize(void);
/* Stream timing: events recorded on the handle's stream. */
int spmspv_timer_start(spmspv_handle_t h);
int spmspv_timer_stop(spmspv_handle_t h, float* ms);
/* No counterpart in the reference (measurement): with on != 0, the next
 * K5/K7 calls record events around their tile kernel alone (not the
 * fix-up, compaction or x conversion); spmspv_main_kernel_ms returns the
 * duration of the last call's tile kernel, or SPMSPV_INVALID_ARGUMENT when
 * the last call ran a slot without one. */
int spmspv_time_main_kernel(spmspv_handle_t h, int on);
int spmspv_main_kernel_ms(spmspv_handle_t h, float* ms);
/* No counterpart in the reference (measurement): the GPU's ceilings for
 * the roofline, each launch after a 512 MiB eviction write, best of reps:
 * out[0] streaming read GB/s (512 MiB, 16-byte nontemporal loads, 8 in
 * flight per lane), out[1] device-to-device copy GB/s (256 MiB read + 256
 * MiB written), out[2] random 4-byte gathers per second (in G) over an x of
 * the handle's n columns. Allocates and frees 1 GiB of scratch. */
int spmspv_measure_ceilings(spmspv_handle_t h, int reps, double out[3]);
/* Writes `bytes` through a scratch buffer to evict the 256 MB MALL. */
int spmspv_flush_cache(spmspv_handle_t h, size_t bytes);

/* ---- matrix and frontier files (host; la/include/spmspv/mtx.hpp and
 * readsvmdata.hpp) -------------------------------------------------------- */
/* ReadMTX + SortMTX (mtx.hpp:218-402) and the harness's coo2csr
 * (la/test/new_test_app.cu:547-559): a MatrixMarket coordinate file as CSR
 * with rows' entries in column order. Pattern and complex entries read as
 * 1.0; a "symmetric" file is mirrored (off-diagonal entries twice, same
 * value); other symmetry tags read as general, as in the reference;
 * duplicates are kept. Call with row_ptr == NULL to get m, n and nnz first;
 * then row_ptr[m+1], col[nnz], val[nnz] (host) are filled. Out-of-range
 * indices, truncated entries and dense (array) files are
 * SPMSPV_MAT_FORMAT_ERR (the reference read past them). */
int spmspv_mtx_read_csr(const char* path, int* m, int* n, long long* nnz, int* row_ptr, int* col,
                        float* val);
/* loadCSR_header (mtx.hpp:750-773): sizes from a binary CSR file (24-byte
 * header: "Hola\1CSR\1", 3 padding bytes, int32 rows, columns, entries). */
int spmspv_csr_load_header(const char* path, int* m, int* n, int* nnz);
/* loadCSR (mtx.hpp:775-804): values, column indices and row pointers
 * (host arrays) of a binary CSR file whose header matches m, n, nnz. */
int spmspv_csr_load(const char* path, int m, int n, int nnz, int* row_ptr, int* col, float* val);
/* storeCSR (mtx.hpp:806-819): writes that file (padding bytes zero). */
int spmspv_csr_store(const char* path, int m, int n, int nnz, const int* row_ptr, const int* col,
                     const float* val);
/* readNNZXFromFile + readSparseXInxFromFile (readsvmdata.hpp:223-258): the
 * x size of every iteration of a recorded BFS / PageRank run (whitespace
 * separated integers). sizes == NULL: *count only. */
int spmspv_frontier_sizes(const char* path, int* count, int* sizes);
/* extractSparseXfromFile (readsvmdata.hpp:308-348): the keys on line `line`
 * (0-based) of a frontier file, every value 1.0; at most cap keys, else
 * SPMSPV_INVALID_ARGUMENT. A line past the end gives xnnz = 0. */
int spmspv_frontier_read(const char* path, int line, int cap, int* xnnz, int* keys, float* vals);

/* ---- synthetic inputs (host; the reference read .mtx files) ------------- */
/* Square n x n CSC with exactly k distinct rows per column (Floyd sampling,
 * rows ascending), values U[-1,1]. Arrays: col_ptr[n+1], row[n*k], val[n*k]. */
int spmspv_gen_uniform_csc(int n, int k, uint64_t seed, int* col_ptr, int* row, float* val);
/* R-MAT (Graph500 quadrant recursion) with 2^scale vertices and
 * edge_factor*2^scale generated edges, symmetrised, deduplicated, self-loops
 * dropped; values U(0,1] equal on (i,j) and (j,i). Writes CSR (== CSC by
 * symmetry). Call with row_ptr == NULL to get the exact nnz in *nnz first. */
int spmspv_gen_rmat_csr(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                        long long* nnz, int* row_ptr, int* col, float* val);
/* k distinct sorted keys in [0,n) and values U[lo,hi] (for x). */
int spmspv_gen_sparse_vector(int n, int k, uint64_t seed, float lo, float hi, int* key,
                             float* val);
/* Banded n x n CSR with half-bandwidth bw, values U(0,1]. Call with
 * row_ptr == NULL to get nnz. */
int spmspv_gen_banded_csr(int n, int bw, uint64_t seed, long long* nnz, int* row_ptr, int* col,
                          float* val);
/* Block-diagonal n x n CSR with dense bs x bs blocks, values U(0,1]. */
int spmspv_gen_blockdiag_csr(int n, int bs, uint64_t seed, long long* nnz, int* row_ptr, int* col,
                             float* val);
/* Uniform random CSR with mean k per row (each row: Binomial-like count,
 * distinct sorted columns). Call with row_ptr == NULL to get nnz. */
int spmspv_gen_random_csr(int m, int n, double mean_per_row, uint64_t seed, long long* nnz,
                          int* row_ptr, int* col, float* val);

#ifdef __cplusplus
}
#endif
#endif /* SPMSPV_AMD_H_ */
