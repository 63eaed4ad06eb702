/*
 * perflibs_sparse_amd.h — MI355X-specific additions to the perflibs_sparse.h
 * interface.  Every symbol here is new (prefix perflibs_amd_); none of them
 * exists in the reference, and nothing in perflibs_sparse.h is changed by them.
 *
 * Groups:
 *   - streams and asynchronous execution on device-resident vectors
 *     (the drop-in perflibs_spmv_exec_* is synchronous),
 *   - forcing an SpMV kernel variant, the GPU counterpart of the reference's
 *     internal set_C/set_sigma audition knobs (src/matrix_state.cpp:445-462),
 *   - value storage precision (FP64 / FP32 / BF16, FP32 accumulation for the
 *     reduced formats),
 *   - format selection: heuristic, exhaustive autotuner, learned MLP,
 *   - inspection of the device layout (tests) and batched execution.
 */
#ifndef PERFLIBS_SPARSE_AMD_H
#define PERFLIBS_SPARSE_AMD_H

#include <stddef.h>
#include <stdint.h>

#include "perflibs_sparse.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- SpMV kernel variants (the selector's label space) ---- */
enum perflibs_amd_variant {
  PERFLIBS_AMD_VARIANT_AUTO = -1,
  PERFLIBS_AMD_CSR_VECTOR = 0,   /* sub-wave lanes per row, width from mean row length */
  PERFLIBS_AMD_CSR_ADAPTIVE = 1, /* nnz-balanced row blocks, LDS-staged products */
  PERFLIBS_AMD_SELL_C32_S0 = 2,  /* SELL-C-sigma, C = 32, no sorting */
  PERFLIBS_AMD_SELL_C32_S64 = 3,
  PERFLIBS_AMD_SELL_C32_S1024 = 4,
  PERFLIBS_AMD_SELL_C32_S16384 = 5,
  PERFLIBS_AMD_SELL_C64_S0 = 6,
  PERFLIBS_AMD_SELL_C64_S64 = 7,
  PERFLIBS_AMD_SELL_C64_S1024 = 8,
  PERFLIBS_AMD_SELL_C64_S16384 = 9,
  PERFLIBS_AMD_BSR_B4 = 10,      /* blocked 4x4 (converted from CSR when needed) */
  PERFLIBS_AMD_BSR_B8 = 11,
  PERFLIBS_AMD_CSR_PANEL = 12,   /* row tiles swept over column panels of x, LDS row sums */
  PERFLIBS_AMD_NUM_VARIANTS = 13
};

/* ---- how perflibs_spmv_optimize picks a variant ---- */
enum perflibs_amd_select_mode {
  /* MLP on matrix features ranks the candidates, the best three predictions
   * are timed (the exhaustive mode's protocol) and the fastest is kept;
   * features outside the model's training range run the exhaustive sweep.
   * The default when a model is loaded. */
  PERFLIBS_AMD_SELECT_LEARNED = 0,
  PERFLIBS_AMD_SELECT_EXHAUSTIVE = 1, /* time every candidate variant, keep the fastest */
  PERFLIBS_AMD_SELECT_HEURISTIC = 2,  /* fixed rules on the same features, no network */
  PERFLIBS_AMD_SELECT_FIXED_CSR = 3,  /* always PERFLIBS_AMD_CSR_VECTOR */
  PERFLIBS_AMD_SELECT_NETWORK = 4     /* the MLP's best prediction alone, nothing timed */
};

/* How the last perflibs_spmv_optimize of a handle chose its variant. */
enum perflibs_amd_select_path {
  PERFLIBS_AMD_PATH_NONE = -1,       /* no selection ran (no optimize, tiny or hinted FEW) */
  PERFLIBS_AMD_PATH_FIXED = 0,       /* forced variant or fixed CSR */
  PERFLIBS_AMD_PATH_HEURISTIC = 1,
  PERFLIBS_AMD_PATH_NETWORK = 2,     /* the network's prediction, nothing timed */
  PERFLIBS_AMD_PATH_AUDITION = 3,    /* the fastest of the best predictions, timed */
  PERFLIBS_AMD_PATH_EXHAUSTIVE = 4   /* every candidate timed */
};

/* ---- value storage precision ---- */
enum perflibs_amd_precision {
  PERFLIBS_AMD_VALUES_NATIVE = 0, /* the handle's own type (FP64 for _d, FP32 for _s) */
  PERFLIBS_AMD_VALUES_BF16 = 2    /* _s handles only: BF16 values, FP32 x/y and accumulation */
};

#define PERFLIBS_AMD_NUM_FEATURES 12

/* Stream used by every kernel launch of the library (default: the null
 * stream).  `stream` is a hipStream_t. */
perflibs_status_t perflibs_amd_set_stream(void *stream);
void *perflibs_amd_get_stream(void);

/* y = alpha*op(A)*x + beta*y with x and y in device memory, enqueued on the
 * library stream and not synchronized.  Same semantics as the drop-in call. */
perflibs_status_t perflibs_amd_spmv_exec_async_d(
    enum perflibs_sparse_hint_value trans, double alpha, perflibs_spmat_t A,
    const double *x_dev, double beta, double *y_dev);
perflibs_status_t perflibs_amd_spmv_exec_async_s(
    enum perflibs_sparse_hint_value trans, float alpha, perflibs_spmat_t A,
    const float *x_dev, float beta, float *y_dev);

/* C = alpha*A*B + beta*C with row-major device operands (ldb, ldc in
 * elements), enqueued on the library stream and not synchronized. */
perflibs_status_t perflibs_amd_spmm_exec_async_d(
    double alpha, perflibs_spmat_t A, perflibs_int_t k, const double *B_dev,
    perflibs_int_t ldb, double beta, double *C_dev, perflibs_int_t ldc);
perflibs_status_t perflibs_amd_spmm_exec_async_s(
    float alpha, perflibs_spmat_t A, perflibs_int_t k, const float *B_dev,
    perflibs_int_t ldb, float beta, float *C_dev, perflibs_int_t ldc);

/* Variant control and introspection. */
perflibs_status_t perflibs_amd_set_select_mode(perflibs_spmat_t A, int mode);
perflibs_status_t perflibs_amd_force_variant(perflibs_spmat_t A, int variant);
perflibs_status_t perflibs_amd_get_variant(perflibs_spmat_t A, int *variant);
perflibs_status_t perflibs_amd_set_precision(perflibs_spmat_t A, int precision);
/* Force SELL-C-sigma with any C in [1, 64] and sigma a multiple of C (or 0). */
perflibs_status_t perflibs_amd_force_sell(perflibs_spmat_t A, perflibs_int_t C,
                                          perflibs_int_t sigma);

/* Force the column-panel variant with panels of 2^panel_log2 columns
 * (1 <= panel_log2 <= 24, or 0 for the default) and at most wg_per_cu resident
 * workgroups of 4 wave tiles per CU (1..16, or 0 for the default). */
perflibs_status_t perflibs_amd_force_panel(perflibs_spmat_t A, perflibs_int_t panel_log2,
                                           perflibs_int_t wg_per_cu);
/* Column-compressed row blocks of the SpMM plan (for tests and tuning):
 * blocks_state 0 = not built, 1 = built and used for k >= 8, -1 = built but
 * too little B-row reuse (direct kernels run); list_max = the list length
 * limit the blocks were cut with; list_total = total list length.
 * blocks_state 2: NOTRANS products run on the blocks of a former BSR object
 * (nblocks = block rows, list_max = block size, list_total = blocks).
 * Status 1 when the handle has neither. */
perflibs_status_t perflibs_amd_spmm_info(perflibs_spmat_t A, perflibs_int_t *blocks_state,
                                         perflibs_int_t *nblocks, perflibs_int_t *list_max,
                                         int64_t *list_total);
/* How the SpMM plan groups rows into those blocks: 0 (default) keeps the cut
 * with the shorter column lists of two, consecutive rows or rows clustered by
 * shared columns (a block grows by the row that adds the fewest unlisted
 * columns); 1 cuts consecutive rows only; 2 takes the clustered cut whenever
 * it completes.  Takes effect at the next SpMM with k >= 8. */
perflibs_status_t perflibs_amd_set_spmm_rows(perflibs_spmat_t A, int rows);
/* *clustered = 1 when the built blocks hold clustered rows, 0 for consecutive
 * rows; status 1 before the blocks are built. */
perflibs_status_t perflibs_amd_spmm_rows_clustered(perflibs_spmat_t A, int *clustered);
/* The row blocks an SpMM plan would cut for an m x n CSR pattern (row_ptr[0]
 * = index base 0 or 1), computed on the host with no device or handle: *used
 * = 1 when the blocks give enough B-row reuse to be used; nblocks, list_max,
 * list_total and clustered as above; row_order[m] (0-based rows, block by
 * block) and block_starts[nblocks + 1] (offsets into row_order; give room for
 * m + 1) are filled when not null.  rows as for perflibs_amd_set_spmm_rows. */
perflibs_status_t perflibs_amd_spmm_blocks_inspect(perflibs_int_t m, perflibs_int_t n,
                                                  const perflibs_int_t *row_ptr,
                                                  const perflibs_int_t *col_indx, int rows,
                                                  int *used, perflibs_int_t *nblocks,
                                                  perflibs_int_t *list_max, int64_t *list_total,
                                                  int *clustered, perflibs_int_t *row_order,
                                                  perflibs_int_t *block_starts);
/* Shape of a column-panel plan (for tests and tuning). */
perflibs_status_t perflibs_amd_panel_info(perflibs_spmat_t A, perflibs_int_t *panel_log2,
                                          perflibs_int_t *npanels, perflibs_int_t *ntiles,
                                          perflibs_int_t *rows_per_tile_max,
                                          perflibs_int_t *grid);

/* Measurement, not a product path: runs the column-panel kernel's index/value
 * stream and x gathers over the handle's panel plan (built by an optimize
 * with PERFLIBS_AMD_CSR_PANEL forced) without the row sums, on the library
 * stream and not synchronized.  out_dev receives one partial sum per thread:
 * grid * 256 values of the handle's type (grid from perflibs_amd_panel_info).
 * Its time is the ceiling the gathers alone set on the panel SpMV.  Status 1
 * without a panel plan. */
perflibs_status_t perflibs_amd_probe_panel_gathers(perflibs_spmat_t A, const void *x_dev,
                                                  void *out_dev);

/* Measurement, not a product path: `count` random 8-byte gathers from
 * table_dev[table_len] (a power of two; indices from an integer hash, 8 in
 * flight per thread), on the library stream, not synchronized.  out_dev
 * receives PERFLIBS_AMD_PROBE_GRID * 256 partial sums.  A table that fits an
 * XCD's L2 (4 MB) gives the random-gather rate of the L1 -> L2 path, a larger
 * one that of the MALL. */
#define PERFLIBS_AMD_PROBE_GRID 8192
perflibs_status_t perflibs_amd_probe_random_gathers(const double *table_dev, int64_t table_len,
                                                   int64_t count, double *out_dev);

/* Matrix features used by the selector (PERFLIBS_AMD_NUM_FEATURES floats),
 * computed on the GPU from the CSR structure. */
perflibs_status_t perflibs_amd_features(perflibs_spmat_t A, float *features);

/* Padding efficiency of SELL-C-sigma(C, sigma) for op(A) (the SpMV hint's
 * operation): beta = non-zeros / padded slots of real rows, and the smallest
 * and largest per-thread beta over the reference's nnz-division chunk split
 * into nthreads threads (nthreads <= 0: the device's CU count).  Counterpart
 * of perflibs_scs::get_beta / get_beta_threads (src/sell_c_sigma.cpp:369-452);
 * the chunk widths are computed on the GPU from the row pointers when C and
 * sigma are powers of two (sigma 0 or at most 16384).  optimize() keeps a
 * SELL-C-sigma variant as a candidate only at beta >= 0.96 and per-thread
 * beta > 0.85 (the rule of src/matvec.cpp:303-325).  min_beta / max_beta may
 * be NULL. */
perflibs_status_t perflibs_amd_sell_beta(perflibs_spmat_t A, perflibs_int_t C,
                                         perflibs_int_t sigma, perflibs_int_t nthreads,
                                         double *beta, double *min_beta, double *max_beta);

/* Time every variant of A (warmup + reps launches each, HIP events on the
 * library stream), writing milliseconds per launch; variants that do not
 * apply get a negative time.  Used by the exhaustive mode and to build
 * selector training data. */
perflibs_status_t perflibs_amd_time_variants(perflibs_spmat_t A, int warmup,
                                             int reps, double *ms);

/* Selector model: 12 -> 64 -> 64 -> PERFLIBS_AMD_NUM_VARIANTS MLP predicting
 * log(time / best time) per variant.  Weights are a flat FP32 array
 * (W1[12x64], b1[64], W2[64x64], b2[64], W3[64xK], b3[K]) plus the feature
 * standardisation (mean[12], std[12]). */
#define PERFLIBS_AMD_MLP_HIDDEN 64
size_t perflibs_amd_mlp_num_params(void);
perflibs_status_t perflibs_amd_set_selector_model(const float *params,
                                                  const float *feat_mean,
                                                  const float *feat_std);
perflibs_status_t perflibs_amd_clear_selector_model(void);
/* Training range of each feature (PERFLIBS_AMD_NUM_FEATURES floats each) of
 * the current model; outside it (beyond a tenth of the range on either side)
 * the learned mode runs the exhaustive sweep.  NULL arguments forget the
 * range (then a standardised feature beyond 4 counts as outside).  Status 2
 * without a model. */
perflibs_status_t perflibs_amd_set_selector_range(const float *feat_min, const float *feat_max);
/* Selection path (perflibs_amd_select_path) and number of variants timed by
 * the last perflibs_spmv_optimize of A; ntimed may be NULL. */
perflibs_status_t perflibs_amd_select_info(perflibs_spmat_t A, int *path, int *ntimed);
/* Load a model file written by the training driver (plsparse.selector).  The
 * library loads <dir of libplsparse_amd.so>/../selector/selector_mlp.txt (or the
 * file named by the environment variable PERFLIBS_AMD_SELECTOR_MODEL) by
 * itself at the first optimize() when no model was set or cleared. */
perflibs_status_t perflibs_amd_load_selector_model(const char *path);

/* MLP training/inference primitives on device buffers (FP32), used by the
 * training driver and by the gradient tests.  n = number of samples.
 *   forward : X[n x 12] -> out[n x K]; h1, h2 are [n x 64] activations saved
 *             for backward (may be NULL for inference).
 *   backward: given targets T[n x K] and mask M[n x K] (1 = measured), the
 *             masked mean squared error loss and its gradient w.r.t. the flat
 *             params into grad (same layout as params).
 *   adam    : one Adam step in place. */
perflibs_status_t perflibs_amd_mlp_forward(const float *params_dev,
                                           const float *X_dev, int n,
                                           float *h1_dev, float *h2_dev,
                                           float *out_dev);
perflibs_status_t perflibs_amd_mlp_backward(const float *params_dev,
                                            const float *X_dev, int n,
                                            const float *h1_dev,
                                            const float *h2_dev,
                                            const float *out_dev,
                                            const float *T_dev,
                                            const float *M_dev,
                                            float *grad_dev, float *loss_dev);
/* perflibs_amd_mlp_backward without any device allocation.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, training loops that forbid hipMalloc).  The
 * per-sample gradient scratch lives in a static device array of the library,
 * one per GPU, sized for n <= PERFLIBS_AMD_MLP_NOALLOC_MAX_SAMPLES (larger n:
 * status 1, nothing launched).  Calls that overlap on one GPU share that
 * array and race; calls on one stream are ordered and safe.  Results are
 * bit-identical to perflibs_amd_mlp_backward. */
#define PERFLIBS_AMD_MLP_NOALLOC_MAX_SAMPLES 4096
perflibs_status_t perflibs_amd_mlp_backward_noalloc(const float *params_dev,
                                                    const float *X_dev, int n,
                                                    const float *h1_dev,
                                                    const float *h2_dev,
                                                    const float *out_dev,
                                                    const float *T_dev,
                                                    const float *M_dev,
                                                    float *grad_dev, float *loss_dev);
perflibs_status_t perflibs_amd_mlp_adam(float *params_dev, const float *grad_dev,
                                        float *m_dev, float *v_dev, int step,
                                        float lr, float beta1, float beta2,
                                        float eps);

/* Device layout of a SELL-C-sigma handle, copied to host arrays sized from
 * perflibs_amd_sell_info (for bit-exact layout tests).  offsets are the
 * logical column offsets widened to int64. */
perflibs_status_t perflibs_amd_sell_info(perflibs_spmat_t A, perflibs_int_t *C,
                                         perflibs_int_t *sigma, int64_t *nC,
                                         int64_t *padded,
                                         perflibs_int_t *col_indx_bytes,
                                         perflibs_int_t *n_full_chunks,
                                         perflibs_int_t *ntidyup);
perflibs_status_t perflibs_amd_sell_export(perflibs_spmat_t A, int64_t *cs,
                                           int64_t *cl, int64_t *row_permd2in,
                                           int64_t *col_indx_min,
                                           int64_t *offsets, double *vals);

/* ---- batched products: one persistent launch, replayed from a hipGraph ----
 * A batch holds nprob independent products Y_p = alpha*A_p*X_p + beta*Y_p:
 * A_p a sparse handle (CSR, or BSR / SELL-C-sigma through their CSR view),
 * NOTRANS, all of one data type; X_p (n_p x k_p) and Y_p (m_p x k_p) row-major
 * device arrays, k_p in {1, 2, 4, 8, 16} (k_p = 1: SpMV).  beta == 0 never
 * reads Y; alpha == 0 gives beta*Y.  The matrices are copied to HBM at
 * create; later changes to the handles are not seen. */
typedef struct perflibs_amd_batch_s *perflibs_amd_batch_t;
/* precision: PERFLIBS_AMD_VALUES_NATIVE, or PERFLIBS_AMD_VALUES_BF16 for _s
 * handles (BF16 values, FP32 X/Y and accumulation). */
perflibs_status_t perflibs_amd_batch_create(perflibs_amd_batch_t *B, int64_t nprob,
                                            const perflibs_spmat_t *A, const perflibs_int_t *k,
                                            int precision);
perflibs_status_t perflibs_amd_batch_destroy(perflibs_amd_batch_t B);
/* Device pointers of every X_p and Y_p (arrays of nprob pointers in host
 * memory).  May be called again between executions; no recapture needed. */
perflibs_status_t perflibs_amd_batch_set_operands(perflibs_amd_batch_t B, const void *const *X,
                                                  void *const *Y);
/* autotune != 0: time candidate tile sizes and per-k row shapes by graph
 * replays (reading the current X, writing scratch outputs) and keep the
 * fastest; autotune == 0 keeps the current shape. */
perflibs_status_t perflibs_amd_batch_optimize(perflibs_amd_batch_t B, int autotune);
/* Set the shape by hand: tile_nnz non-zeros per tile at most, and for each k
 * class (k = 1, 2, 4, 8, 16) lanes per row (a power of two in [k/cpl, 64]) and
 * RHS columns per lane cpl (1, 2 or 4, at most k).  lanes / cpl may be NULL. */
perflibs_status_t perflibs_amd_batch_set_shape(perflibs_amd_batch_t B, perflibs_int_t tile_nnz,
                                               const perflibs_int_t *lanes,
                                               const perflibs_int_t *cpl);
/* Enqueue the whole batch on the library stream as one kernel launch, replayed
 * from a hipGraph (captured on first use and after a change of alpha, beta or
 * shape).  Not synchronized. */
perflibs_status_t perflibs_amd_batch_exec(perflibs_amd_batch_t B, double alpha, double beta);
/* Run the k = 1 problems of the batch as SELL-C-sigma tiles: C = 32 or 64 rows
 * per chunk (one lane per row), rows sorted by length inside sigma-row windows
 * (sigma 0: unsorted, else a multiple of C), the layout of
 * src/sell_c_sigma.cpp:145-336 per problem; C = 0 returns them to CSR row
 * tiles.  Built on the host from the batch's device copy; optimize(autotune)
 * also tries (0, 0), (32, 256), (64, 256) and (64, 1024) and keeps the fastest. */
perflibs_status_t perflibs_amd_batch_set_sell(perflibs_amd_batch_t B, perflibs_int_t C,
                                              perflibs_int_t sigma);
/* Current SELL-C-sigma choice of the k = 1 problems (C 0: CSR row tiles) and
 * the padded entries of their layout. */
perflibs_status_t perflibs_amd_batch_sell_info(perflibs_amd_batch_t B, perflibs_int_t *C,
                                               perflibs_int_t *sigma, int64_t *padded);
/* Current shape: tiles, tile_nnz, resident workgroups, lanes[5], cpl[5]. */
perflibs_status_t perflibs_amd_batch_info(perflibs_amd_batch_t B, int64_t *ntiles,
                                          perflibs_int_t *tile_nnz, perflibs_int_t *grid,
                                          perflibs_int_t *lanes, perflibs_int_t *cpl);

/* Last error message of the library (thread-unsafe, for diagnostics). */
const char *perflibs_amd_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* PERFLIBS_SPARSE_AMD_H */
