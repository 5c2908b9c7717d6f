/*
 * rivrs_ldlt.h -- C ABI of the MI355X numeric phase for rivrs-sparse's multifrontal
 * APTP LDL^T solver.
 *
 * The reference keeps the numeric phase behind three Rust functions, called from
 * SparseLDLT (sparse/src/symmetric/solver.rs:296-428):
 *   AptpNumeric::factor(&AptpSymbolic, &SparseColMat<usize,f64>, &AptpOptions, Option<&[f64]>)
 *                                                   (sparse/src/symmetric/numeric.rs:844-849)
 *   aptp_solve(&AptpSymbolic, &AptpNumeric, &mut [f64], &mut MemStack, Par)
 *                                                   (sparse/src/symmetric/solve.rs:57-63)
 *   aptp_solve_scratch(&AptpNumeric, usize)         (sparse/src/symmetric/solve.rs:150)
 * plus the accessors SparseLDLT::inertia / stats / per_supernode_stats
 * (solver.rs:504-543).  Each entry point below names the reference item it replaces.
 *
 * Conventions: plain C types only; caller-owned input buffers; library-owned opaque
 * handles; every fallible call returns an rv_status and records a message readable with
 * rv_last_error().  Matrices are full-symmetric CSC (both triangles stored) with 64-bit
 * column pointers and row indices, as SparseColMat<usize, f64>.  Zero pivots and unresolved
 * delayed columns are not errors (numeric.rs:829-835); they show up in rv_stats.
 * One GPU: the library uses the current HIP device.
 */
#ifndef RIVRS_LDLT_H
#define RIVRS_LDLT_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes mirror SparseError (sparse/src/error.rs:20-98). */
typedef enum rv_status {
    RV_OK = 0,
    RV_DIMENSION_MISMATCH = 1,    /* SparseError::DimensionMismatch */
    RV_NOT_SQUARE = 2,            /* SparseError::NotSquare */
    RV_INVALID_INPUT = 3,         /* SparseError::InvalidInput */
    RV_ANALYSIS_FAILURE = 6,      /* SparseError::AnalysisFailure */
    RV_SOLVE_BEFORE_FACTOR = 9,   /* SparseError::SolveBeforeFactor */
    RV_DEVICE_ERROR = 100,        /* HIP runtime failure */
    RV_NO_DEVICE = 101            /* no usable GPU: the numeric phase never runs on the CPU */
} rv_status;

typedef struct rv_symbolic_s rv_symbolic;
typedef struct rv_numeric_s rv_numeric;

/* Mirrors AptpOptions (sparse/src/symmetric/factor.rs:86-130) and FactorOptions
 * (solver.rs:72-116).  inner_block_size must be <= 32 on the device. */
typedef struct rv_factor_options {
    double threshold;             /* u in (0, 1], default 0.01 */
    double small;                 /* >= 0, default 1e-20 */
    int32_t outer_block_size;     /* nb, default 256 */
    int32_t inner_block_size;     /* ib, default 32 */
    int32_t failed_pivot_tpp;     /* 1 = FailedPivotMethod::Tpp (default), 0 = Pass */
    int32_t nemin;                /* amalgamation, default 32 */
    int32_t small_leaf_threshold; /* default 256, 0 disables the small-leaf TPP path */
    int32_t update_precision;     /* RV_PREC_*: arithmetic of the Schur-complement updates */
} rv_factor_options;

/* Arithmetic of the Schur-complement updates:
 *   RV_PREC_FP64 (default): FP64 throughout, the contribution GEMM on FP64 MFMA;
 *   RV_PREC_FP64_EXACT: FP64 with the contribution GEMM accumulated in the reference's fixed
 *     order (each entry summed over ascending pivot columns, then subtracted), so every factor
 *     entry is bitwise reproducible against a sequential FP64 implementation;
 *   RV_PREC_FP32 / RV_PREC_BF16: contribution and outer-block updates on FP32 / BF16 MFMA with
 *     FP32 accumulation, updates beyond the current outer block deferred to one rank-nb GEMM;
 *     pivot tests, D and the solve stay FP64; full accuracy through rv_solve_refine;
 *   RV_PREC_FP64_MFMA: FP64 throughout with the outer-block updates, like the contribution
 *     GEMM, deferred to one rank-nb FP64 MFMA GEMM per outer block (RV_PREC_FP64 runs them in the
 *     reference's order on the vector units, so that pivot decisions within a front follow the
 *     reference's rounding; here they may differ where candidates tie to rounding).
 * The reference's factor is FP64 throughout (factor.rs) and has no such switch. */
enum { RV_PREC_FP64 = 0, RV_PREC_FP32 = 1, RV_PREC_BF16 = 2, RV_PREC_FP64_EXACT = 3, RV_PREC_FP64_MFMA = 4 };

/* Ordering for rv_analyze (OrderingStrategy, solver.rs:42-55).  METIS is not available:
 * RV_ORDER_ND_GRAPH is the library's own nested dissection; RV_ORDER_ND_GRID is nested
 * dissection by planes for matrices on an nx x ny x nz grid numbered x-fastest (extra
 * unknowns beyond the grid are ordered last). */
enum { RV_ORDER_NATURAL = 0, RV_ORDER_USER = 1, RV_ORDER_ND_GRAPH = 2, RV_ORDER_ND_GRID = 3,
       RV_ORDER_MATCH_ND = 4 /* MatchOrderMetis: needs values, rv_analyze_with_matrix only */ };

/* Message of the last failing call on this thread ("" if none). */
const char *rv_last_error(void);

/* Fills the defaults of AptpOptions::default() (factor.rs:117-131). */
void rv_factor_options_default(rv_factor_options *opts);

/* SparseLDLT::analyze (solver.rs:195-259) / AptpSymbolic::analyze (symbolic.rs:248-310):
 * ordering, elimination tree, supernode partition.  user_perm (perm_fwd[new] = old) for
 * RV_ORDER_USER, grid_dims[3] for RV_ORDER_ND_GRID; otherwise may be NULL. */
int rv_analyze(int64_t n, const int64_t *col_ptr, const int64_t *row_idx, int32_t ordering,
               const int32_t *user_perm, const int32_t *grid_dims, rv_symbolic **out);

/* SparseLDLT::analyze_with_matrix (solver.rs:261-285).  With RV_ORDER_MATCH_ND the ordering is
 * rv_match_order's and elim_scaling (length n, may be NULL) receives the MC64 scaling in
 * elimination order (elim_scaling[k] = scaling[perm_fwd[k]], as solver.rs:276-279), with
 * *has_scaling = 1; any other ordering behaves as rv_analyze and sets *has_scaling = 0. */
int rv_analyze_with_matrix(int64_t n, const int64_t *col_ptr, const int64_t *row_idx,
                           const double *values, int32_t ordering, const int32_t *user_perm,
                           const int32_t *grid_dims, rv_symbolic **out, double *elim_scaling,
                           int32_t *has_scaling);

/* mc64_matching (matching.rs:121-248), job MaximumProduct: maximum-product matching of a
 * symmetric matrix in CSC (one triangle or both; explicit zeros ignored) and the symmetric
 * scaling S of Duff & Pralet (A_scaled = S A S has |entries| <= 1 between matched indices and
 * unit matched entries).  matching[i] = column matched to row i (unmatched rows take the free
 * columns in ascending order); is_matched[i] = 1 if row i has a matching edge; *matched =
 * cardinality.  Any output may be NULL.  Errors: n = 0 or non-finite values (InvalidInput). */
int rv_mc64_matching(int64_t n, const int64_t *col_ptr, const int64_t *row_idx, const double *values,
                     int32_t *matching, double *scaling, int32_t *is_matched, int64_t *matched);

/* match_order_metis (metis.rs:118-190) with the library's nested dissection in place of METIS:
 * order[new] = old (matched 2-cycles at adjacent positions), scaling in original index order,
 * info[4] = {matched, condensed dimension, singletons, two-cycles}.  Outputs may be NULL. */
int rv_match_order(int64_t n, const int64_t *col_ptr, const int64_t *row_idx, const double *values,
                   int32_t *order, double *scaling, int64_t *info);

/* Import the supernode partition of an existing AptpSymbolic, i.e. what build_supernode_info
 * reads (numeric.rs:2179-2236): per supernode [sn_begin, sn_end), sorted pattern (CSR
 * pat_ptr/pat, permuted global indices), parent (-1 = root, AptpSymbolic::supernode_parent
 * symbolic.rs:576-598), and perm_fwd[new] = old (AptpSymbolic::perm_vecs). */
int rv_symbolic_from_structure(int64_t n, int32_t nsn, const int32_t *sn_begin,
                               const int32_t *sn_end, const int64_t *pat_ptr, const int32_t *pat,
                               const int32_t *parent, const int32_t *perm_fwd, rv_symbolic **out);

/* out[3] = {n, number of supernodes, predicted nnz(L)} (AptpSymbolic::statistics). */
int rv_symbolic_info(const rv_symbolic *sym, int64_t *out);

/* Export the partition (arrays sized from rv_symbolic_info; pat sized pat_ptr[nsn]; pass
 * pat == NULL first to obtain pat_ptr). */
int rv_symbolic_export(const rv_symbolic *sym, int32_t *sn_begin, int32_t *sn_end,
                       int64_t *pat_ptr, int32_t *pat, int32_t *parent, int32_t *perm_fwd);

void rv_symbolic_free(rv_symbolic *sym);

/* amalgamate (sparse/src/symmetric/amalgamation.rs:115-231), the host step AptpNumeric::factor
 * runs first (numeric.rs:862-868), on a bare partition.  Outputs are sized by the input count
 * (pat_out by pat_ptr[nsn]); owned column ranges come back as CSR (rng_ptr, rng_lo, rng_hi).
 * Returns the number of supernodes after merging, or -1 on invalid input. */
int32_t rv_amalgamate(int32_t nsn, const int32_t *sn_begin, const int32_t *sn_end,
                      const int64_t *pat_ptr, const int32_t *pat, const int32_t *parent,
                      int32_t nemin, int32_t *begin_out, int32_t *end_out, int32_t *parent_out,
                      int64_t *rng_ptr, int32_t *rng_lo, int32_t *rng_hi, int64_t *pat_ptr_out,
                      int32_t *pat_out);

/* AptpNumeric::factor (numeric.rs:844-1023) on the GPU.  col_ptr/row_idx/values are host
 * arrays (values length col_ptr[n]); scaling (nullable) is in elimination order as passed by
 * SparseLDLT::factor (solver.rs:296-324).  The factors stay resident in device memory. */
int rv_factor(rv_symbolic *sym, const int64_t *col_ptr, const int64_t *row_idx,
              const double *values, const double *scaling, const rv_factor_options *opts,
              rv_numeric **out);

/* As rv_factor with values (and scaling) already in device memory.  The handle keeps using
 * d_values (not a copy) for the residuals of rv_solve_refine*: the caller keeps that buffer
 * valid and unchanged until the handle is freed or refactored with other values.  d_scaling is
 * copied.  A numeric handle holds its own reference to the prepared structure: it stays valid
 * when sym is freed or factored again with other options or another pattern. */
int rv_factor_device(rv_symbolic *sym, const int64_t *col_ptr, const int64_t *row_idx,
                     const double *d_values, const double *d_scaling,
                     const rv_factor_options *opts, rv_numeric **out);

/* SparseLDLT::refactor (solver.rs:335-341): new values, same pattern and options; later
 * refinements use the new values.  rv_refactor copies the host values into a buffer the
 * handle owns.  rv_refactor_device copies d_values into that buffer when the handle owns one
 * (made by rv_factor or an earlier rv_refactor); a handle made by rv_factor_device adopts
 * d_values instead, under the same lifetime rule as rv_factor_device. */
int rv_refactor(rv_numeric *num, const double *values);
int rv_refactor_device(rv_numeric *num, const double *d_values);

/* SparseLDLT::solve_in_place (solver.rs:370-428) with aptp_solve (solve.rs:57-144) for
 * nrhs columns (ldrhs >= n), original coordinates: permute, scale, L, D, L^T, scale,
 * unpermute.  rv_solve takes host memory, rv_solve_device device memory. */
int rv_solve(rv_numeric *num, double *rhs, int64_t ldrhs, int32_t nrhs);
int rv_solve_device(rv_numeric *num, double *d_rhs, int64_t ldrhs, int32_t nrhs);

/* No counterpart in the reference: rv_solve_device for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same kernels and results as rv_solve_device; its workspace is a fixed device array of
 * RV_SOLVE_NOALLOC_BYTES (64 MiB) per GPU, held from the first call on a GPU until the process
 * ends, so the call allocates nothing.  A solve whose rv_solve_scratch(num, nrhs) exceeds it
 * returns RV_INVALID_INPUT.  Every call on a GPU shares that GPU's array (the solve clears what
 * it needs at its start): calls that overlap on one GPU, from any handle or thread, race. */
enum { RV_SOLVE_NOALLOC_BYTES = 64 << 20 };
int rv_solve_device_noalloc(rv_numeric *num, double *d_rhs, int64_t ldrhs, int32_t nrhs);

/* Solve with FP64 iterative refinement (one right-hand side, original coordinates):
 * x0 = solve(b), x_{k+1} = x_k + solve(b - A x_k), where A is the matrix of the last
 * factorization, until the scaled backward error ||A x - b|| / (||A||_F ||x|| + ||b||)
 * (sparse_backward_error, validate.rs:209-255) is <= tol or max_iter corrections were
 * applied.  The reference solve has no refinement (solve.rs:57-63); this completes the
 * reduced-precision update forms (RV_PREC_FP32 / RV_PREC_BF16).  iters and backward_error
 * (of the returned x) may be NULL.  rv_solve_refine takes host memory, the _device form
 * device memory (b and x must not overlap). */
int rv_solve_refine(rv_numeric *num, const double *b, double *x, int32_t max_iter, double tol,
                    int32_t *iters, double *backward_error);
int rv_solve_refine_device(rv_numeric *num, const double *d_b, double *d_x, int32_t max_iter,
                           double tol, int32_t *iters, double *backward_error);

/* aptp_solve_scratch (solve.rs:150): bytes of device workspace a solve of nrhs columns uses. */
int64_t rv_solve_scratch(const rv_numeric *num, int32_t nrhs);

/* SparseLDLT::inertia (solver.rs:504-519): out[3] = {positive, negative, zero}. */
int rv_inertia(const rv_numeric *num, int64_t *out);

/* SparseLDLT::stats (FactorizationStats, numeric.rs:604-719): out[10] = {total_1x1,
 * total_2x2, total_delayed, zero_pivots, max_front_size, supernodes_before_amalgamation,
 * supernodes_after_amalgamation, merges_performed, small_leaf_subtrees, small_leaf_nodes}. */
int rv_stats(const rv_numeric *num, int64_t *out);

/* SparseLDLT::per_supernode_stats (PerSupernodeStats): out[8] = {front_size,
 * num_fully_summed, num_eliminated, num_delayed, num_1x1, num_2x2, parent, in_small_leaf}. */
int32_t rv_num_supernodes(const rv_numeric *num);

/* Per-front update precision (no counterpart in the reference, whose factor is FP64 only):
 * prec[s] (one RV_PREC_* code per supernode, rv_num_supernodes entries) sets the arithmetic of
 * supernode s's Schur-complement updates from the next rv_refactor / factorization on; NULL
 * restores the options' update_precision for every front.  Pivoting, D and the solve stay FP64.
 * rv_front_precision reads the current assignment. */
int rv_set_front_precision(rv_numeric *num, const int8_t *prec);
int rv_front_precision(const rv_numeric *num, int8_t *out);

/* Classifier features of every front from the last factorization: RV_MLP_FEATURES floats per
 * supernode, row-major: log2(1+m), log2(1+k), log2(1+m-k), k/m, log2(1+contribution flops),
 * log2(1+front flops), tree level / (levels-1), log2(1+children), log2(1+sum of children's
 * contribution entries), delayed-in / k, delayed-out / k, 2x2 fraction, log10(1+max|L|),
 * negative-pivot fraction, is-root, m / m_parent. */
int rv_front_features(const rv_numeric *num, float *out);

/* The per-front precision classifier (no counterpart in the reference): a multilayer perceptron
 * RV_MLP_FEATURES -> 64 -> 64 -> RV_MLP_CLASSES (ReLU, softmax; classes 0 = BF16, 1 = FP32,
 * 2 = FP64), FP32 parameters in one array of RV_MLP_PARAMS: W1[16][64], b1[64], W2[64][64],
 * b2[64], W3[64][3], b3[3].  All arrays are host memory; the work runs on the GPU.
 *   rv_mlp_forward: class probabilities (n x 3) of n feature rows (n x 16);
 *   rv_mlp_loss_grad: class-weighted mean cross-entropy sum w[y](-log p_y) / sum w[y] and its
 *     gradient with respect to every parameter (class_weight NULL = all 1);
 *   rv_mlp_train: `epochs` full-batch Adam steps (lr, beta 0.9/0.999), params updated in place,
 *     final_loss (may be NULL) = loss of the returned parameters. */
enum { RV_MLP_FEATURES = 16, RV_MLP_CLASSES = 3, RV_MLP_PARAMS = 5443 };
int rv_mlp_forward(const float *params, int64_t n, const float *x, float *probs);
int rv_mlp_loss_grad(const float *params, int64_t n, const float *x, const int32_t *labels,
                     const float *class_weight, float *loss, float *grad);
int rv_mlp_train(float *params, int64_t n, const float *x, const int32_t *labels, const float *class_weight,
                 int32_t epochs, float lr, float *final_loss);
int rv_supernode_stats(const rv_numeric *num, int32_t s, int64_t *out, double *max_l);

/* Owned columns and pattern of amalgamated supernode s; out[2] = {n_owned, n_pattern}. */
int rv_supernode_structure(const rv_numeric *num, int32_t s, int32_t *owned, int32_t *pattern,
                           int64_t *out);

/* FrontFactors of supernode s (numeric.rs:543-596) copied to the host: l11[ne*ne],
 * l21[r*ne] (column-major, r = front_size - ne), dkind/ddiag/doff[ne] (dkind 1 = 1x1,
 * 2 = both columns of a 2x2, off-diagonal at the first column), col_indices[ne],
 * row_indices[r], local_perm[k].  Any pointer may be NULL. */
int rv_export_front(const rv_numeric *num, int32_t s, double *l11, double *l21, int8_t *dkind,
                    double *ddiag, double *doff, int32_t *col_indices, int32_t *row_indices,
                    int32_t *local_perm);

/* ContributionBlock of supernode s (numeric.rs:516-523) as extract_contribution builds it
 * (numeric.rs:2739-2809) after compute_contribution_gemm (factor.rs:1590-1652), copied to the
 * host: out[2] = {size, num_delayed}; data[size*size] column-major (lower triangle holds the
 * block), row_indices[size] = [delayed columns, pattern rows] (global permuted indices).
 * size = 0 for a root or a fully eliminated front.  data/row_indices may be NULL.  A non-root
 * front's working memory is handed to later levels once its parent's level has assembled it:
 * RV_INVALID_INPUT when a later level has overwritten it (with RV_KEEP_FRONTS=1 in the
 * environment when the numeric handle is made, every front keeps memory of its own). */
int rv_export_contribution(const rv_numeric *num, int32_t s, double *data, int32_t *row_indices,
                           int64_t *out);

/* Dense APTP kernel on the GPU, aptp_factor_in_place (factor.rs:301-404): a is an m x m
 * column-major host matrix (lower triangle read and overwritten), p fully-summed columns.
 * Outputs as the reference's AptpFactorResult: perm[m], dkind/ddiag/doff[p] (0 = delayed),
 * stats[4] = {num_eliminated, num_1x1, num_2x2, num_delayed}, max_l. */
int rv_dense_aptp_factor(double *a, int64_t lda, int32_t m, int32_t p,
                         const rv_factor_options *opts, int32_t *perm, int8_t *dkind,
                         double *ddiag, double *doff, int64_t *stats, double *max_l);

/* Graph replay (no counterpart in the reference): with on != 0, a refactorization whose levels
 * all run on unchanged plans (same pattern and options, the previous factorization's delays
 * reproduced, the same device value/scaling pointers) is one launch of a hipGraph holding the
 * whole level sequence, captured at the first such call.  Per-launch timing (rv_set_timing)
 * records nothing for replayed factorizations.  rv_graph_launches counts the replays. */
int rv_set_graph_mode(rv_numeric *num, int32_t on);
int64_t rv_graph_launches(const rv_numeric *num);

/* Persistent outer-block launches (no counterpart in the reference; a diagnostic like
 * rv_graph_launches): the APTP inner blocks of a level's outer block run as one launch (a
 * leader workgroup per front for the pivot chain, panel workgroups for the rows) when all of
 * the level's workgroups fit on the device at once, and as three launches per inner block
 * otherwise.  Counts those launches enqueued by factorizations outside graph replays; the
 * environment variable RV_APTP_CHAIN=0 (read when the numeric handle is created) disables them. */
int64_t rv_chain_launches(const rv_numeric *num);

/* Diagnostic (no counterpart in the reference): outer-block trailing updates that ran on a second
 * stream beside the next outer block (the outer look-ahead of two_level_factor's deferred updates,
 * factor.rs:2108-2252, done in the same per-column order).  A block is run so when no front of
 * its level had failed columns in that block in the previous factorization; a factorization that
 * meets an unplanned failure is repeated in stream order.  RV_OUTER_LOOKAHEAD=0 (read when the
 * numeric handle is created) disables it. */
int64_t rv_lookahead_launches(const rv_numeric *num);

/* Device memory the numeric handle holds for fronts, plans and working space, in bytes
 * (FrontFactors storage, numeric.rs:543-596, lives here). */
int64_t rv_device_bytes(const rv_numeric *num);

/* Device timings of the last factor/solve with timing enabled (milliseconds / flops):
 * out[6] = {schur_ms, schur_flops, schur_launches, and the same three for the contribution
 * GEMMs of levels whose largest contribution block has order >= 2048}. */
int rv_set_timing(rv_numeric *num, int32_t on);
int rv_timings(const rv_numeric *num, double *out);

void rv_numeric_free(rv_numeric *num);

#ifdef __cplusplus
}
#endif
#endif
