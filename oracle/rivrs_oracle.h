/*
 * rivrs_oracle.h -- TEST INFRASTRUCTURE ONLY (the parity checker, never the product).
 *
 * CPU FP64 restatement of rivrs-sparse's numeric phase (multifrontal APTP LDL^T
 * factorization and supernodal solve).  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this library.  The product path
 * (rivrs-linalg_amd/) never links, loads or calls it.
 *
 * Every function follows a reference file:line, cited in rivrs_oracle.cpp.
 * Index conventions: full-symmetric CSC with 64-bit column pointers/row indices
 * (mirrors SparseColMat<usize, f64>); supernode structure arrays are 32-bit.
 */
#ifndef RIVRS_ORACLE_H
#define RIVRS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors AptpOptions (sparse/src/symmetric/factor.rs:86-130). */
typedef struct orc_options {
    double threshold;            /* u, default 0.01 */
    double small;                /* default 1e-20 */
    int32_t outer_block_size;    /* nb, default 256 */
    int32_t inner_block_size;    /* ib, default 32 */
    int32_t failed_pivot_tpp;    /* 1 = FailedPivotMethod::Tpp (default), 0 = Pass */
    int32_t nemin;               /* default 32 */
    int32_t small_leaf_threshold;/* default 256, 0 disables */
} orc_options;

/* Dense kernel: aptp_factor_in_place (factor.rs:301-404) on a column-major m x m
 * matrix (lower triangle), p fully-summed columns.
 * Outputs: perm[m] (col_order), dkind[p] (0 delayed, 1 1x1, 2 2x2), ddiag[p], doff[p]
 * (off-diagonal stored at the first column of a 2x2), stats[4] = {ne, n1x1, n2x2, ndelayed},
 * max_l = AptpStatistics::max_l_entry.  Returns 0, or -1 on InvalidInput. */
int orc_aptp_factor_in_place(double *a, int64_t lda, int32_t m, int32_t p,
                             const orc_options *o, int32_t *perm, int8_t *dkind,
                             double *ddiag, double *doff, int64_t *stats, double *max_l);

/* Multifrontal factorization: AptpNumeric::factor (numeric.rs:844-1023).
 * Structure is the pre-amalgamation supernode partition (what build_supernode_info,
 * numeric.rs:2179-2236, reads from AptpSymbolic): sn_begin/sn_end, sorted patterns
 * (CSR pat_ptr/pat, global permuted indices), parent (-1 = root), perm_fwd[new] = old.
 * scaling (nullable) is in elimination order.  Returns an opaque handle or NULL. */
void *orc_factor(int64_t n, const int64_t *col_ptr, const int64_t *row_idx, const double *val,
                 int32_t nsn, const int32_t *sn_begin, const int32_t *sn_end,
                 const int64_t *pat_ptr, const int32_t *pat, const int32_t *parent,
                 const int32_t *perm_fwd, const double *scaling, const orc_options *o);

/* SparseLDLT::solve_in_place (solver.rs:370-428) for nrhs columns, original coordinates. */
int orc_solve(void *h, double *rhs, int64_t ldrhs, int32_t nrhs);

/* Inertia summed over fronts (solver.rs:504-519): out[3] = {pos, neg, zero}. */
void orc_inertia(void *h, int64_t *out);

/* FactorizationStats (numeric.rs:604-719): out[10] = {total_1x1, total_2x2, total_delayed,
 * zero_pivots, max_front_size, sn_before, sn_after, merges, small_leaf_subtrees,
 * small_leaf_nodes}. */
void orc_stats(void *h, int64_t *out);

int32_t orc_num_supernodes(void *h);

/* Per-supernode: out[8] = {front_size, num_fully_summed, num_eliminated, num_delayed,
 * num_1x1, num_2x2, parent, in_small_leaf}; max_l written to *max_l. */
void orc_supernode_info(void *h, int32_t s, int64_t *out, double *max_l);

/* Amalgamated structure of supernode s: owned columns (concatenated owned_ranges) and
 * pattern. Returns counts in out[2] = {n_owned, n_pattern}; arrays may be NULL to query. */
void orc_supernode_structure(void *h, int32_t s, int32_t *owned, int32_t *pattern, int64_t *out);

/* FrontFactors of supernode s (numeric.rs:543-596).  Sizes: ne, r = row count.
 * l11[ne*ne] column-major, l21[r*ne] column-major, dkind/ddiag/doff[ne],
 * col_indices[ne], row_indices[r], local_perm[k].  Any pointer may be NULL. */
void orc_front_factors(void *h, int32_t s, double *l11, double *l21, int8_t *dkind,
                       double *ddiag, double *doff, int32_t *col_indices,
                       int32_t *row_indices, int32_t *local_perm);

void orc_free(void *h);

/* Keep every front's contribution block (ContributionBlock, numeric.rs:516-523, as built by
 * compute_contribution_gemm factor.rs:1590-1652 + extract_contribution numeric.rs:2739-2809)
 * in factorizations started after the call, for inspection by orc_contribution. */
void orc_set_keep_contributions(int32_t on);

/* Kept contribution block of supernode s: data[size*size] column-major (lower triangle),
 * rows[size] = global permuted indices [delayed columns, pattern rows], *num_delayed.
 * Returns size, or -1 if s has none (root, fully eliminated, or not kept). */
int64_t orc_contribution(void *h, int32_t s, double *data, int32_t *rows, int64_t *num_delayed);

/* extend_add_mapped (numeric.rs:2507-2549): parent (column-major, leading dimension ldp) +=
 * child (size x size lower triangle) scattered through row_map. */
void orc_extend_add_mapped(double *parent, int64_t ldp, const double *child, int32_t size,
                           const uint32_t *row_map);

/* Thread count of later calls (OpenMP): concurrent small-leaf subtrees and fronts of a wave
 * (numeric.rs:2111-2142), large products split by column; results do not depend on it. */
void orc_set_threads(int32_t threads);
int32_t orc_max_threads(void);

/* sparse_backward_error (validate.rs:209-255). */
double orc_backward_error(int64_t n, const int64_t *col_ptr, const int64_t *row_idx,
                          const double *val, const double *x, const double *b);

/* amalgamate (amalgamation.rs:115-231) on a bare structure; for the reference's KATs.
 * Outputs (caller sizes them >= input): parent_out, owned (CSR of ranges: rng_ptr, rng_lo,
 * rng_hi), pattern (CSR pat_ptr_out/pat_out), begin_out/end_out. Returns new count. */
int32_t orc_amalgamate(int32_t nsn, const int32_t *sn_begin, const int32_t *sn_end,
                       const int64_t *pat_ptr, const int32_t *pat, const int32_t *parent,
                       int32_t nemin, int32_t *begin_out, int32_t *end_out,
                       int32_t *parent_out, int64_t *rng_ptr, int32_t *rng_lo,
                       int32_t *rng_hi, int64_t *pat_ptr_out, int32_t *pat_out);

#ifdef __cplusplus
}
#endif
#endif
