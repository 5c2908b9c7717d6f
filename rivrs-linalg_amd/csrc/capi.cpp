// capi.cpp -- the extern "C" boundary (include/rivrs_ldlt.h).  Translates C arguments to the
// host driver, converts exceptions into rv_status codes, and keeps the last error message.
#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <string>

#include "../../include/rivrs_ldlt.h"
#include "rv_mlp.h"
#include "rv_numeric.h"

using namespace rv;

struct rv_symbolic_s {
    Symbolic sym;
    // shared with every numeric handle built on it: a later factorization with other options or
    // another pattern replaces it here without touching the handles that use the old one
    std::shared_ptr<Prepared> prep;
    // key of the prepared structure
    int nemin = -1, small_leaf = -1;
    i64 nnz = -1;
    std::vector<i64> col_ptr, row_idx;
};

struct rv_numeric_s {
    std::shared_ptr<const Prepared> prep;  // the structure the factors were built on
    i64 n = 0;
    FactorOptions opts;
    hipStream_t stream = nullptr;
    std::unique_ptr<NumericGPU> gpu;
    double *d_val = nullptr;
    double *d_scal = nullptr;
    bool owns_val = false;
    i64 nnz = 0;
    // the factored matrix's pattern, for the residuals of rv_solve_refine (device copy on demand)
    std::vector<i64> cp, ri;
    i64 *d_cp = nullptr, *d_ri = nullptr;
    DevBuf refine_work, refine_io;
    ~rv_numeric_s() {
        gpu.reset();
        if (d_val && owns_val) (void)hipFree(d_val);
        if (d_scal) (void)hipFree(d_scal);
        if (d_cp) (void)hipFree(d_cp);
        if (d_ri) (void)hipFree(d_ri);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

struct StatusError : std::runtime_error {
    int code;
    StatusError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

template <class F>
int guarded(F f) {
    try {
        g_err.clear();
        f();
        return RV_OK;
    } catch (const StatusError &e) {
        return fail(e.code, e.what());
    } catch (const HipError &e) {
        return fail(RV_DEVICE_ERROR, e.what());
    } catch (const std::bad_alloc &) {
        return fail(RV_DEVICE_ERROR, "host allocation failed");
    } catch (const std::exception &e) {
        return fail(RV_DEVICE_ERROR, e.what());
    }
}

void require_device() {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0)
        throw StatusError(RV_NO_DEVICE,
                          "no HIP device available: the numeric phase runs only on the GPU");
}

FactorOptions to_opts(const rv_factor_options *o) {
    FactorOptions f;
    if (o) {
        f.threshold = o->threshold;
        f.small = o->small;
        f.outer_block_size = o->outer_block_size;
        f.inner_block_size = o->inner_block_size;
        f.failed_pivot_tpp = o->failed_pivot_tpp;
        f.nemin = o->nemin;
        f.small_leaf_threshold = o->small_leaf_threshold;
        f.update_precision = o->update_precision;
    }
    // validation as aptp_factor_in_place (factor.rs:306-347)
    if (!(f.threshold > 0.0 && f.threshold <= 1.0))
        throw StatusError(RV_INVALID_INPUT, "threshold must be in (0, 1], got " + std::to_string(f.threshold));
    if (f.small < 0.0) throw StatusError(RV_INVALID_INPUT, "small must be >= 0.0");
    if (f.outer_block_size <= 0) throw StatusError(RV_INVALID_INPUT, "outer_block_size must be > 0");
    if (f.inner_block_size <= 0) throw StatusError(RV_INVALID_INPUT, "inner_block_size must be > 0");
    if (f.inner_block_size > f.outer_block_size)
        throw StatusError(RV_INVALID_INPUT, "inner_block_size must be <= outer_block_size");
    if (f.inner_block_size > MAX_IB)
        throw StatusError(RV_INVALID_INPUT, "inner_block_size must be <= 32 on the device");
    if (f.nemin < 1) f.nemin = 1;
    if (f.small_leaf_threshold < 0) throw StatusError(RV_INVALID_INPUT, "small_leaf_threshold must be >= 0");
    if (f.update_precision < RV_PREC_FP64 || f.update_precision > RV_PREC_FP64_MFMA)
        throw StatusError(RV_INVALID_INPUT,
                          "update_precision must be RV_PREC_FP64, RV_PREC_FP64_EXACT, RV_PREC_FP64_MFMA, RV_PREC_FP32 or "
                          "RV_PREC_BF16");
    return f;
}

void check_csc(const rv_symbolic *sym, const i64 *cp, const i64 *ri) {
    const i64 n = sym->sym.n;
    if (!cp || !ri) throw StatusError(RV_INVALID_INPUT, "null CSC structure");
    if (cp[0] != 0) throw StatusError(RV_INVALID_INPUT, "col_ptr[0] must be 0");
    for (i64 j = 0; j < n; ++j)
        if (cp[j + 1] < cp[j]) throw StatusError(RV_INVALID_INPUT, "col_ptr not monotone");
    for (i64 p = 0; p < cp[n]; ++p)
        if (ri[p] < 0 || ri[p] >= n) throw StatusError(RV_INVALID_INPUT, "row index out of range");
}

// amalgamation + assembly maps for this pattern and options (numeric.rs:862-881), cached
void ensure_prepared(rv_symbolic *s, const i64 *cp, const i64 *ri, const FactorOptions &o) {
    const i64 n = s->sym.n;
    const i64 nnz = cp[n];
    bool same = s->prep && s->nemin == o.nemin && s->small_leaf == o.small_leaf_threshold &&
                s->nnz == nnz && (i64)s->col_ptr.size() == n + 1 &&
                std::memcmp(s->col_ptr.data(), cp, (n + 1) * sizeof(i64)) == 0 &&
                std::memcmp(s->row_idx.data(), ri, nnz * sizeof(i64)) == 0;
    if (same) return;
    check_csc(s, cp, ri);
    auto fresh = std::make_shared<Prepared>();
    fresh->build(s->sym, cp, ri, o.nemin, o.small_leaf_threshold);
    s->prep = std::move(fresh);
    s->nemin = o.nemin;
    s->small_leaf = o.small_leaf_threshold;
    s->nnz = nnz;
    s->col_ptr.assign(cp, cp + n + 1);
    s->row_idx.assign(ri, ri + nnz);
}

RefineResult do_refine(rv_numeric *num, const double *d_b, double *d_x, int32_t max_iter, double tol) {
    if (max_iter < 0) throw StatusError(RV_INVALID_INPUT, "max_iter must be >= 0");
    if (!(tol >= 0.0)) throw StatusError(RV_INVALID_INPUT, "tol must be >= 0");
    const i64 n = num->n;
    if (!num->d_cp) {
        RV_HIP(hipMalloc(&num->d_cp, (size_t)(n + 1) * sizeof(i64)));
        RV_HIP(hipMalloc(&num->d_ri, (size_t)std::max<i64>(num->nnz, 1) * sizeof(i64)));
        RV_HIP(hipMemcpy(num->d_cp, num->cp.data(), (size_t)(n + 1) * sizeof(i64), hipMemcpyHostToDevice));
        if (num->nnz)
            RV_HIP(hipMemcpy(num->d_ri, num->ri.data(), (size_t)num->nnz * sizeof(i64), hipMemcpyHostToDevice));
    }
    return solve_refine(*num->gpu, num->stream, n, num->nnz, num->d_cp, num->d_ri, num->d_val, num->d_scal, d_b,
                        d_x, max_iter, tol, num->refine_work);
}


int do_factor(rv_symbolic *sym, const i64 *cp, const i64 *ri, const double *values,
              const double *scaling, bool device, const rv_factor_options *opts,
              rv_numeric **out) {
    return guarded([&] {
        if (!sym || !out) throw StatusError(RV_INVALID_INPUT, "null argument");
        *out = nullptr;
        FactorOptions o = to_opts(opts);
        check_csc(sym, cp, ri);
        require_device();
        ensure_prepared(sym, cp, ri, o);
        const i64 n = sym->sym.n;
        const i64 nnz = cp[n];
        std::unique_ptr<rv_numeric_s> num(new rv_numeric_s());
        num->prep = sym->prep;
        num->n = n;
        num->opts = o;
        num->nnz = nnz;
        num->cp.assign(cp, cp + n + 1);
        num->ri.assign(ri, ri + nnz);
        RV_HIP(hipStreamCreate(&num->stream));
        if (device) {
            num->d_val = const_cast<double *>(values);
            num->owns_val = false;
        } else {
            RV_HIP(hipMalloc(&num->d_val, std::max<i64>(nnz, 1) * sizeof(double)));
            num->owns_val = true;
            if (nnz)
                RV_HIP(hipMemcpy(num->d_val, values, nnz * sizeof(double), hipMemcpyHostToDevice));
        }
        if (scaling) {
            RV_HIP(hipMalloc(&num->d_scal, std::max<i64>(n, 1) * sizeof(double)));
            RV_HIP(hipMemcpy(num->d_scal, scaling, n * sizeof(double),
                             device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice));
        }
        num->gpu.reset(new NumericGPU(num->prep.get(), o, num->stream));
        num->gpu->factor(num->d_val, num->d_scal);
        RV_HIP(hipStreamSynchronize(num->stream));
        *out = num.release();
    });
}

}  // namespace

extern "C" {

const char *rv_last_error(void) { return g_err.c_str(); }

void rv_factor_options_default(rv_factor_options *o) {
    if (!o) return;
    o->threshold = 0.01;
    o->small = 1e-20;
    o->outer_block_size = 256;
    o->inner_block_size = 32;
    o->failed_pivot_tpp = 1;
    o->nemin = 32;
    o->small_leaf_threshold = 256;
    o->update_precision = RV_PREC_FP64;
}

int rv_analyze(int64_t n, const int64_t *col_ptr, const int64_t *row_idx, int32_t ordering,
               const int32_t *user_perm, const int32_t *grid_dims, rv_symbolic **out) {
    return guarded([&] {
        if (!out) throw StatusError(RV_INVALID_INPUT, "null output");
        *out = nullptr;
        if (n < 0) throw StatusError(RV_INVALID_INPUT, "negative dimension");
        if (!col_ptr || (n > 0 && !row_idx)) throw StatusError(RV_INVALID_INPUT, "null CSC structure");
        if (ordering == RV_ORDER_MATCH_ND)  // solver.rs:241-245
            throw StatusError(RV_ANALYSIS_FAILURE,
                              "MatchOrderMetis requires numeric matrix values; use analyze_with_matrix() instead");
        std::unique_ptr<rv_symbolic_s> s(new rv_symbolic_s());
        std::string err;
        if (!analyze(n, col_ptr, row_idx, ordering, user_perm, grid_dims, s->sym, err))
            throw StatusError(ordering == RV_ORDER_USER && err.find("perm") != std::string::npos
                                  ? RV_DIMENSION_MISMATCH
                                  : RV_ANALYSIS_FAILURE,
                              err);
        *out = s.release();
    });
}

int rv_analyze_with_matrix(int64_t n, const int64_t *col_ptr, const int64_t *row_idx,
                           const double *values, int32_t ordering, const int32_t *user_perm,
                           const int32_t *grid_dims, rv_symbolic **out, double *elim_scaling,
                           int32_t *has_scaling) {
    return guarded([&] {
        if (has_scaling) *has_scaling = 0;
        if (ordering != RV_ORDER_MATCH_ND) {
            const int rc = rv_analyze(n, col_ptr, row_idx, ordering, user_perm, grid_dims, out);
            if (rc != RV_OK) throw StatusError(rc, rv_last_error());
            return;
        }
        if (!out) throw StatusError(RV_INVALID_INPUT, "null output");
        *out = nullptr;
        if (n < 0) throw StatusError(RV_INVALID_INPUT, "negative dimension");
        if (!col_ptr || (n > 0 && (!row_idx || !values))) throw StatusError(RV_INVALID_INPUT, "null CSC arrays");
        MatchOrderResult mo;
        std::string err;
        if (n > 0 && !match_order(n, col_ptr, row_idx, values, mo, err))
            throw StatusError(err.find("finite") != std::string::npos || err.find("empty") != std::string::npos
                                  ? RV_INVALID_INPUT : RV_ANALYSIS_FAILURE, err);
        std::unique_ptr<rv_symbolic_s> s(new rv_symbolic_s());
        if (!analyze(n, col_ptr, row_idx, RV_ORDER_USER, mo.order.data(), nullptr, s->sym, err))
            throw StatusError(RV_ANALYSIS_FAILURE, err);
        if (elim_scaling)
            for (int64_t k = 0; k < n; ++k) elim_scaling[k] = mo.scaling[s->sym.perm_fwd[k]];
        if (has_scaling) *has_scaling = 1;
        *out = s.release();
    });
}

int rv_mc64_matching(int64_t n, const int64_t *col_ptr, const int64_t *row_idx, const double *values,
                     int32_t *matching, double *scaling, int32_t *is_matched, int64_t *matched) {
    return guarded([&] {
        if (n <= 0) throw StatusError(RV_INVALID_INPUT, "MC64 requires non-empty matrix");
        if (!col_ptr || !row_idx || !values) throw StatusError(RV_INVALID_INPUT, "null CSC arrays");
        Mc64Result r;
        std::string err;
        if (!mc64_matching(n, col_ptr, row_idx, values, r, err)) throw StatusError(RV_INVALID_INPUT, err);
        for (int64_t i = 0; i < n; ++i) {
            if (matching) matching[i] = r.matching[i];
            if (scaling) scaling[i] = r.scaling[i];
            if (is_matched) is_matched[i] = r.is_matched[i];
        }
        if (matched) *matched = r.matched;
    });
}

int rv_match_order(int64_t n, const int64_t *col_ptr, const int64_t *row_idx, const double *values,
                   int32_t *order, double *scaling, int64_t *info) {
    return guarded([&] {
        if (n < 0) throw StatusError(RV_INVALID_INPUT, "negative dimension");
        if (!col_ptr || (n > 0 && (!row_idx || !values))) throw StatusError(RV_INVALID_INPUT, "null CSC arrays");
        MatchOrderResult r;
        std::string err;
        if (!match_order(n, col_ptr, row_idx, values, r, err)) throw StatusError(RV_INVALID_INPUT, err);
        for (int64_t i = 0; i < n; ++i) {
            if (order) order[i] = r.order[i];
            if (scaling) scaling[i] = r.scaling[i];
        }
        if (info) { info[0] = r.matched; info[1] = r.condensed_dim; info[2] = r.singletons; info[3] = r.two_cycles; }
    });
}

int rv_symbolic_from_structure(int64_t n, int32_t nsn, const int32_t *sn_begin,
                               const int32_t *sn_end, const int64_t *pat_ptr, const int32_t *pat,
                               const int32_t *parent, const int32_t *perm_fwd, rv_symbolic **out) {
    return guarded([&] {
        if (!out) throw StatusError(RV_INVALID_INPUT, "null output");
        *out = nullptr;
        std::unique_ptr<rv_symbolic_s> s(new rv_symbolic_s());
        Symbolic &S = s->sym;
        S.n = n;
        S.perm_fwd.assign(perm_fwd, perm_fwd + n);
        S.perm_inv.assign(n, -1);
        for (i64 i = 0; i < n; ++i) {
            const int v = perm_fwd[i];
            if (v < 0 || v >= n || S.perm_inv[v] != -1)
                throw StatusError(RV_INVALID_INPUT, "perm_fwd is not a permutation");
            S.perm_inv[v] = (int)i;
        }
        S.sn.resize(nsn);
        i64 covered = 0;
        for (int q = 0; q < nsn; ++q) {
            Supernode &X = S.sn[q];
            X.col_begin = sn_begin[q];
            X.col_end = sn_end[q];
            if (X.col_end <= X.col_begin || X.col_begin < 0 || X.col_end > n)
                throw StatusError(RV_INVALID_INPUT, "invalid supernode column range");
            X.owned = {ColRange{X.col_begin, X.col_end}};
            X.pattern.assign(pat + pat_ptr[q], pat + pat_ptr[q + 1]);
            for (size_t t = 0; t < X.pattern.size(); ++t)
                if (X.pattern[t] < X.col_end || X.pattern[t] >= n ||
                    (t > 0 && X.pattern[t] <= X.pattern[t - 1]))
                    throw StatusError(RV_INVALID_INPUT, "pattern must be sorted and below the supernode");
            X.parent = parent[q];
            if (X.parent >= 0 && X.parent <= q)
                throw StatusError(RV_INVALID_INPUT, "supernodes must be in postorder (parent > child)");
            covered += X.col_end - X.col_begin;
        }
        if (covered != n) throw StatusError(RV_INVALID_INPUT, "supernodes do not partition the columns");
        *out = s.release();
    });
}

int rv_symbolic_info(const rv_symbolic *sym, int64_t *out) {
    return guarded([&] {
        if (!sym || !out) throw StatusError(RV_INVALID_INPUT, "null argument");
        out[0] = sym->sym.n;
        out[1] = (int64_t)sym->sym.sn.size();
        out[2] = sym->sym.predicted_nnz;
    });
}

int rv_symbolic_export(const rv_symbolic *sym, int32_t *sn_begin, int32_t *sn_end, int64_t *pat_ptr,
                       int32_t *pat, int32_t *parent, int32_t *perm_fwd) {
    return guarded([&] {
        if (!sym) throw StatusError(RV_INVALID_INPUT, "null argument");
        const Symbolic &S = sym->sym;
        int64_t p = 0;
        if (pat_ptr) pat_ptr[0] = 0;
        for (size_t q = 0; q < S.sn.size(); ++q) {
            if (sn_begin) sn_begin[q] = S.sn[q].col_begin;
            if (sn_end) sn_end[q] = S.sn[q].col_end;
            if (parent) parent[q] = S.sn[q].parent;
            for (int v : S.sn[q].pattern) {
                if (pat) pat[p] = v;
                ++p;
            }
            if (pat_ptr) pat_ptr[q + 1] = p;
        }
        if (perm_fwd)
            for (i64 i = 0; i < S.n; ++i) perm_fwd[i] = S.perm_fwd[i];
    });
}

void rv_symbolic_free(rv_symbolic *sym) { delete sym; }

int32_t rv_amalgamate(int32_t nsn, const int32_t *sn_begin, const int32_t *sn_end,
                      const int64_t *pat_ptr, const int32_t *pat, const int32_t *parent,
                      int32_t nemin, int32_t *begin_out, int32_t *end_out, int32_t *parent_out,
                      int64_t *rng_ptr, int32_t *rng_lo, int32_t *rng_hi, int64_t *pat_ptr_out,
                      int32_t *pat_out) {
    int32_t result = -1;
    int rc = guarded([&] {
        if (nsn < 0 || (nsn > 0 && (!sn_begin || !sn_end || !pat_ptr || !parent)))
            throw StatusError(RV_INVALID_INPUT, "invalid partition");
        std::vector<Supernode> in(nsn);
        for (int q = 0; q < nsn; ++q) {
            in[q].col_begin = sn_begin[q];
            in[q].col_end = sn_end[q];
            in[q].owned = {ColRange{sn_begin[q], sn_end[q]}};
            in[q].pattern.assign(pat + pat_ptr[q], pat + pat_ptr[q + 1]);
            in[q].parent = parent[q];
        }
        std::vector<Supernode> out = amalgamate(std::move(in), std::max(nemin, 1));
        int64_t rp = 0, pp = 0;
        rng_ptr[0] = 0;
        pat_ptr_out[0] = 0;
        for (size_t q = 0; q < out.size(); ++q) {
            begin_out[q] = out[q].col_begin;
            end_out[q] = out[q].col_end;
            parent_out[q] = out[q].parent;
            for (const ColRange &r : out[q].owned) { rng_lo[rp] = r.lo; rng_hi[rp] = r.hi; ++rp; }
            rng_ptr[q + 1] = rp;
            for (int v : out[q].pattern) pat_out[pp++] = v;
            pat_ptr_out[q + 1] = pp;
        }
        result = (int32_t)out.size();
    });
    return rc == RV_OK ? result : -1;
}

int rv_factor(rv_symbolic *sym, const int64_t *col_ptr, const int64_t *row_idx, const double *values,
              const double *scaling, const rv_factor_options *opts, rv_numeric **out) {
    return do_factor(sym, col_ptr, row_idx, values, scaling, false, opts, out);
}

int rv_factor_device(rv_symbolic *sym, const int64_t *col_ptr, const int64_t *row_idx,
                     const double *d_values, const double *d_scaling, const rv_factor_options *opts,
                     rv_numeric **out) {
    return do_factor(sym, col_ptr, row_idx, d_values, d_scaling, true, opts, out);
}

int rv_refactor(rv_numeric *num, const double *values) {
    return guarded([&] {
        if (!num || !values) throw StatusError(RV_INVALID_INPUT, "null argument");
        if (!num->owns_val) {
            RV_HIP(hipMalloc(&num->d_val, std::max<i64>(num->nnz, 1) * sizeof(double)));
            num->owns_val = true;
        }
        if (num->nnz)
            RV_HIP(hipMemcpyAsync(num->d_val, values, num->nnz * sizeof(double), hipMemcpyHostToDevice,
                                  num->stream));
        num->gpu->factor(num->d_val, num->d_scal);
        RV_HIP(hipStreamSynchronize(num->stream));
    });
}

int rv_refactor_device(rv_numeric *num, const double *d_values) {
    return guarded([&] {
        if (!num || !d_values) throw StatusError(RV_INVALID_INPUT, "null argument");
        // the residuals of later refinements use these values: a handle that owns its copy
        // takes them into it, otherwise it adopts the caller's pointer
        if (num->owns_val) {
            if (num->nnz)
                RV_HIP(hipMemcpyAsync(num->d_val, d_values, num->nnz * sizeof(double), hipMemcpyDeviceToDevice,
                                      num->stream));
        } else {
            num->d_val = const_cast<double *>(d_values);
        }
        num->gpu->factor(num->d_val, num->d_scal);
        RV_HIP(hipStreamSynchronize(num->stream));
    });
}

int rv_solve_device(rv_numeric *num, double *d_rhs, int64_t ldrhs, int32_t nrhs) {
    return guarded([&] {
        if (!num) throw StatusError(RV_SOLVE_BEFORE_FACTOR, "factor() must be called before solve()");
        const i64 n = num->n;
        if (ldrhs < n) throw StatusError(RV_DIMENSION_MISMATCH, "RHS leading dimension smaller than n");
        if (nrhs < 0) throw StatusError(RV_INVALID_INPUT, "negative nrhs");
        if (n == 0 || nrhs == 0) return;
        num->gpu->solve(d_rhs, ldrhs, nrhs, num->d_scal);
        RV_HIP(hipStreamSynchronize(num->stream));
    });
}

int rv_solve_device_noalloc(rv_numeric *num, double *d_rhs, int64_t ldrhs, int32_t nrhs) {
    return guarded([&] {
        if (!num) throw StatusError(RV_SOLVE_BEFORE_FACTOR, "factor() must be called before solve()");
        const i64 n = num->n;
        if (ldrhs < n) throw StatusError(RV_DIMENSION_MISMATCH, "RHS leading dimension smaller than n");
        if (nrhs < 0) throw StatusError(RV_INVALID_INPUT, "negative nrhs");
        if (n == 0 || nrhs == 0) return;
        if (!num->gpu->solve_ready()) throw StatusError(RV_SOLVE_BEFORE_FACTOR, "no completed factorization");
        const long long need = num->gpu->solve_scratch_bytes(nrhs);
        if (need > (long long)noalloc_scratch_bytes())
            throw StatusError(RV_INVALID_INPUT, "solve needs " + std::to_string(need) +
                                                    " bytes of workspace, more than the fixed " +
                                                    std::to_string(noalloc_scratch_bytes()) +
                                                    " bytes of rv_solve_device_noalloc");
        num->gpu->solve_with(d_rhs, ldrhs, nrhs, num->d_scal, noalloc_scratch());
        int err = 0;
        RV_HIP(hipMemcpyAsync(&err, num->gpu->solve_error_word(), sizeof(int), hipMemcpyDeviceToHost, num->stream));
        RV_HIP(hipStreamSynchronize(num->stream));
        if (err) throw StatusError(RV_DEVICE_ERROR, solve_error_text(err));
    });
}

int rv_solve(rv_numeric *num, double *rhs, int64_t ldrhs, int32_t nrhs) {
    return guarded([&] {
        if (!num) throw StatusError(RV_SOLVE_BEFORE_FACTOR, "factor() must be called before solve()");
        const i64 n = num->n;
        if (ldrhs < n) throw StatusError(RV_DIMENSION_MISMATCH, "RHS leading dimension smaller than n");
        if (nrhs < 0) throw StatusError(RV_INVALID_INPUT, "negative nrhs");
        if (n == 0 || nrhs == 0) return;
        double *d = nullptr;
        RV_HIP(hipMalloc(&d, (size_t)n * nrhs * sizeof(double)));
        try {
            RV_HIP(hipMemcpy2D(d, n * sizeof(double), rhs, ldrhs * sizeof(double), n * sizeof(double),
                               nrhs, hipMemcpyHostToDevice));
            num->gpu->solve(d, n, nrhs, num->d_scal);
            RV_HIP(hipStreamSynchronize(num->stream));
            RV_HIP(hipMemcpy2D(rhs, ldrhs * sizeof(double), d, n * sizeof(double), n * sizeof(double),
                               nrhs, hipMemcpyDeviceToHost));
        } catch (...) {
            (void)hipFree(d);
            throw;
        }
        RV_HIP(hipFree(d));
    });
}

int rv_solve_refine(rv_numeric *num, const double *b, double *x, int32_t max_iter, double tol, int32_t *iters,
                    double *backward_error) {
    return guarded([&] {
        if (!num) throw StatusError(RV_SOLVE_BEFORE_FACTOR, "factor() must be called before solve()");
        if (!b || !x) throw StatusError(RV_INVALID_INPUT, "null argument");
        const i64 n = num->n;
        if (n == 0) {
            if (iters) *iters = 0;
            if (backward_error) *backward_error = 0.0;
            return;
        }
        double *d = (double *)num->refine_io.get((size_t)2 * n * sizeof(double));
        RV_HIP(hipMemcpy(d, b, (size_t)n * sizeof(double), hipMemcpyHostToDevice));
        const RefineResult r = do_refine(num, d, d + n, max_iter, tol);
        RV_HIP(hipMemcpy(x, d + n, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
        if (iters) *iters = r.iters;
        if (backward_error) *backward_error = r.backward_error;
    });
}

int rv_solve_refine_device(rv_numeric *num, const double *d_b, double *d_x, int32_t max_iter, double tol,
                           int32_t *iters, double *backward_error) {
    return guarded([&] {
        if (!num) throw StatusError(RV_SOLVE_BEFORE_FACTOR, "factor() must be called before solve()");
        if (!d_b || !d_x) throw StatusError(RV_INVALID_INPUT, "null argument");
        if (num->n == 0) {
            if (iters) *iters = 0;
            if (backward_error) *backward_error = 0.0;
            return;
        }
        const RefineResult r = do_refine(num, d_b, d_x, max_iter, tol);
        if (iters) *iters = r.iters;
        if (backward_error) *backward_error = r.backward_error;
    });
}

int64_t rv_solve_scratch(const rv_numeric *num, int32_t nrhs) {
    if (!num) return 0;
    return num->gpu->solve_scratch_bytes(nrhs);
}

int rv_inertia(const rv_numeric *num, int64_t *out) {
    return guarded([&] {
        if (!num || !out) throw StatusError(RV_INVALID_INPUT, "null argument");
        long long v[3];
        num->gpu->inertia(v);
        out[0] = v[0]; out[1] = v[1]; out[2] = v[2];
    });
}

int rv_stats(const rv_numeric *num, int64_t *out) {
    return guarded([&] {
        if (!num || !out) throw StatusError(RV_INVALID_INPUT, "null argument");
        const FactorStats &s = num->gpu->stats();
        out[0] = s.total_1x1; out[1] = s.total_2x2; out[2] = s.total_delayed; out[3] = s.zero_pivots;
        out[4] = s.max_front_size; out[5] = s.sn_before; out[6] = s.sn_after; out[7] = s.merges;
        out[8] = s.small_leaf_subtrees; out[9] = s.small_leaf_nodes;
    });
}

int32_t rv_num_supernodes(const rv_numeric *num) { return num ? num->gpu->num_supernodes() : -1; }

int rv_supernode_stats(const rv_numeric *num, int32_t s, int64_t *out, double *max_l) {
    return guarded([&] {
        if (!num || !out) throw StatusError(RV_INVALID_INPUT, "null argument");
        if (s < 0 || s >= num->gpu->num_supernodes()) throw StatusError(RV_INVALID_INPUT, "supernode out of range");
        long long v[8];
        num->gpu->supernode_info(s, v, max_l);
        for (int i = 0; i < 8; ++i) out[i] = v[i];
    });
}

int rv_set_front_precision(rv_numeric *num, const int8_t *prec) {
    return guarded([&] {
        if (!num) throw StatusError(RV_INVALID_INPUT, "null handle");
        const int ns = num->gpu->num_supernodes();
        if (prec)
            for (int s = 0; s < ns; ++s)
                if (prec[s] < RV_PREC_FP64 || prec[s] > RV_PREC_FP64_MFMA)
                    throw StatusError(RV_INVALID_INPUT, "front precision must be an RV_PREC_* code");
        num->gpu->set_front_precision((const signed char *)prec);
    });
}

int rv_front_precision(const rv_numeric *num, int8_t *out) {
    return guarded([&] {
        if (!num || !out) throw StatusError(RV_INVALID_INPUT, "null argument");
        const auto &p = num->gpu->front_precision();
        for (size_t s = 0; s < p.size(); ++s) out[s] = p[s];
    });
}

int rv_front_features(const rv_numeric *num, float *out) {
    return guarded([&] {
        if (!num || !out) throw StatusError(RV_INVALID_INPUT, "null argument");
        num->gpu->front_features(out);
    });
}

namespace {
void mlp_check(int64_t n, const float *params, const float *x) {
    if (n <= 0 || n > (int64_t)1 << 30) throw StatusError(RV_INVALID_INPUT, "sample count out of range");
    if (!params || !x) throw StatusError(RV_INVALID_INPUT, "null argument");
    require_device();
}
void mlp_labels(int64_t n, const int32_t *labels) {
    if (!labels) throw StatusError(RV_INVALID_INPUT, "null labels");
    for (int64_t i = 0; i < n; ++i)
        if (labels[i] < 0 || labels[i] >= RV_MLP_CLASSES) throw StatusError(RV_INVALID_INPUT, "label out of range");
}
}  // namespace

int rv_mlp_forward(const float *params, int64_t n, const float *x, float *probs) {
    return guarded([&] {
        mlp_check(n, params, x);
        if (!probs) throw StatusError(RV_INVALID_INPUT, "null output");
        mlp_forward(params, (int)n, x, probs);
    });
}

int rv_mlp_loss_grad(const float *params, int64_t n, const float *x, const int32_t *labels,
                     const float *class_weight, float *loss, float *grad) {
    return guarded([&] {
        mlp_check(n, params, x);
        mlp_labels(n, labels);
        if (!loss || !grad) throw StatusError(RV_INVALID_INPUT, "null output");
        mlp_loss_grad(params, (int)n, x, labels, class_weight, loss, grad);
    });
}

int rv_mlp_train(float *params, int64_t n, const float *x, const int32_t *labels, const float *class_weight,
                 int32_t epochs, float lr, float *final_loss) {
    return guarded([&] {
        mlp_check(n, params, x);
        mlp_labels(n, labels);
        if (epochs < 0 || !(lr > 0.0f)) throw StatusError(RV_INVALID_INPUT, "epochs >= 0 and lr > 0 required");
        mlp_train(params, (int)n, x, labels, class_weight, epochs, lr, final_loss);
    });
}

int rv_supernode_structure(const rv_numeric *num, int32_t s, int32_t *owned, int32_t *pattern,
                           int64_t *out) {
    return guarded([&] {
        if (!num || !out) throw StatusError(RV_INVALID_INPUT, "null argument");
        const Prepared &P = *num->prep;
        if (s < 0 || s >= (int)P.sns.size()) throw StatusError(RV_INVALID_INPUT, "supernode out of range");
        const Supernode &S = P.sns[s];
        int q = 0;
        for (const ColRange &r : S.owned)
            for (int j = r.lo; j < r.hi; ++j) {
                if (owned) owned[q] = j;
                ++q;
            }
        out[0] = q;
        out[1] = (int64_t)S.pattern.size();
        if (pattern)
            for (size_t i = 0; i < S.pattern.size(); ++i) pattern[i] = S.pattern[i];
    });
}

int rv_export_front(const rv_numeric *num, int32_t s, double *l11, double *l21, int8_t *dkind,
                    double *ddiag, double *doff, int32_t *col_indices, int32_t *row_indices,
                    int32_t *local_perm) {
    return guarded([&] {
        if (!num) throw StatusError(RV_INVALID_INPUT, "null argument");
        if (s < 0 || s >= num->gpu->num_supernodes()) throw StatusError(RV_INVALID_INPUT, "supernode out of range");
        num->gpu->export_front(s, l11, l21, (signed char *)dkind, ddiag, doff, col_indices, row_indices,
                               local_perm);
    });
}

int rv_export_contribution(const rv_numeric *num, int32_t s, double *data, int32_t *row_indices,
                           int64_t *out) {
    return guarded([&] {
        if (!num || !out) throw StatusError(RV_INVALID_INPUT, "null argument");
        if (s < 0 || s >= num->gpu->num_supernodes()) throw StatusError(RV_INVALID_INPUT, "supernode out of range");
        if (!num->gpu->front_held(s))
            throw StatusError(RV_INVALID_INPUT, "the contribution block of this supernode was overwritten by a later "
                                                "level (RV_KEEP_FRONTS=1 when the handle is made keeps every one)");
        int nd = 0;
        const int size = num->gpu->export_contribution(s, data, row_indices, &nd);
        out[0] = size;
        out[1] = size > 0 ? nd : 0;
    });
}

int rv_dense_aptp_factor(double *a, int64_t lda, int32_t m, int32_t p, const rv_factor_options *opts,
                         int32_t *perm, int8_t *dkind, double *ddiag, double *doff, int64_t *stats,
                         double *max_l) {
    return guarded([&] {
        if (!a || m < 0 || p < 0 || lda < m) throw StatusError(RV_INVALID_INPUT, "invalid dense arguments");
        if (p > m)
            throw StatusError(RV_INVALID_INPUT, "num_fully_summed (" + std::to_string(p) +
                                                    ") > matrix dimension (" + std::to_string(m) + ")");
        FactorOptions o = to_opts(opts);
        require_device();
        o.small_leaf_threshold = 0;
        if (m == 0) {
            if (stats) stats[0] = stats[1] = stats[2] = stats[3] = 0;
            if (max_l) *max_l = 0.0;
            return;
        }
        // one front: owned columns [0, p), pattern [p, m), natural order
        std::vector<i64> cp(m + 1, 0), ri;
        std::vector<double> val;
        ri.reserve((size_t)m * m);
        val.reserve((size_t)m * m);
        for (int j = 0; j < m; ++j) {
            for (int i = 0; i < m; ++i) {
                ri.push_back(i);
                val.push_back(i >= j ? a[i + j * lda] : a[j + i * lda]);
            }
            cp[j + 1] = (i64)ri.size();
        }
        rv_symbolic_s sym;
        sym.sym.n = m;
        sym.sym.perm_fwd.resize(m);
        sym.sym.perm_inv.resize(m);
        for (int i = 0; i < m; ++i) sym.sym.perm_fwd[i] = sym.sym.perm_inv[i] = i;
        if (p > 0) {
            Supernode S;
            S.col_begin = 0;
            S.col_end = p;
            S.owned = {ColRange{0, p}};
            for (int i = p; i < m; ++i) S.pattern.push_back(i);
            S.parent = -1;
            sym.sym.sn.push_back(S);
            // the rows beyond p form a second (root) supernode that is never factored here
            if (p < m) {
                Supernode R;
                R.col_begin = p;
                R.col_end = m;
                R.owned = {ColRange{p, m}};
                R.parent = -1;
                sym.sym.sn.push_back(R);
            }
        } else {
            if (stats) stats[0] = stats[1] = stats[2] = stats[3] = 0;
            if (perm) for (int i = 0; i < m; ++i) perm[i] = i;
            if (max_l) *max_l = 0.0;
            return;
        }
        o.nemin = 1;  // keep the two supernodes apart
        ensure_prepared(&sym, cp.data(), ri.data(), o);
        hipStream_t st;
        RV_HIP(hipStreamCreate(&st));
        double *dv = nullptr;
        try {
            RV_HIP(hipMalloc(&dv, val.size() * sizeof(double)));
            RV_HIP(hipMemcpy(dv, val.data(), val.size() * sizeof(double), hipMemcpyHostToDevice));
            NumericGPU g(sym.prep.get(), o, st);
            g.factor(dv, nullptr);
            RV_HIP(hipStreamSynchronize(st));
            // supernode 0 is the dense front
            const int mm = g.front_m(0);
            std::vector<double> F((size_t)mm * mm);
            std::vector<int> pm(mm), rows(mm);
            g.export_raw(0, F.data(), pm.data(), rows.data());
            for (int j = 0; j < p; ++j)
                for (int i = j; i < mm; ++i) a[i + j * lda] = F[i + (size_t)j * mm];
            if (perm) for (int i = 0; i < mm; ++i) perm[i] = pm[i];
            const int ne = g.front_ne(0);
            std::vector<signed char> kd(std::max(ne, 1));
            std::vector<double> dg(std::max(ne, 1)), of(std::max(ne, 1));
            g.front_d(0, kd.data(), dg.data(), of.data());
            for (int c = 0; c < p; ++c) {
                if (dkind) dkind[c] = c < ne ? kd[c] : 0;
                if (ddiag) ddiag[c] = c < ne ? dg[c] : 0.0;
                if (doff) doff[c] = c < ne ? of[c] : 0.0;
            }
            long long info[8];
            double ml = 0.0;
            g.supernode_info(0, info, &ml);
            if (stats) { stats[0] = ne; stats[1] = info[4]; stats[2] = info[5]; stats[3] = p - ne; }
            if (max_l) *max_l = ml;
        } catch (...) {
            if (dv) (void)hipFree(dv);
            (void)hipStreamDestroy(st);
            throw;
        }
        (void)hipFree(dv);
        (void)hipStreamDestroy(st);
    });
}

int rv_set_graph_mode(rv_numeric *num, int32_t on) {
    return guarded([&] {
        if (!num) throw StatusError(RV_INVALID_INPUT, "null argument");
        num->gpu->graph_mode = on != 0;
    });
}

int64_t rv_graph_launches(const rv_numeric *num) { return num ? num->gpu->graph_launches() : -1; }

int64_t rv_chain_launches(const rv_numeric *num) { return num ? num->gpu->chain_launches() : -1; }

int64_t rv_lookahead_launches(const rv_numeric *num) { return num ? num->gpu->lookahead_launches() : -1; }

int64_t rv_device_bytes(const rv_numeric *num) { return num ? (int64_t)num->gpu->device_bytes() : -1; }

int rv_set_timing(rv_numeric *num, int32_t on) {
    return guarded([&] {
        if (!num) throw StatusError(RV_INVALID_INPUT, "null argument");
        num->gpu->timing = on != 0;
    });
}

int rv_timings(const rv_numeric *num, double *out) {
    return guarded([&] {
        if (!num || !out) throw StatusError(RV_INVALID_INPUT, "null argument");
        const PhaseTimes &t = num->gpu->times();
        out[0] = t.schur_ms;
        out[1] = t.schur_flops;
        out[2] = t.schur_launches;
        out[3] = t.big_ms;
        out[4] = t.big_flops;
        out[5] = t.big_launches;
    });
}

void rv_numeric_free(rv_numeric *num) { delete num; }

}  // extern "C"
