// rivrs_oracle.cpp -- TEST INFRASTRUCTURE ONLY: the CPU FP64 parity oracle.
//
// A decision-for-decision C++ restatement of rivrs-sparse's numeric phase:
//   sparse/src/symmetric/factor.rs    (dense APTP kernel, TPP, contribution GEMM)
//   sparse/src/symmetric/numeric.rs   (multifrontal driver, assembly, extend-add, extraction)
//   sparse/src/symmetric/amalgamation.rs (supernode amalgamation)
//   sparse/src/symmetric/diagonal.rs  (MixedDiagonal solve / inertia)
//   sparse/src/symmetric/solve.rs     (supernodal forward / diagonal / backward solve)
//   sparse/src/symmetric/solver.rs    (permute/scale order around the core solve)
//   sparse/src/validate.rs            (scaled backward error)
// The dense BLAS the reference delegates to faer (TRSM / GEMM) is written out as plain
// loops with a fixed summation order (ascending contraction index), which the GPU kernels
// for the same steps reproduce.  Nothing here is used by the product library.
//
// Build: oracle/Makefile (gcc, -O3 -ffp-contract=off so scalar sequences are unfused).
#include "rivrs_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using std::vector;
typedef int64_t i64;

struct Opts {
    double threshold = 0.01, small = 1e-20;
    int nb = 256, ib = 32;
    bool tpp = true;
    int nemin = 32, small_leaf = 256;
};

inline double &A(double *a, i64 lda, i64 i, i64 j) { return a[i + j * lda]; }

// ---------------------------------------------------------------------------
// MixedDiagonal (diagonal.rs:49-462)
// kind: 0 = Delayed, 1 = OneByOne, 2 = TwoByTwo{partner}
// ---------------------------------------------------------------------------
struct Diag {
    int n = 0;
    vector<int8_t> kind;
    vector<int> partner;
    vector<double> diag, off;
    void init(int nn) {
        n = nn;
        kind.assign(nn, 0);
        partner.assign(nn, -1);
        diag.assign(nn, 0.0);
        off.assign(nn, 0.0);
    }
    void set1(int c, double v) { kind[c] = 1; diag[c] = v; }
    void set2(int c, double a, double b, double cc) {  // diagonal.rs set_2x2
        kind[c] = 2; partner[c] = c + 1;
        kind[c + 1] = 2; partner[c + 1] = c;
        diag[c] = a; diag[c + 1] = cc; off[c] = b;
    }
    bool two_first(int c) const { return kind[c] == 2 && partner[c] > c; }
    void grow(int nn) {  // diagonal.rs grow
        if (nn > n) {
            kind.resize(nn, 0); partner.resize(nn, -1);
            diag.resize(nn, 0.0); off.resize(nn, 0.0); n = nn;
        }
    }
    void truncate(int nn) {
        kind.resize(nn); partner.resize(nn); diag.resize(nn); off.resize(nn); n = nn;
    }
    // diagonal.rs copy_from_offset
    void copy_from_offset(const Diag &s, int off0, int count) {
        int col = 0;
        while (col < count) {
            if (s.kind[col] == 1) {
                kind[off0 + col] = 1; diag[off0 + col] = s.diag[col]; col += 1;
            } else if (s.kind[col] == 2 && col + 1 < count) {
                int d = off0 + col;
                kind[d] = 2; partner[d] = d + 1; kind[d + 1] = 2; partner[d + 1] = d;
                diag[d] = s.diag[col]; diag[d + 1] = s.diag[col + 1]; off[d] = s.off[col];
                col += 2;
            } else {
                col += 1;
            }
        }
    }
};

struct Stats {
    i64 n1 = 0, n2 = 0, ndel = 0;
    double max_l = 0.0;
};

struct FactorResult {  // AptpFactorResult (factor.rs:211-224)
    Diag d;
    vector<int> perm;
    int ne = 0;
    Stats st;
};

// ---------------------------------------------------------------------------
// Symmetric swaps (factor.rs:452-505)
// ---------------------------------------------------------------------------
void swap_symmetric_block(double *a, i64 lda, int i, int j, int col_start, int row_limit) {
    if (i == j) return;
    if (i > j) std::swap(i, j);
    std::swap(A(a, lda, i, i), A(a, lda, j, j));
    for (int k = col_start; k < i; ++k) std::swap(A(a, lda, i, k), A(a, lda, j, k));
    for (int k = i + 1; k < j; ++k) std::swap(A(a, lda, k, i), A(a, lda, j, k));
    for (int k = j + 1; k < row_limit; ++k) std::swap(A(a, lda, k, i), A(a, lda, k, j));
}
void swap_symmetric(double *a, i64 lda, int m, int i, int j) {
    swap_symmetric_block(a, lda, i, j, 0, m);
}

// ---------------------------------------------------------------------------
// factor_block_diagonal (factor.rs:891-1153): complete pivoting on an ib x ib block
// ---------------------------------------------------------------------------
struct BlockRes {
    Diag d;
    vector<int> perm;
    int nelim = 0;
    Stats st;
};

BlockRes factor_block_diagonal(double *a, i64 lda, int col_start, int bs, double small,
                               int row_limit) {
    const int be = col_start + bs;
    BlockRes r;
    r.perm.resize(bs);
    for (int i = 0; i < bs; ++i) r.perm[i] = i;
    r.d.init(bs);
    int k = 0;
    while (k < bs) {
        const int cur = col_start + k;
        const int remaining = bs - k;
        double max_val = 0.0;
        int max_row = cur, max_col = cur;
        for (int j = cur; j < be; ++j)
            for (int i = j; i < be; ++i) {
                double v = std::fabs(A(a, lda, i, j));
                if (v > max_val) { max_val = v; max_row = i; max_col = j; }
            }
        if (max_val < small) {
            r.st.ndel += remaining;
            break;
        }
        if (max_row == max_col) {
            if (max_row != cur) {
                swap_symmetric_block(a, lda, cur, max_row, col_start, row_limit);
                std::swap(r.perm[k], r.perm[max_row - col_start]);
            }
            const double dkk = A(a, lda, cur, cur);
            const double inv = 1.0 / dkk;
            double max_l = 0.0;
            for (int i = cur + 1; i < be; ++i) {
                double l = A(a, lda, i, cur) * inv;
                A(a, lda, i, cur) = l;
                if (std::fabs(l) > max_l) max_l = std::fabs(l);
            }
            for (int j = cur + 1; j < be; ++j) {
                const double ldl = A(a, lda, j, cur) * dkk;
                for (int i = j; i < be; ++i) A(a, lda, i, j) -= A(a, lda, i, cur) * ldl;
            }
            r.d.set1(k, dkk);
            r.st.n1 += 1;
            if (max_l > r.st.max_l) r.st.max_l = max_l;
            k += 1;
        } else {
            const int t = max_row, mi = max_col;
            const double amm = A(a, lda, mi, mi), att = A(a, lda, t, t), atm = A(a, lda, t, mi);
            const double delta = amm * att - atm * atm;
            if (remaining < 2) {
                const int pos = (std::fabs(amm) >= std::fabs(att)) ? mi : t;
                if (pos != cur) {
                    swap_symmetric_block(a, lda, cur, pos, col_start, row_limit);
                    std::swap(r.perm[k], r.perm[pos - col_start]);
                }
                const double dkk = A(a, lda, cur, cur);
                if (std::fabs(dkk) < small) { r.st.ndel += 1; break; }
                const double inv = 1.0 / dkk;
                for (int i = cur + 1; i < be; ++i) A(a, lda, i, cur) *= inv;
                r.d.set1(k, dkk);
                r.st.n1 += 1;
                k += 1;
                continue;
            }
            if (std::fabs(delta) >= 0.5 * atm * atm) {
                if (mi != cur) {
                    swap_symmetric_block(a, lda, cur, mi, col_start, row_limit);
                    std::swap(r.perm[k], r.perm[mi - col_start]);
                }
                const int new_t = (t == cur) ? mi : t;
                if (new_t != cur + 1) {
                    swap_symmetric_block(a, lda, cur + 1, new_t, col_start, row_limit);
                    std::swap(r.perm[k + 1], r.perm[new_t - col_start]);
                }
                const double a11 = A(a, lda, cur, cur), a22 = A(a, lda, cur + 1, cur + 1),
                             a21 = A(a, lda, cur + 1, cur);
                const double det = a11 * a22 - a21 * a21;
                const double inv_det = 1.0 / det;
                double max_l = 0.0;
                for (int i = cur + 2; i < be; ++i) {
                    const double ai1 = A(a, lda, i, cur), ai2 = A(a, lda, i, cur + 1);
                    const double l1 = (ai1 * a22 - ai2 * a21) * inv_det;
                    const double l2 = (ai2 * a11 - ai1 * a21) * inv_det;
                    A(a, lda, i, cur) = l1;
                    A(a, lda, i, cur + 1) = l2;
                    if (std::fabs(l1) > max_l) max_l = std::fabs(l1);
                    if (std::fabs(l2) > max_l) max_l = std::fabs(l2);
                }
                for (int j = cur + 2; j < be; ++j) {
                    const double l1j = A(a, lda, j, cur), l2j = A(a, lda, j, cur + 1);
                    const double w1 = l1j * a11 + l2j * a21;
                    const double w2 = l1j * a21 + l2j * a22;
                    for (int i = j; i < be; ++i)
                        A(a, lda, i, j) -= A(a, lda, i, cur) * w1 + A(a, lda, i, cur + 1) * w2;
                }
                r.d.set2(k, a11, a21, a22);
                r.st.n2 += 1;
                if (max_l > r.st.max_l) r.st.max_l = max_l;
                k += 2;
            } else {
                const int pos = (std::fabs(amm) >= std::fabs(att)) ? mi : t;
                if (pos != cur) {
                    swap_symmetric_block(a, lda, cur, pos, col_start, row_limit);
                    std::swap(r.perm[k], r.perm[pos - col_start]);
                }
                const double dkk = A(a, lda, cur, cur);
                const double inv = 1.0 / dkk;
                double max_l = 0.0;
                for (int i = cur + 1; i < be; ++i) {
                    double l = A(a, lda, i, cur) * inv;
                    A(a, lda, i, cur) = l;
                    if (std::fabs(l) > max_l) max_l = std::fabs(l);
                }
                for (int j = cur + 1; j < be; ++j) {
                    const double ldl = A(a, lda, j, cur) * dkk;
                    for (int i = j; i < be; ++i) A(a, lda, i, j) -= A(a, lda, i, cur) * ldl;
                }
                r.d.set1(k, dkk);
                r.st.n1 += 1;
                if (max_l > r.st.max_l) r.st.max_l = max_l;
                k += 1;
            }
        }
    }
    r.nelim = k;
    return r;
}

// adjust_for_2x2_boundary (factor.rs:1160-1172)
int adjust_for_2x2_boundary(int e, const Diag &d) {
    if (e == 0) return 0;
    const int last = e - 1;
    if (d.kind[last] == 2 && d.partner[last] > last) return e - 1;
    return e;
}

// BlockBackup (factor.rs:1183-1276)
struct Backup {
    vector<double> data;  // rows x cols column-major, rows = m - col_start
    int rows = 0;
    void create(const double *a, i64 lda, int col_start, int cols, int m) {
        rows = m - col_start;
        data.resize((size_t)rows * cols);
        for (int j = 0; j < cols; ++j)
            for (int i = 0; i < rows; ++i)
                data[i + (size_t)j * rows] = a[(col_start + i) + (i64)(col_start + j) * lda];
    }
    double at(int i, int j) const { return data[i + (size_t)j * rows]; }
    void restore_failed(double *a, i64 lda, int col_start, int nelim, int cols) const {
        for (int j = nelim; j < cols; ++j)
            for (int i = 0; i < rows; ++i) A(a, lda, col_start + i, col_start + j) = at(i, j);
    }
    void restore_diagonal_with_perm(double *a, i64 lda, int col_start, int nelim, int cols,
                                    const vector<int> &bp) const {
        for (int j = nelim; j < cols; ++j) {
            const int c = bp[j];
            for (int i = nelim; i < cols; ++i) {
                const int r = bp[i];
                A(a, lda, col_start + i, col_start + j) = (r >= c) ? at(r, c) : at(c, r);
            }
        }
        for (int j = nelim; j < cols; ++j) {
            const int c = bp[j];
            for (int i = cols; i < rows; ++i) A(a, lda, col_start + i, col_start + j) = at(i, c);
        }
    }
};

// ---------------------------------------------------------------------------
// apply_and_check (factor.rs:1295-1408).  The TRSM (faer solve_unit_lower_triangular,
// factor.rs:1338) is written as the row-wise forward substitution
//   x_j = a_j - sum_{i<j} L11[j,i] x_i   (i ascending)
// ---------------------------------------------------------------------------
int apply_and_check(double *a, i64 lda, int col_start, int bn, int bcols, int m, const Diag &d,
                    double threshold) {
    if (bn == 0) return 0;
    const int panel_rows = m - col_start - bcols;
    if (panel_rows == 0) return bn;
    const int ps = col_start + bcols;
    // L11 copy (aliasing avoidance; values identical)
    vector<double> l11((size_t)bn * bn);
    for (int j = 0; j < bn; ++j)
        for (int i = 0; i < bn; ++i) l11[i + (size_t)j * bn] = A(a, lda, col_start + i, col_start + j);
    for (int r = ps; r < m; ++r) {
        for (int j = 0; j < bn; ++j) {
            double s = A(a, lda, r, col_start + j);
            for (int i = 0; i < j; ++i) s -= l11[j + (size_t)i * bn] * A(a, lda, r, col_start + i);
            A(a, lda, r, col_start + j) = s;
        }
    }
    int col = 0;
    while (col < bn) {
        if (d.kind[col] == 1) {
            const double inv = 1.0 / d.diag[col];
            for (int i = 0; i < panel_rows; ++i) A(a, lda, ps + i, col_start + col) *= inv;
            col += 1;
        } else if (d.two_first(col)) {
            const double ba = d.diag[col], bb = d.off[col], bc = d.diag[col + 1];
            const double det = ba * bc - bb * bb;
            const double inv_det = 1.0 / det;
            for (int i = 0; i < panel_rows; ++i) {
                const double x1 = A(a, lda, ps + i, col_start + col);
                const double x2 = A(a, lda, ps + i, col_start + col + 1);
                A(a, lda, ps + i, col_start + col) = (x1 * bc - x2 * bb) * inv_det;
                A(a, lda, ps + i, col_start + col + 1) = (x2 * ba - x1 * bb) * inv_det;
            }
            col += 2;
        } else {
            col += 1;
        }
    }
    const double bound = 1.0 / threshold;
    int eff = bn;
    int sc = 0;
    while (sc < bn) {
        const int w = d.two_first(sc) ? 2 : 1;
        bool ok = true;
        for (int c = sc; c < sc + w && ok; ++c) {
            if (c >= bn) break;
            for (int i = 0; i < panel_rows; ++i)
                if (std::fabs(A(a, lda, ps + i, col_start + c)) >= bound) { ok = false; break; }
        }
        if (!ok) { eff = sc; break; }
        sc += w;
    }
    return eff;
}

// compute_ld_into (factor.rs:1536-1563): dst[i,col] = L[i,:] * D
void compute_ld(const double *l, i64 ldl, int nrows, const Diag &d, int nelim, double *dst,
                i64 ldd) {
    int col = 0;
    while (col < nelim) {
        if (d.kind[col] == 1) {
            const double dv = d.diag[col];
            for (int i = 0; i < nrows; ++i) dst[i + col * ldd] = l[i + col * ldl] * dv;
            col += 1;
        } else if (d.two_first(col)) {
            const double ba = d.diag[col], bb = d.off[col], bc = d.diag[col + 1];
            for (int i = 0; i < nrows; ++i) {
                const double l1 = l[i + col * ldl], l2 = l[i + (col + 1) * ldl];
                dst[i + col * ldd] = l1 * ba + l2 * bb;
                dst[i + (col + 1) * ldd] = l1 * bb + l2 * bc;
            }
            col += 2;
        } else {
            col += 1;
        }
    }
}

// C[i,j] -= sum_c W[i,c] * L[j,c]  (c ascending, sum formed first) over the rectangle
// rows [0,nr) x cols [0,nc), optionally restricted to i >= j (lower part).
// This is the fixed-order form of faer's matmul / triangular matmul calls used by
// update_trailing (factor.rs:1491,1513) and compute_contribution_gemm (factor.rs:1639).
// Every entry is summed on its own, over c ascending from zero, and subtracted once; the loops
// are blocked (8 rows x 4 columns held in registers over the whole sum) without changing that
// order.  Column groups are independent, so large products are split over threads by column
// group: results do not depend on the thread count.
namespace {
constexpr int GR = 8, GC = 4;

// one 8 x 4 register tile at rows [i, i + GR), columns [j, j + GC), every row and column inside
inline void wlt_tile_full(double *C, i64 ldc, int i, int j, const double *W, i64 ldw, const double *L,
                          i64 ldl, int kk, bool lower) {
    double acc[GC][GR];
    for (int q = 0; q < GC; ++q)
        for (int r = 0; r < GR; ++r) acc[q][r] = 0.0;
    const double *w = W + i;
    const double *l = L + j;
    for (int c = 0; c < kk; ++c, w += ldw, l += ldl) {
        double wv[GR];
        for (int r = 0; r < GR; ++r) wv[r] = w[r];
        for (int q = 0; q < GC; ++q) {
            const double lq = l[q];
            for (int r = 0; r < GR; ++r) acc[q][r] += wv[r] * lq;
        }
    }
    for (int q = 0; q < GC; ++q) {
        double *cc = C + (j + q) * ldc + i;
        for (int r = 0; r < GR; ++r)
            if (!lower || i + r >= j + q) cc[r] -= acc[q][r];
    }
}

// the same for a partial tile (mr rows, nq columns)
inline void wlt_tile_edge(double *C, i64 ldc, int i, int j, int mr, int nq, const double *W, i64 ldw,
                          const double *L, i64 ldl, int kk, bool lower) {
    for (int q = 0; q < nq; ++q) {
        double acc[GR];
        for (int r = 0; r < mr; ++r) acc[r] = 0.0;
        for (int c = 0; c < kk; ++c) {
            const double lq = L[j + q + c * ldl];
            const double *w = W + i + c * ldw;
            for (int r = 0; r < mr; ++r) acc[r] += w[r] * lq;
        }
        double *cc = C + (j + q) * ldc + i;
        for (int r = 0; r < mr; ++r)
            if (!lower || i + r >= j + q) cc[r] -= acc[r];
    }
}
}  // namespace

void gemm_sub_wlt(double *C, i64 ldc, int nr, int nc, const double *W, i64 ldw,
                  const double *L, i64 ldl, int kk, bool lower) {
    const int ngrp = (nc + GC - 1) / GC;
#ifdef _OPENMP
    const bool par = !omp_in_parallel() && omp_get_max_threads() > 1 && (double)nr * nc * kk > 4e6;
#pragma omp parallel for schedule(dynamic, 4) if (par)
#endif
    for (int g = 0; g < ngrp; ++g) {
        const int j = g * GC, nq = std::min(GC, nc - j);
        const int i0 = lower ? j : 0;
        for (int i = i0 - i0 % GR; i < nr; i += GR) {
            const int mr = std::min(GR, nr - i);
            if (mr == GR && nq == GC)
                wlt_tile_full(C, ldc, i, j, W, ldw, L, ldl, kk, lower);
            else
                wlt_tile_edge(C, ldc, i, j, mr, nq, W, ldw, L, ldl, kk, lower);
        }
    }
}

// update_trailing (factor.rs:1418-1526)
void update_trailing(double *a, i64 lda, int col_start, int nelim, int bcols, int m, int p,
                     const Diag &d) {
    if (nelim == 0) return;
    const int ts = col_start + bcols;
    const int tsz = m - ts;
    if (tsz == 0) return;
    vector<double> w((size_t)tsz * nelim), l21((size_t)tsz * nelim);
    compute_ld(&A(a, lda, ts, col_start), lda, tsz, d, nelim, w.data(), tsz);
    for (int c = 0; c < nelim; ++c)
        for (int i = 0; i < tsz; ++i) l21[i + (size_t)c * tsz] = A(a, lda, ts + i, col_start + c);
    const int fs = std::max(p - ts, 0);
    if (fs > 0)
        gemm_sub_wlt(&A(a, lda, ts, ts), lda, fs, fs, w.data(), tsz, l21.data(), tsz, nelim, true);
    const int nfs = m - p;
    if (nfs > 0 && fs > 0)
        gemm_sub_wlt(&A(a, lda, p, ts), lda, nfs, fs, w.data() + fs, tsz, l21.data(), tsz, nelim,
                     false);
}

// update_cross_terms (factor.rs:1669-1766)
void update_cross_terms(double *a, i64 lda, int col_start, int nelim, int bcols, int m,
                        const Diag &d) {
    if (nelim == 0 || nelim >= bcols) return;
    const int nf = bcols - nelim;
    const int ts = col_start + bcols;
    const int tsz = m - ts;
    vector<double> lb((size_t)nf * nelim), wb((size_t)nf * nelim);
    for (int c = 0; c < nelim; ++c)
        for (int i = 0; i < nf; ++i) lb[i + (size_t)c * nf] = A(a, lda, col_start + nelim + i, col_start + c);
    compute_ld(lb.data(), nf, nf, d, nelim, wb.data(), nf);
    for (int j = 0; j < nf; ++j)
        for (int i = j; i < nf; ++i) {
            double s = 0.0;
            for (int c = 0; c < nelim; ++c) s += wb[i + (size_t)c * nf] * lb[j + (size_t)c * nf];
            A(a, lda, col_start + nelim + i, col_start + nelim + j) -= s;
        }
    if (tsz > 0) {
        vector<double> lp((size_t)tsz * nelim), wp((size_t)tsz * nelim);
        for (int c = 0; c < nelim; ++c)
            for (int i = 0; i < tsz; ++i) lp[i + (size_t)c * tsz] = A(a, lda, ts + i, col_start + c);
        compute_ld(lp.data(), tsz, tsz, d, nelim, wp.data(), tsz);
        for (int j = 0; j < nf; ++j)
            for (int i = 0; i < tsz; ++i) {
                double s = 0.0;
                for (int c = 0; c < nelim; ++c) s += wp[i + (size_t)c * tsz] * lb[j + (size_t)c * nf];
                A(a, lda, ts + i, col_start + nelim + j) -= s;
            }
    }
}

// ---------------------------------------------------------------------------
// factor_inner (factor.rs:1811-2095)
// ---------------------------------------------------------------------------
FactorResult factor_inner(double *a, i64 lda, int m, int p, int nfs_boundary, const Opts &o) {
    const int ib = o.ib;
    FactorResult res;
    res.perm.resize(m);
    for (int i = 0; i < m; ++i) res.perm[i] = i;
    vector<int> &col_order = res.perm;
    res.d.init(p);
    int k = 0, end_pos = p;
    vector<double> buf(ib);
    vector<int> cobuf(ib);
    while (k < end_pos) {
        const int bs = std::min(end_pos - k, ib);
        const int be = k + bs;
        Backup bk;
        bk.create(a, lda, k, bs, m);
        BlockRes br = factor_block_diagonal(a, lda, k, bs, o.small, be);
        if (br.nelim == 0) {
            bk.restore_failed(a, lda, k, 0, bs);
            for (int off = 0; off < bs; ++off) {
                end_pos -= 1;
                if (k + off < end_pos) {
                    swap_symmetric(a, lda, m, k + off, end_pos);
                    std::swap(col_order[k + off], col_order[end_pos]);
                }
                res.st.ndel += 1;
            }
            continue;
        }
        // 3. permute panel columns
        for (int r = be; r < m; ++r) {
            for (int i = 0; i < bs; ++i) buf[i] = A(a, lda, r, k + br.perm[i]);
            for (int i = 0; i < bs; ++i) A(a, lda, r, k + i) = buf[i];
        }
        // 4. zero D off-diagonals
        {
            int bc = 0;
            while (bc < br.nelim) {
                if (br.d.two_first(bc)) { A(a, lda, k + bc + 1, k + bc) = 0.0; bc += 2; }
                else bc += 1;
            }
        }
        // 5. row perm propagation to columns 0..k
        for (int c = 0; c < k; ++c) {
            for (int i = 0; i < bs; ++i) buf[i] = A(a, lda, k + br.perm[i], c);
            for (int i = 0; i < bs; ++i) A(a, lda, k + i, c) = buf[i];
        }
        // 6. col_order
        for (int i = 0; i < bs; ++i) cobuf[i] = col_order[k + i];
        for (int i = 0; i < bs; ++i) col_order[k + i] = cobuf[br.perm[i]];
        // 7-8. apply + adjust
        int eff = apply_and_check(a, lda, k, br.nelim, bs, m, br.d, o.threshold);
        eff = adjust_for_2x2_boundary(eff, br.d);
        // 9. partial restore
        if (eff < br.nelim) bk.restore_diagonal_with_perm(a, lda, k, eff, bs, br.perm);
        const int nelim = eff;
        // 10. Schur updates
        if (nelim > 0) {
            update_trailing(a, lda, k, nelim, bs, m, nfs_boundary, br.d);
            update_cross_terms(a, lda, k, nelim, bs, m, br.d);
        }
        // 11. accumulate D and stats
        res.d.copy_from_offset(br.d, k, nelim);
        {
            int sc = 0;
            while (sc < nelim) {
                if (br.d.kind[sc] == 1) { res.st.n1 += 1; sc += 1; }
                else if (br.d.two_first(sc)) { res.st.n2 += 1; sc += 2; }
                else sc += 1;
            }
        }
        if (br.st.max_l > res.st.max_l) res.st.max_l = br.st.max_l;
        for (int c = 0; c < nelim; ++c)
            for (int i = be; i < m; ++i) {
                double v = std::fabs(A(a, lda, i, k + c));
                if (v > res.st.max_l) res.st.max_l = v;
            }
        // 12. delay failed columns
        if (nelim < bs) {
            const int nd = bs - nelim;
            for (int i = nd - 1; i >= 0; --i) {
                const int dp = k + nelim + i;
                end_pos -= 1;
                if (dp < end_pos) {
                    swap_symmetric(a, lda, m, dp, end_pos);
                    std::swap(col_order[dp], col_order[end_pos]);
                }
                res.st.ndel += 1;
            }
        }
        k += nelim;
    }
    res.ne = k;
    res.d.truncate(k);
    return res;
}

// two_level_factor (factor.rs:2108-2252)
FactorResult two_level_factor(double *a, i64 lda, int m, int p, const Opts &o) {
    const int nb = o.nb;
    FactorResult res;
    res.perm.resize(m);
    for (int i = 0; i < m; ++i) res.perm[i] = i;
    vector<int> &col_order = res.perm;
    res.d.init(p);
    int gne = 0, rem = p;
    while (rem > 0) {
        const int cs = gne;
        const int bcols = std::min(rem, nb);
        const int bm = m - cs;
        FactorResult br = factor_inner(a + cs + (i64)cs * lda, lda, bm, bcols, p - cs, o);
        const int bne = br.ne;
        if (cs > 0) {
            vector<double> tmp(bcols);
            for (int c = 0; c < cs; ++c) {
                for (int i = 0; i < bcols; ++i) tmp[i] = A(a, lda, cs + br.perm[i], c);
                for (int i = 0; i < bcols; ++i) A(a, lda, cs + i, c) = tmp[i];
            }
        }
        {
            vector<int> orig(col_order.begin() + cs, col_order.begin() + cs + bcols);
            for (int i = 0; i < bcols; ++i)
                if (br.perm[i] < bcols) col_order[cs + i] = orig[br.perm[i]];
        }
        if (bne < bcols) {
            const int nf = bcols - bne;
            for (int i = 0; i < nf; ++i) {
                const int fp = cs + bne + i;
                const int end = cs + rem - 1 - i;
                if (fp < end) {
                    swap_symmetric(a, lda, m, fp, end);
                    std::swap(col_order[fp], col_order[end]);
                }
            }
            res.st.ndel += nf;
        }
        res.d.copy_from_offset(br.d, gne, bne);
        res.st.n1 += br.st.n1;
        res.st.n2 += br.st.n2;
        if (br.st.max_l > res.st.max_l) res.st.max_l = br.st.max_l;
        gne += bne;
        rem -= bcols;
    }
    res.d.truncate(gne);
    res.ne = gne;
    return res;
}

// ---------------------------------------------------------------------------
// TPP (factor.rs:2268-2947).  One implementation over an m x ncols column-major view:
// ncols == m gives the square-matrix variant (swap_symmetric), ncols == k the
// rectangular small-leaf variant (swap_rect); see swap_rect (factor.rs:2521-2568).
// ---------------------------------------------------------------------------
struct View {
    double *a;
    i64 lda;
    int m, n;  // rows, stored columns
    double &operator()(int i, int j) const { return a[i + (i64)j * lda]; }
};

void swap_rect(const View &v, int i, int j) {
    if (i == j) return;
    if (i > j) std::swap(i, j);
    const int n = v.n;
    if (i < n && j < n) std::swap(v(i, i), v(j, j));
    for (int k = 0; k < std::min(i, n); ++k) std::swap(v(i, k), v(j, k));
    for (int k = i + 1; k < j; ++k) {
        if (k < n) std::swap(v(k, i), v(j, k));
        else if (i < n) break;
    }
    if (i < n && j < n)
        for (int k = j + 1; k < v.m; ++k) std::swap(v(k, i), v(k, j));
}

double tpp_sym_entry(const View &v, int row, int col) {
    if (col < v.n && row >= col) return v(row, col);
    return v(col, row);  // row < n && col >= row
}

bool tpp_is_col_small(const View &v, int idx, int from, int to, double small) {
    for (int c = from; c < std::min(idx, v.n); ++c)
        if (std::fabs(v(idx, c)) >= small) return false;
    if (idx < v.n)
        for (int r = idx; r < to; ++r)
            if (std::fabs(v(r, idx)) >= small) return false;
    return true;
}

int tpp_find_row_abs_max(const View &v, int p, int from, int to) {
    if (from >= to) return from;
    int best = from;
    double bv = std::fabs(tpp_sym_entry(v, p, from));
    for (int c = from + 1; c < to; ++c) {
        double x = std::fabs(tpp_sym_entry(v, p, c));
        if (x > bv) { best = c; bv = x; }
    }
    return best;
}

double tpp_find_rc_abs_max_exclude(const View &v, int col, int nelim, int m, int exclude) {
    double best = 0.0;
    for (int c = nelim; c < std::min(col, v.n); ++c) {
        if (c == exclude) continue;
        best = std::max(best, std::fabs(v(col, c)));
    }
    if (col < v.n)
        for (int r = col + 1; r < m; ++r) {
            if (r == exclude) continue;
            best = std::max(best, std::fabs(v(r, col)));
        }
    return best;
}

bool tpp_test_2x2(const View &v, int t, int p, double maxt, double maxp, double u, double small) {
    const double a11 = v(t, t), a21 = v(p, t), a22 = v(p, p);
    const double maxpiv = std::max(std::max(std::fabs(a11), std::fabs(a21)), std::fabs(a22));
    if (maxpiv < small) return false;
    const double detscale = 1.0 / maxpiv;
    const double d0 = (a11 * detscale) * a22;
    const double d1 = (a21 * detscale) * a21;
    const double det = d0 - d1;
    if (std::fabs(det) < std::max(std::max(small, std::fabs(d0 / 2.0)), std::fabs(d1 / 2.0)))
        return false;
    const double i11 = (a22 * detscale) / det;
    const double i12 = (-a21 * detscale) / det;
    const double i22 = (a11 * detscale) / det;
    if (std::max(maxt, maxp) < small) return true;
    const double x1 = std::fabs(i11) * maxt + std::fabs(i12) * maxp;
    const double x2 = std::fabs(i12) * maxt + std::fabs(i22) * maxp;
    return u * std::max(x1, x2) < 1.0;
}

double tpp_apply_1x1(const View &v, int nelim, int m, int p) {
    const double d = v(nelim, nelim);
    const double inv = 1.0 / d;
    double max_l = 0.0;
    for (int i = nelim + 1; i < m; ++i) {
        double l = v(i, nelim) * inv;
        v(i, nelim) = l;
        max_l = std::max(max_l, std::fabs(l));
    }
    const int ce = std::min(p, v.n);
    for (int j = nelim + 1; j < ce; ++j) {
        const double ldlj = v(j, nelim) * d;
        for (int i = j; i < m; ++i) v(i, j) -= v(i, nelim) * ldlj;
    }
    return max_l;
}

double tpp_apply_2x2(const View &v, int nelim, int m, int p) {
    const double a11 = v(nelim, nelim), a21 = v(nelim + 1, nelim), a22 = v(nelim + 1, nelim + 1);
    const double det = a11 * a22 - a21 * a21;
    const double inv_det = 1.0 / det;
    double max_l = 0.0;
    const int st = nelim + 2;
    for (int i = st; i < m; ++i) {
        const double ai1 = v(i, nelim), ai2 = v(i, nelim + 1);
        const double l1 = (ai1 * a22 - ai2 * a21) * inv_det;
        const double l2 = (ai2 * a11 - ai1 * a21) * inv_det;
        v(i, nelim) = l1;
        v(i, nelim + 1) = l2;
        max_l = std::max(std::max(max_l, std::fabs(l1)), std::fabs(l2));
    }
    const int ce = std::min(p, v.n);
    for (int j = st; j < ce; ++j) {
        const double wj1 = v(j, nelim) * a11 + v(j, nelim + 1) * a21;
        const double wj2 = v(j, nelim) * a21 + v(j, nelim + 1) * a22;
        for (int i = j; i < m; ++i) v(i, j) -= v(i, nelim) * wj1 + v(i, nelim + 1) * wj2;
    }
    v(nelim + 1, nelim) = 0.0;
    return max_l;
}

// tpp_factor (factor.rs:2335-2506) / tpp_factor_rect (factor.rs:2761-2947) loop body.
// Returns the final nelim.
int tpp_core(const View &v, int start, int n, vector<int> &col_order, Diag &d, Stats &st,
             double u, double small) {
    const int m = v.m;
    int nelim = start;
    while (nelim < n) {
        if (tpp_is_col_small(v, nelim, nelim, m, small)) {
            d.set1(nelim, 0.0);
            st.n1 += 1;
            if (nelim < v.n)
                for (int r = nelim + 1; r < m; ++r) v(r, nelim) = 0.0;
            nelim += 1;
            continue;
        }
        bool found = false;
        for (int p = nelim + 1; p < n; ++p) {
            if (tpp_is_col_small(v, p, nelim, m, small)) {
                if (p != nelim) {
                    swap_rect(v, p, nelim);
                    std::swap(col_order[p], col_order[nelim]);
                }
                d.set1(nelim, 0.0);
                st.n1 += 1;
                for (int r = nelim + 1; r < m; ++r) v(r, nelim) = 0.0;
                nelim += 1;
                found = true;
                break;
            }
            const int t = tpp_find_row_abs_max(v, p, nelim, p);
            const double maxt = tpp_find_rc_abs_max_exclude(v, t, nelim, m, p);
            const double maxp = tpp_find_rc_abs_max_exclude(v, p, nelim, m, t);
            if (tpp_test_2x2(v, t, p, maxt, maxp, u, small)) {
                if (t != nelim) {
                    swap_rect(v, t, nelim);
                    std::swap(col_order[t], col_order[nelim]);
                }
                const int np = (p == nelim) ? t : p;
                if (np != nelim + 1) {
                    swap_rect(v, np, nelim + 1);
                    std::swap(col_order[np], col_order[nelim + 1]);
                }
                const double d11 = v(nelim, nelim), d21 = v(nelim + 1, nelim),
                             d22 = v(nelim + 1, nelim + 1);
                const double ml = tpp_apply_2x2(v, nelim, m, n);
                d.set2(nelim, d11, d21, d22);
                st.n2 += 1;
                st.max_l = std::max(st.max_l, ml);
                nelim += 2;
                found = true;
                break;
            }
            const double maxp_t = std::max(maxp, std::fabs(tpp_sym_entry(v, p, t)));
            if (std::fabs(v(p, p)) >= u * maxp_t) {
                if (p != nelim) {
                    swap_rect(v, p, nelim);
                    std::swap(col_order[p], col_order[nelim]);
                }
                const double ml = tpp_apply_1x1(v, nelim, m, n);
                d.set1(nelim, v(nelim, nelim));
                st.n1 += 1;
                st.max_l = std::max(st.max_l, ml);
                nelim += 1;
                found = true;
                break;
            }
        }
        if (!found) {
            const double maxp = tpp_find_rc_abs_max_exclude(v, nelim, nelim, m, -1);
            if (std::fabs(v(nelim, nelim)) >= u * maxp) {
                const double ml = tpp_apply_1x1(v, nelim, m, n);
                d.set1(nelim, v(nelim, nelim));
                st.n1 += 1;
                st.max_l = std::max(st.max_l, ml);
                nelim += 1;
            } else {
                break;
            }
        }
    }
    return nelim;
}

FactorResult tpp_factor_rect(const View &v, int n, const Opts &o) {
    FactorResult r;
    r.perm.resize(v.m);
    for (int i = 0; i < v.m; ++i) r.perm[i] = i;
    r.d.init(n);
    const int ne = tpp_core(v, 0, n, r.perm, r.d, r.st, o.threshold, o.small);
    r.d.truncate(ne);
    r.st.ndel = n - ne;
    r.ne = ne;
    return r;
}

bool valid_opts(const Opts &o) {
    if (!(o.threshold > 0.0 && o.threshold <= 1.0)) return false;
    if (o.small < 0.0) return false;
    if (o.nb <= 0 || o.ib <= 0 || o.ib > o.nb) return false;
    return true;
}

// aptp_factor_in_place (factor.rs:301-404); options already validated.
FactorResult aptp_factor_in_place(double *a, i64 lda, int m, int p, const Opts &o) {
    FactorResult res;
    if (p < o.ib) {
        View v{a, lda, m, m};
        res = tpp_factor_rect(v, p, o);
    } else if (p > o.nb) {
        res = two_level_factor(a, lda, m, p, o);
    } else {
        res = factor_inner(a, lda, m, p, p, o);
    }
    if (res.ne < p && o.tpp) {
        const int q = res.ne;
        res.d.grow(p);
        View v{a, lda, m, m};
        const int ne = tpp_core(v, q, p, res.perm, res.d, res.st, o.threshold, o.small);
        res.ne = ne;
        res.d.truncate(ne);
        res.st.ndel = p - ne;
    }
    return res;
}

// ---------------------------------------------------------------------------
// Multifrontal driver (numeric.rs)
// ---------------------------------------------------------------------------
struct Range { int lo, hi; };
struct SnInfo {  // SupernodeInfo (numeric.rs:431-448)
    int col_begin, col_end;
    vector<int> pattern;
    int parent;  // -1 root
    vector<Range> owned;
    bool in_small_leaf = false;
    int n_owned() const {
        int s = 0;
        for (auto &r : owned) s += r.hi - r.lo;
        return s;
    }
};

// amalgamation.rs:38-54
bool do_merge(int pn, int pcc, int cn, int ccc, int nemin) {
    if (pn == 1 && pcc == std::max(ccc - 1, 0)) return true;
    if (pn < nemin && cn < nemin) return true;
    return false;
}

vector<int> sorted_union_excluding(const vector<int> &a, const vector<int> &b, int lo, int hi) {
    vector<int> r;
    r.reserve(a.size() + b.size());
    size_t ia = 0, ib = 0;
    auto push = [&](int v) { if (!(v >= lo && v < hi)) r.push_back(v); };
    while (ia < a.size() && ib < b.size()) {
        int va = a[ia], vb = b[ib], nx;
        if (va < vb) { ++ia; nx = va; }
        else if (vb < va) { ++ib; nx = vb; }
        else { ++ia; ++ib; nx = va; }
        push(nx);
    }
    for (; ia < a.size(); ++ia) push(a[ia]);
    for (; ib < b.size(); ++ib) push(b[ib]);
    return r;
}

// amalgamate (amalgamation.rs:115-231)
vector<SnInfo> amalgamate(vector<SnInfo> sns, int nemin) {
    const int n = (int)sns.size();
    if (n <= 1) return sns;
    vector<char> deleted(n, 0);
    vector<int> nel(n);
    for (int s = 0; s < n; ++s) nel[s] = sns[s].col_end - sns[s].col_begin;
    vector<vector<int>> ch(n);
    for (int s = 0; s < n; ++s)
        if (sns[s].parent >= 0) ch[sns[s].parent].push_back(s);
    for (int p = 0; p < n; ++p) {
        if (deleted[p]) continue;
        vector<int> pch;
        pch.swap(ch[p]);
        for (int c : pch) {
            if (deleted[c]) continue;
            const int pn = nel[p], pcc = nel[p] + (int)sns[p].pattern.size();
            const int cn = nel[c], ccc = nel[c] + (int)sns[c].pattern.size();
            if (do_merge(pn, pcc, cn, ccc, nemin)) {
                const int nb = std::min(sns[p].col_begin, sns[c].col_begin);
                const int ne = std::max(sns[p].col_end, sns[c].col_end);
                vector<int> cp, pp;
                cp.swap(sns[c].pattern);
                pp.swap(sns[p].pattern);
                sns[p].pattern = sorted_union_excluding(pp, cp, nb, ne);
                vector<Range> co;
                co.swap(sns[c].owned);
                sns[p].owned.insert(sns[p].owned.end(), co.begin(), co.end());
                sns[p].col_begin = nb;
                sns[p].col_end = ne;
                nel[p] += nel[c];
                vector<int> cch;
                cch.swap(ch[c]);
                for (int gc : cch)
                    if (!deleted[gc]) sns[gc].parent = p;
                ch[p].insert(ch[p].end(), cch.begin(), cch.end());
                deleted[c] = 1;
            }
        }
        for (int c : pch)
            if (!deleted[c]) ch[p].push_back(c);
    }
    vector<int> o2n(n, 0);
    int ni = 0;
    for (int s = 0; s < n; ++s)
        if (!deleted[s]) o2n[s] = ni++;
    vector<SnInfo> out;
    out.reserve(ni);
    for (int s = 0; s < n; ++s)
        if (!deleted[s]) {
            SnInfo x = std::move(sns[s]);
            if (x.parent >= 0) x.parent = o2n[x.parent];
            out.push_back(std::move(x));
        }
    return out;
}

vector<vector<int>> build_children_map(const vector<SnInfo> &s) {
    vector<vector<int>> ch(s.size());
    for (size_t i = 0; i < s.size(); ++i)
        if (s[i].parent >= 0) ch[s[i].parent].push_back((int)i);
    return ch;
}

struct SmallLeaf { int root; vector<int> nodes; int max_front; int parent_of_root; };

// classify_small_leaf_subtrees (numeric.rs:2278-2369)
vector<SmallLeaf> classify_small_leaf(vector<SnInfo> &sns, const vector<vector<int>> &ch,
                                      int threshold) {
    vector<SmallLeaf> out;
    if (threshold == 0) return out;
    const int ns = (int)sns.size();
    for (int s = 0; s < ns; ++s) {
        const int fs = sns[s].n_owned() + (int)sns[s].pattern.size();
        if (fs >= threshold) { sns[s].in_small_leaf = false; continue; }
        if (ch[s].empty()) sns[s].in_small_leaf = true;
        else {
            bool all = true;
            for (int c : ch[s]) all = all && sns[c].in_small_leaf;
            sns[s].in_small_leaf = all;
        }
    }
    for (int s = 0; s < ns; ++s) {
        if (!sns[s].in_small_leaf) continue;
        const bool is_root = sns[s].parent < 0 || !sns[sns[s].parent].in_small_leaf;
        if (!is_root) continue;
        vector<int> nodes, stack{s}, visit;
        while (!stack.empty()) {
            int nd = stack.back();
            stack.pop_back();
            visit.push_back(nd);
            for (int c : ch[nd])
                if (sns[c].in_small_leaf) stack.push_back(c);
        }
        while (!visit.empty()) { nodes.push_back(visit.back()); visit.pop_back(); }
        if (nodes.size() < 2) { sns[s].in_small_leaf = false; continue; }
        int mf = 0;
        for (int nd : nodes) mf = std::max(mf, sns[nd].n_owned() + (int)sns[nd].pattern.size());
        out.push_back(SmallLeaf{s, nodes, mf, sns[s].parent});
    }
    return out;
}

struct Contribution {  // ContributionBlock (numeric.rs:516-523)
    vector<double> data;  // size x size column-major (lower)
    int size = 0;
    vector<int> rows;
    int ndelayed = 0;
    bool present = false;
};

struct FrontFactorsO {  // FrontFactors (numeric.rs:543-596)
    vector<double> l11, l21;
    Diag d11;
    vector<int> local_perm, col_idx, row_idx;
    int ne = 0;
};

struct SnStats { int m = 0, k = 0, ne = 0, n1 = 0, n2 = 0; double max_l = 0.0; };

struct Matrix {
    i64 n;
    const i64 *cp, *ri;
    const double *v;
};

struct AmapEntry { uint32_t src, dest, sr, sc; };

struct Ctx {
    Matrix A;
    vector<int> perm_fwd, perm_inv;
    const double *scaling = nullptr;
    Opts o;
    vector<SnInfo> sns;
    vector<vector<int>> children;
    vector<SmallLeaf> small_leaves;
    vector<vector<AmapEntry>> amap;      // per supernode
    vector<vector<vector<uint32_t>>> ea; // [parent][child ordinal] -> parent-local rows
    vector<FrontFactorsO> ff;
    vector<SnStats> st;
    i64 sn_before = 0, merges = 0;
    vector<Contribution> contrib;
    bool keep = false;             // keep every contribution block (kept[s]) for inspection
    vector<Contribution> kept;
};

const int NOT_IN_FRONT = -1;

// Global-to-local row table of the calling thread (fronts factored concurrently each need
// their own); every front resets the entries it set before returning.
vector<int> &g2l_table(i64 n) {
    static thread_local vector<int> t;
    if ((i64)t.size() != n) t.assign(n, NOT_IN_FRONT);
    return t;
}


// build_assembly_maps (numeric.rs:264-408)
void build_assembly_maps(Ctx &c) {
    const int ns = (int)c.sns.size();
    const i64 n = c.A.n;
    vector<int> g2l(n, NOT_IN_FRONT);
    c.amap.assign(ns, {});
    c.ea.assign(ns, {});
    for (int s = 0; s < ns; ++s) {
        const SnInfo &sn = c.sns[s];
        vector<int> rows;
        for (auto &r : sn.owned)
            for (int j = r.lo; j < r.hi; ++j) rows.push_back(j);
        const int total_owned = (int)rows.size();
        rows.insert(rows.end(), sn.pattern.begin(), sn.pattern.end());
        const int m = (int)rows.size();
        for (int i = 0; i < m; ++i) g2l[rows[i]] = i;
        for (auto &r : sn.owned)
            for (int pj = r.lo; pj < r.hi; ++pj) {
                const int oc = c.perm_fwd[pj];
                const int lc = g2l[pj];
                for (i64 idx = c.A.cp[oc]; idx < c.A.cp[oc + 1]; ++idx) {
                    const int orow = (int)c.A.ri[idx];
                    if (orow < oc) {
                        const int peer = g2l[c.perm_inv[orow]];
                        if (peer != NOT_IN_FRONT && peer < total_owned) continue;
                    }
                    const int gr = c.perm_inv[orow];
                    const int lr = g2l[gr];
                    if (lr == NOT_IN_FRONT) continue;
                    int dr, dc;
                    if (lr >= lc) { dr = lr; dc = lc; } else { dr = lc; dc = lr; }
                    c.amap[s].push_back(AmapEntry{(uint32_t)idx, (uint32_t)((i64)dc * m + dr),
                                                  (uint32_t)gr, (uint32_t)pj});
                }
            }
        for (int g : rows) g2l[g] = NOT_IN_FRONT;
    }
    for (int s = 0; s < ns; ++s) {
        const SnInfo &sn = c.sns[s];
        vector<int> rows;
        for (auto &r : sn.owned)
            for (int j = r.lo; j < r.hi; ++j) rows.push_back(j);
        rows.insert(rows.end(), sn.pattern.begin(), sn.pattern.end());
        for (size_t i = 0; i < rows.size(); ++i) g2l[rows[i]] = (int)i;
        for (int ch : c.children[s]) {
            vector<uint32_t> mp;
            for (int r : c.sns[ch].pattern) mp.push_back((uint32_t)g2l[r]);
            c.ea[s].push_back(std::move(mp));
        }
        for (int g : rows) g2l[g] = NOT_IN_FRONT;
    }
}

// scatter_original_entries_multi (numeric.rs:2376-2435); writes via callback add(r,c,val)
template <class F>
void scatter_fallback(Ctx &c, const SnInfo &sn, int k, F add) {
    const vector<int> &g2l = g2l_table(c.A.n);
    const int total_owned = sn.n_owned();
    for (auto &r : sn.owned)
        for (int pj = r.lo; pj < r.hi; ++pj) {
            const int oc = c.perm_fwd[pj];
            const int lc = g2l[pj];
            for (i64 idx = c.A.cp[oc]; idx < c.A.cp[oc + 1]; ++idx) {
                const int orow = (int)c.A.ri[idx];
                if (orow < oc) {
                    const int peer = g2l[c.perm_inv[orow]];
                    if (peer != NOT_IN_FRONT && peer < total_owned) continue;
                }
                const int gr = c.perm_inv[orow];
                const int lr = g2l[gr];
                if (lr == NOT_IN_FRONT) continue;
                if (lr >= total_owned && lr < k) continue;
                double val = c.A.v[idx];
                if (c.scaling) val *= c.scaling[c.perm_inv[orow]] * c.scaling[c.perm_inv[oc]];
                if (lr >= lc) add(lr, lc, val); else add(lc, lr, val);
            }
        }
}

// extract_front_factors (numeric.rs:2579-2713) from an m x ncols storage with leading dim lda
FrontFactorsO extract_front_factors(const double *F, i64 lda, int m, int k,
                                    const vector<int> &rows, const FactorResult &r) {
    FrontFactorsO ff;
    const int ne = r.ne;
    ff.ne = ne;
    ff.l11.assign((size_t)ne * ne, 0.0);
    int col = 0;
    while (col < ne) {
        ff.l11[col + (size_t)col * ne] = 1.0;
        if (r.d.kind[col] == 1) {
            for (int i = col + 1; i < ne; ++i) ff.l11[i + (size_t)col * ne] = F[i + col * lda];
            col += 1;
        } else if (r.d.two_first(col)) {
            ff.l11[(col + 1) + (size_t)(col + 1) * ne] = 1.0;
            for (int i = col + 2; i < ne; ++i) {
                ff.l11[i + (size_t)col * ne] = F[i + col * lda];
                ff.l11[i + (size_t)(col + 1) * ne] = F[i + (col + 1) * lda];
            }
            col += 2;
        } else {
            col += 1;
        }
    }
    ff.d11.init(ne);
    col = 0;
    while (col < ne) {
        if (r.d.kind[col] == 1) { ff.d11.set1(col, r.d.diag[col]); col += 1; }
        else { ff.d11.set2(col, r.d.diag[col], r.d.off[col], r.d.diag[col + 1]); col += 2; }
    }
    const int rr = m - ne;
    ff.l21.assign((size_t)rr * ne, 0.0);
    if (ne > 0 && rr > 0)
        for (int j = 0; j < ne; ++j)
            for (int i = 0; i < rr; ++i) ff.l21[i + (size_t)j * rr] = F[(ne + i) + j * lda];
    ff.local_perm.assign(r.perm.begin(), r.perm.begin() + k);
    for (int i = 0; i < ne; ++i) ff.col_idx.push_back(rows[ff.local_perm[i]]);
    for (int i = ne; i < k; ++i) ff.row_idx.push_back(rows[r.perm[i]]);
    for (int i = k; i < m; ++i) ff.row_idx.push_back(rows[i]);
    return ff;
}

// extend_add_mapped (numeric.rs:2507-2549): parent[li, lj] += child[i, j] over the child's
// lower triangle, li = map[i], lj = map[j]; the entry goes to (max, min) of the two local rows.
void extend_add_mapped(double *F, i64 ldf, const double *cb, int sz, const uint32_t *mp) {
    for (int j = 0; j < sz; ++j) {
        const int lj = (int)mp[j];
        for (int i = j; i < sz; ++i) {
            const int li = (int)mp[i];
            const double val = cb[i + (size_t)j * sz];
            if (li >= lj) F[li + (size_t)lj * ldf] += val;
            else F[lj + (size_t)li * ldf] += val;
        }
    }
}

// extend_add (numeric.rs:2445-2485): as extend_add_mapped with the rows mapped through the
// parent's global-to-local table, and exact zeros skipped.
void extend_add_g2l(double *F, i64 ldf, const double *cb, int sz, const vector<int> &rows,
                    const vector<int> &g2l) {
    for (int j = 0; j < sz; ++j) {
        const int lj = g2l[rows[j]];
        for (int i = j; i < sz; ++i) {
            const double val = cb[i + (size_t)j * sz];
            if (val != 0.0) {
                const int li = g2l[rows[i]];
                if (li >= lj) F[li + (size_t)lj * ldf] += val;
                else F[lj + (size_t)li * ldf] += val;
            }
        }
    }
}

// The contribution block of front s is kept for inspection when requested (it is consumed by
// the parent's extend-add otherwise).
void keep_contribution(Ctx &c, int s) {
    if (c.keep) c.kept[s] = c.contrib[s];
}

// factor_single_supernode (numeric.rs:1226-1533)
void factor_single(Ctx &c, int s) {
    vector<int> &g2l = g2l_table(c.A.n);
    const SnInfo &sn = c.sns[s];
    vector<int> delayed;
    for (int ch : c.children[s])
        if (c.contrib[ch].present)
            delayed.insert(delayed.end(), c.contrib[ch].rows.begin(),
                           c.contrib[ch].rows.begin() + c.contrib[ch].ndelayed);
    const int k = sn.n_owned() + (int)delayed.size();
    vector<int> rows;
    for (auto &r : sn.owned)
        for (int j = r.lo; j < r.hi; ++j) rows.push_back(j);
    rows.insert(rows.end(), delayed.begin(), delayed.end());
    rows.insert(rows.end(), sn.pattern.begin(), sn.pattern.end());
    const int m = (int)rows.size();
    vector<double> F((size_t)m * m, 0.0);
    for (int i = 0; i < m; ++i) g2l[rows[i]] = i;
    const int ndelay_in = (int)delayed.size();
    if (ndelay_in == 0) {
        for (const AmapEntry &e : c.amap[s]) {
            const i64 dl = e.dest;
            const int dc = (int)(dl / m), dr = (int)(dl % m);
            double val = c.A.v[e.src];
            if (c.scaling) val = c.A.v[e.src] * c.scaling[e.sr] * c.scaling[e.sc];
            F[dr + (size_t)dc * m] += val;
        }
    } else {
        scatter_fallback(c, sn, k, [&](int r, int cc, double v) { F[r + (size_t)cc * m] += v; });
    }
    // extend-add (numeric.rs:1353-1377)
    for (size_t ci = 0; ci < c.children[s].size(); ++ci) {
        const int ch = c.children[s][ci];
        Contribution &cb = c.contrib[ch];
        if (!cb.present) continue;
        const int sz = cb.size;
        if (cb.ndelayed > 0 || ndelay_in > 0) {
            extend_add_g2l(F.data(), m, cb.data.data(), sz, cb.rows, g2l);
        } else {
            extend_add_mapped(F.data(), m, cb.data.data(), sz, c.ea[s][ci].data());
        }
        cb.data.clear();
        cb.data.shrink_to_fit();
        cb.present = false;
    }
    FactorResult r = aptp_factor_in_place(F.data(), m, m, k, c.o);
    const int ne = r.ne;
    const int nfs = m - k;
    Contribution out;
    if (sn.parent >= 0 && ne < m) {
        // compute_contribution_gemm (factor.rs:1590-1652) + extract_contribution (2739-2809)
        const int size = m - ne, nd = k - ne;
        out.present = true;
        out.size = size;
        out.ndelayed = nd;
        out.data.assign((size_t)size * size, 0.0);
        for (int j = 0; j < nd; ++j)
            for (int i = j; i < nd; ++i)
                out.data[i + (size_t)j * size] = F[(ne + i) + (size_t)(ne + j) * m];
        for (int j = 0; j < nd; ++j)
            for (int i = 0; i < nfs; ++i)
                out.data[(nd + i) + (size_t)j * size] = F[(k + i) + (size_t)(ne + j) * m];
        double *C = out.data.data() + nd + (size_t)nd * size;
        for (int j = 0; j < nfs; ++j)
            for (int i = j; i < nfs; ++i) C[i + (size_t)j * size] = F[(k + i) + (size_t)(k + j) * m];
        if (nfs > 0 && ne > 0) {
            vector<double> W((size_t)nfs * ne, 0.0);
            compute_ld(&F[k], m, nfs, r.d, ne, W.data(), nfs);
            gemm_sub_wlt(C, size, nfs, nfs, W.data(), nfs, &F[k], m, ne, true);
        }
        for (int i = ne; i < k; ++i) out.rows.push_back(rows[r.perm[i]]);
        for (int i = k; i < m; ++i) out.rows.push_back(rows[i]);
    }
    c.ff[s] = extract_front_factors(F.data(), m, m, k, rows, r);
    SnStats &ss = c.st[s];
    ss.m = m; ss.k = k; ss.ne = ne; ss.n1 = (int)r.st.n1; ss.n2 = (int)r.st.n2; ss.max_l = r.st.max_l;
    for (int g : rows) g2l[g] = NOT_IN_FRONT;
    c.contrib[s] = std::move(out);
    keep_contribution(c, s);
}

// factor_small_leaf_subtree (numeric.rs:1562-1938)
void factor_small_leaf(Ctx &c, const SmallLeaf &sl) {
    vector<int> &g2l = g2l_table(c.A.n);
    for (int node : sl.nodes) {
        const SnInfo &sn = c.sns[node];
        vector<int> delayed;
        for (int ch : c.children[node])
            if (c.contrib[ch].present)
                delayed.insert(delayed.end(), c.contrib[ch].rows.begin(),
                               c.contrib[ch].rows.begin() + c.contrib[ch].ndelayed);
        const int total_owned = sn.n_owned();
        const int k = total_owned + (int)delayed.size();
        vector<int> rows;
        for (auto &r : sn.owned)
            for (int j = r.lo; j < r.hi; ++j) rows.push_back(j);
        rows.insert(rows.end(), delayed.begin(), delayed.end());
        rows.insert(rows.end(), sn.pattern.begin(), sn.pattern.end());
        const int m = (int)rows.size();
        vector<double> L((size_t)m * std::max(k, 1), 0.0);  // m x k rectangular
        for (int i = 0; i < m; ++i) g2l[rows[i]] = i;
        const int ndelay_in = (int)delayed.size();
        const int nfs = m - k;
        const bool has_nfs = nfs > 0 && sn.parent >= 0;
        vector<double> C;
        if (has_nfs) C.assign((size_t)nfs * nfs, 0.0);
        auto add = [&](int r, int cc, double val) {
            if (cc < k) L[r + (size_t)cc * m] += val;
            else if (has_nfs && r >= k) C[(r - k) + (size_t)(cc - k) * nfs] += val;
        };
        if (ndelay_in == 0) {
            for (const AmapEntry &e : c.amap[node]) {
                const i64 dl = e.dest;
                const int dc = (int)(dl / m), dr = (int)(dl % m);
                double val = c.A.v[e.src];
                if (c.scaling) val = c.A.v[e.src] * c.scaling[e.sr] * c.scaling[e.sc];
                add(dr, dc, val);
            }
        } else {
            scatter_fallback(c, sn, k, add);
        }
        for (int ch : c.children[node]) {
            Contribution &cb = c.contrib[ch];
            if (!cb.present) continue;
            const int sz = cb.size;
            for (int i = 0; i < sz; ++i) {
                const int li0 = g2l[cb.rows[i]];
                for (int j = 0; j <= i; ++j) {
                    const int lj0 = g2l[cb.rows[j]];
                    const double val = cb.data[i + (size_t)j * sz];
                    if (val != 0.0) {
                        int li = li0, lj = lj0;
                        if (li < lj) std::swap(li, lj);
                        if (lj < k) L[li + (size_t)lj * m] += val;
                        else if (has_nfs) C[(li - k) + (size_t)(lj - k) * nfs] += val;
                    }
                }
            }
            cb.data.clear();
            cb.data.shrink_to_fit();
            cb.present = false;
        }
        View v{L.data(), m, m, k};
        FactorResult r = tpp_factor_rect(v, k, c.o);
        const int ne = r.ne;
        Contribution out;
        if (sn.parent >= 0 && ne < m) {
            const int size = m - ne, nd = k - ne;
            out.present = true;
            out.size = size;
            out.ndelayed = nd;
            out.data.assign((size_t)size * size, 0.0);
            if (nfs > 0 && ne > 0) {
                vector<double> W((size_t)nfs * ne, 0.0);
                compute_ld(&L[k], m, nfs, r.d, ne, W.data(), nfs);
                gemm_sub_wlt(C.data(), nfs, nfs, nfs, W.data(), nfs, &L[k], m, ne, true);
            }
            for (int j = 0; j < nd; ++j)
                for (int i = j; i < nd; ++i)
                    out.data[i + (size_t)j * size] = L[(ne + i) + (size_t)(ne + j) * m];
            for (int j = 0; j < nd; ++j)
                for (int i = 0; i < nfs; ++i)
                    out.data[(nd + i) + (size_t)j * size] = L[(k + i) + (size_t)(ne + j) * m];
            for (int j = 0; j < nfs; ++j)
                for (int i = j; i < nfs; ++i)
                    out.data[(nd + i) + (size_t)(nd + j) * size] = C[i + (size_t)j * nfs];
            for (int i = ne; i < k; ++i) out.rows.push_back(rows[r.perm[i]]);
            for (int i = k; i < m; ++i) out.rows.push_back(rows[i]);
        }
        c.ff[node] = extract_front_factors(L.data(), m, m, k, rows, r);
        SnStats &ss = c.st[node];
        ss.m = m; ss.k = k; ss.ne = ne; ss.n1 = (int)r.st.n1; ss.n2 = (int)r.st.n2; ss.max_l = r.st.max_l;
        for (int g : rows) g2l[g] = NOT_IN_FRONT;
        c.contrib[node] = std::move(out);
        keep_contribution(c, node);
    }
}

// factor_tree_levelset (numeric.rs:1957-2163), sequential mode
// With several threads, independent small-leaf subtrees and the fronts of one wave run
// concurrently (numeric.rs:2111-2142); a wave with fewer fronts than threads runs its fronts
// one after another, each splitting its large products over the threads instead.  Every front
// is computed the same way either way, so results do not depend on the thread count.
void factor_tree_levelset(Ctx &c) {
    const int ns = (int)c.sns.size();
    c.contrib.assign(ns, Contribution());
    if (c.keep) c.kept.assign(ns, Contribution());
    vector<int> remaining(ns);
    for (int s = 0; s < ns; ++s) remaining[s] = (int)c.children[s].size();
    const int nsl = (int)c.small_leaves.size();
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) if (nsl > 1 && omp_get_max_threads() > 1)
#endif
    for (int q = 0; q < nsl; ++q) factor_small_leaf(c, c.small_leaves[q]);
    for (const SmallLeaf &sl : c.small_leaves)
        if (sl.parent_of_root >= 0) remaining[sl.parent_of_root] -= 1;
    vector<int> ready;
    for (int s = 0; s < ns; ++s)
        if (remaining[s] == 0 && !c.sns[s].in_small_leaf) ready.push_back(s);
    while (!ready.empty()) {
        vector<int> next;
        const int nr = (int)ready.size();
#ifdef _OPENMP
        const bool wide = omp_get_max_threads() > 1 && nr >= omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) if (wide)
#endif
        for (int q = 0; q < nr; ++q) factor_single(c, ready[q]);
        for (int s : ready) {
            const int p = c.sns[s].parent;
            if (p >= 0) {
                remaining[p] -= 1;
                if (remaining[p] == 0) next.push_back(p);
            }
        }
        ready.swap(next);
    }
}

// MixedDiagonal::solve_in_place (diagonal.rs:312-365) on nrhs columns
void d_solve(const Diag &d, double *x, i64 ldx, int nrhs) {
    int col = 0;
    while (col < d.n) {
        if (d.kind[col] == 1) {
            const double dv = d.diag[col];
            for (int r = 0; r < nrhs; ++r) {
                double &xv = x[col + r * ldx];
                xv = (dv == 0.0) ? 0.0 : xv / dv;
            }
            col += 1;
        } else {
            const int p = d.partner[col];
            if (p > col) {
                const double a = d.diag[col], b = d.off[col], cc = d.diag[p];
                const double det = a * cc - b * b;
                for (int r = 0; r < nrhs; ++r) {
                    double &x1 = x[col + r * ldx];
                    double &x2 = x[p + r * ldx];
                    if (det == 0.0) { x1 = 0.0; x2 = 0.0; }
                    else {
                        const double r1 = x1, r2 = x2;
                        x1 = (cc * r1 - b * r2) / det;
                        x2 = (a * r2 - b * r1) / det;
                    }
                }
            }
            col += 1;
        }
    }
}

// compute_inertia (diagonal.rs:391-461)
void d_inertia(const Diag &d, i64 *out) {
    int col = 0;
    while (col < d.n) {
        if (d.kind[col] == 1) {
            const double v = d.diag[col];
            if (v > 0.0) out[0] += 1; else if (v < 0.0) out[1] += 1; else out[2] += 1;
            col += 1;
        } else {
            const int p = d.partner[col];
            if (p > col) {
                const double a = d.diag[col], b = d.off[col], cc = d.diag[p];
                const double det = a * cc - b * b, tr = a + cc;
                if (det > 0.0) { if (tr > 0.0) out[0] += 2; else out[1] += 2; }
                else if (det < 0.0) { out[0] += 1; out[1] += 1; }
                else {
                    if (tr > 0.0) { out[0] += 1; out[2] += 1; }
                    else if (tr < 0.0) { out[1] += 1; out[2] += 1; }
                    else out[2] += 2;
                }
            }
            col += 1;
        }
    }
}

// aptp_solve (solve.rs:57-144) in permuted coordinates, nrhs columns.
// L11 solves and L21 products use ascending-index sums (faer's TRSV/GEMV order is not fixed).
void aptp_solve(Ctx &c, double *x, i64 ldx, int nrhs) {
    const int ns = (int)c.ff.size();
    vector<double> w, w2;
    for (int s = 0; s < ns; ++s) {  // forward (solve.rs:161-225)
        const FrontFactorsO &f = c.ff[s];
        const int ne = f.ne, r = (int)f.row_idx.size();
        if (ne == 0) continue;
        w.assign(ne, 0.0);
        w2.assign(r, 0.0);
        for (int q = 0; q < nrhs; ++q) {
            double *xq = x + q * ldx;
            for (int i = 0; i < ne; ++i) w[i] = xq[f.col_idx[i]];
            for (int j = 0; j < ne; ++j)
                for (int i = j + 1; i < ne; ++i) w[i] -= f.l11[i + (size_t)j * ne] * w[j];
            for (int i = 0; i < ne; ++i) xq[f.col_idx[i]] = w[i];
            if (r > 0) {
                for (int i = 0; i < r; ++i) w2[i] = 0.0;
                for (int j = 0; j < ne; ++j) {
                    const double wj = w[j];
                    const double *lc = &f.l21[(size_t)j * r];
                    for (int i = 0; i < r; ++i) w2[i] += lc[i] * wj;
                }
                for (int i = 0; i < r; ++i) xq[f.row_idx[i]] -= w2[i];
            }
        }
    }
    for (int s = 0; s < ns; ++s) {  // diagonal (solve.rs:232-252)
        const FrontFactorsO &f = c.ff[s];
        const int ne = f.ne;
        if (ne == 0) continue;
        w.assign((size_t)ne * nrhs, 0.0);
        for (int q = 0; q < nrhs; ++q)
            for (int i = 0; i < ne; ++i) w[i + (size_t)q * ne] = x[f.col_idx[i] + q * ldx];
        d_solve(f.d11, w.data(), ne, nrhs);
        for (int q = 0; q < nrhs; ++q)
            for (int i = 0; i < ne; ++i) x[f.col_idx[i] + q * ldx] = w[i + (size_t)q * ne];
    }
    for (int s = ns - 1; s >= 0; --s) {  // backward (solve.rs:261-326)
        const FrontFactorsO &f = c.ff[s];
        const int ne = f.ne, r = (int)f.row_idx.size();
        if (ne == 0) continue;
        w.assign(ne, 0.0);
        w2.assign(r, 0.0);
        for (int q = 0; q < nrhs; ++q) {
            double *xq = x + q * ldx;
            for (int i = 0; i < ne; ++i) w[i] = xq[f.col_idx[i]];
            if (r > 0) {
                for (int i = 0; i < r; ++i) w2[i] = xq[f.row_idx[i]];
                for (int j = 0; j < ne; ++j) {
                    const double *lc = &f.l21[(size_t)j * r];
                    double s2 = 0.0;
                    for (int i = 0; i < r; ++i) s2 += lc[i] * w2[i];
                    w[j] -= s2;
                }
            }
            for (int j = ne - 1; j >= 0; --j) {
                double s2 = 0.0;
                for (int i = j + 1; i < ne; ++i) s2 += f.l11[i + (size_t)j * ne] * w[i];
                w[j] -= s2;
            }
            for (int i = 0; i < ne; ++i) xq[f.col_idx[i]] = w[i];
        }
    }
}

bool g_keep_contributions = false;

Opts to_opts(const orc_options *o) {
    Opts r;
    if (o) {
        r.threshold = o->threshold;
        r.small = o->small;
        r.nb = o->outer_block_size;
        r.ib = o->inner_block_size;
        r.tpp = o->failed_pivot_tpp != 0;
        r.nemin = o->nemin;
        r.small_leaf = o->small_leaf_threshold;
    }
    return r;
}

vector<SnInfo> input_structure(int32_t nsn, const int32_t *b, const int32_t *e,
                               const int64_t *pp, const int32_t *pat, const int32_t *par) {
    vector<SnInfo> s(nsn);
    for (int i = 0; i < nsn; ++i) {
        s[i].col_begin = b[i];
        s[i].col_end = e[i];
        s[i].pattern.assign(pat + pp[i], pat + pp[i + 1]);
        s[i].parent = par[i];
        s[i].owned = {Range{b[i], e[i]}};
    }
    return s;
}

}  // namespace

extern "C" {

int orc_aptp_factor_in_place(double *a, int64_t lda, int32_t m, int32_t p, const orc_options *o,
                             int32_t *perm, int8_t *dkind, double *ddiag, double *doff,
                             int64_t *stats, double *max_l) {
    Opts op = to_opts(o);
    if (p > m || !valid_opts(op)) return -1;
    FactorResult r = aptp_factor_in_place(a, lda, m, p, op);
    if (perm)
        for (int i = 0; i < m; ++i) perm[i] = r.perm[i];
    for (int i = 0; i < p; ++i) {
        if (dkind) dkind[i] = (i < r.d.n) ? r.d.kind[i] : 0;
        if (ddiag) ddiag[i] = (i < r.d.n) ? r.d.diag[i] : 0.0;
        if (doff) doff[i] = (i < r.d.n) ? r.d.off[i] : 0.0;
    }
    if (stats) {
        stats[0] = r.ne; stats[1] = r.st.n1; stats[2] = r.st.n2; stats[3] = r.st.ndel;
    }
    if (max_l) *max_l = r.st.max_l;
    return 0;
}

void *orc_factor(int64_t n, const int64_t *col_ptr, const int64_t *row_idx, const double *val,
                 int32_t nsn, const int32_t *sn_begin, const int32_t *sn_end,
                 const int64_t *pat_ptr, const int32_t *pat, const int32_t *parent,
                 const int32_t *perm_fwd, const double *scaling, const orc_options *o) {
    Ctx *c = new Ctx();
    c->A = Matrix{n, col_ptr, row_idx, val};
    c->o = to_opts(o);
    if (!valid_opts(c->o)) { delete c; return nullptr; }
    c->perm_fwd.assign(perm_fwd, perm_fwd + n);
    c->perm_inv.assign(n, 0);
    for (i64 i = 0; i < n; ++i) c->perm_inv[perm_fwd[i]] = (int)i;
    c->scaling = scaling;
    vector<SnInfo> s0 = input_structure(nsn, sn_begin, sn_end, pat_ptr, pat, parent);
    c->sn_before = nsn;
    c->sns = amalgamate(std::move(s0), c->o.nemin);
    c->merges = c->sn_before - (i64)c->sns.size();
    c->children = build_children_map(c->sns);
    c->small_leaves = classify_small_leaf(c->sns, c->children, c->o.small_leaf);
    build_assembly_maps(*c);
    const int ns = (int)c->sns.size();
    c->ff.assign(ns, FrontFactorsO());
    c->st.assign(ns, SnStats());
    c->keep = g_keep_contributions;
    factor_tree_levelset(*c);
    c->contrib.clear();
    c->amap.clear();
    c->ea.clear();
    return c;
}

int orc_solve(void *h, double *rhs, int64_t ldrhs, int32_t nrhs) {
    Ctx *c = (Ctx *)h;
    const i64 n = c->A.n;
    if (n == 0) return 0;
    vector<double> xp((size_t)n * nrhs);
    for (int q = 0; q < nrhs; ++q) {
        for (i64 i = 0; i < n; ++i) xp[i + q * n] = rhs[c->perm_fwd[i] + q * ldrhs];
        if (c->scaling)
            for (i64 i = 0; i < n; ++i) xp[i + q * n] *= c->scaling[i];
    }
    aptp_solve(*c, xp.data(), n, nrhs);
    for (int q = 0; q < nrhs; ++q) {
        if (c->scaling)
            for (i64 i = 0; i < n; ++i) xp[i + q * n] *= c->scaling[i];
        for (i64 i = 0; i < n; ++i) rhs[c->perm_fwd[i] + q * ldrhs] = xp[i + q * n];
    }
    return 0;
}

void orc_inertia(void *h, int64_t *out) {
    Ctx *c = (Ctx *)h;
    out[0] = out[1] = out[2] = 0;
    for (auto &f : c->ff) d_inertia(f.d11, out);
}

void orc_stats(void *h, int64_t *out) {
    Ctx *c = (Ctx *)h;
    i64 t1 = 0, t2 = 0, td = 0, zp = 0, mf = 0;
    for (auto &s : c->st) {
        t1 += s.n1; t2 += s.n2; td += s.k - s.ne;
        mf = std::max<i64>(mf, s.m);
    }
    for (auto &f : c->ff)
        for (int i = 0; i < f.d11.n; ++i)
            if (f.d11.kind[i] == 1 && f.d11.diag[i] == 0.0) zp += 1;
    for (size_t s = 0; s < c->sns.size(); ++s)
        if (c->sns[s].parent < 0 && c->st[s].ne < c->st[s].m) zp += c->st[s].k - c->st[s].ne;
    i64 sln = 0;
    for (auto &sl : c->small_leaves) sln += (i64)sl.nodes.size();
    out[0] = t1; out[1] = t2; out[2] = td; out[3] = zp; out[4] = mf;
    out[5] = c->sn_before; out[6] = (i64)c->sns.size(); out[7] = c->merges;
    out[8] = (i64)c->small_leaves.size(); out[9] = sln;
}

int32_t orc_num_supernodes(void *h) { return (int32_t)((Ctx *)h)->sns.size(); }

void orc_supernode_info(void *h, int32_t s, int64_t *out, double *max_l) {
    Ctx *c = (Ctx *)h;
    const SnStats &st = c->st[s];
    out[0] = st.m; out[1] = st.k; out[2] = st.ne; out[3] = st.k - st.ne;
    out[4] = st.n1; out[5] = st.n2; out[6] = c->sns[s].parent; out[7] = c->sns[s].in_small_leaf;
    if (max_l) *max_l = st.max_l;
}

void orc_supernode_structure(void *h, int32_t s, int32_t *owned, int32_t *pattern, int64_t *out) {
    Ctx *c = (Ctx *)h;
    const SnInfo &sn = c->sns[s];
    int q = 0;
    for (auto &r : sn.owned)
        for (int j = r.lo; j < r.hi; ++j) { if (owned) owned[q] = j; ++q; }
    out[0] = q;
    out[1] = (i64)sn.pattern.size();
    if (pattern)
        for (size_t i = 0; i < sn.pattern.size(); ++i) pattern[i] = sn.pattern[i];
}

void orc_front_factors(void *h, int32_t s, double *l11, double *l21, int8_t *dkind, double *ddiag,
                       double *doff, int32_t *col_indices, int32_t *row_indices,
                       int32_t *local_perm) {
    Ctx *c = (Ctx *)h;
    const FrontFactorsO &f = c->ff[s];
    if (l11) std::memcpy(l11, f.l11.data(), f.l11.size() * sizeof(double));
    if (l21) std::memcpy(l21, f.l21.data(), f.l21.size() * sizeof(double));
    for (int i = 0; i < f.ne; ++i) {
        if (dkind) dkind[i] = f.d11.kind[i];
        if (ddiag) ddiag[i] = f.d11.diag[i];
        if (doff) doff[i] = f.d11.off[i];
        if (col_indices) col_indices[i] = f.col_idx[i];
    }
    if (row_indices)
        for (size_t i = 0; i < f.row_idx.size(); ++i) row_indices[i] = f.row_idx[i];
    if (local_perm)
        for (size_t i = 0; i < f.local_perm.size(); ++i) local_perm[i] = f.local_perm[i];
}

void orc_set_keep_contributions(int32_t on) { g_keep_contributions = on != 0; }

int64_t orc_contribution(void *h, int32_t s, double *data, int32_t *rows, int64_t *num_delayed) {
    Ctx *c = (Ctx *)h;
    if (s < 0 || s >= (int32_t)c->kept.size() || !c->kept[s].present) return -1;
    const Contribution &cb = c->kept[s];
    if (data) std::memcpy(data, cb.data.data(), cb.data.size() * sizeof(double));
    if (rows)
        for (int i = 0; i < cb.size; ++i) rows[i] = cb.rows[i];
    if (num_delayed) *num_delayed = cb.ndelayed;
    return cb.size;
}

void orc_extend_add_mapped(double *parent, int64_t ldp, const double *child, int32_t size,
                           const uint32_t *row_map) {
    extend_add_mapped(parent, ldp, child, size, row_map);
}

void orc_set_threads(int32_t threads) {
#ifdef _OPENMP
    omp_set_num_threads(threads > 0 ? threads : 1);
#else
    (void)threads;
#endif
}

int32_t orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_free(void *h) { delete (Ctx *)h; }

double orc_backward_error(int64_t n, const int64_t *cp, const int64_t *ri, const double *v,
                          const double *x, const double *b) {
    vector<double> ax(n, 0.0);
    for (i64 j = 0; j < n; ++j)
        for (i64 idx = cp[j]; idx < cp[j + 1]; ++idx) ax[ri[idx]] += v[idx] * x[j];
    double rn = 0.0, xn = 0.0, bn = 0.0, an = 0.0;
    for (i64 k = 0; k < n; ++k) { double r = ax[k] - b[k]; rn += r * r; }
    for (i64 k = 0; k < n; ++k) xn += x[k] * x[k];
    for (i64 k = 0; k < n; ++k) bn += b[k] * b[k];
    for (i64 j = 0; j < n; ++j)
        for (i64 idx = cp[j]; idx < cp[j + 1]; ++idx) an += v[idx] * v[idx];
    const double den = std::sqrt(an) * std::sqrt(xn) + std::sqrt(bn);
    if (den == 0.0) return 0.0;
    return std::sqrt(rn) / den;
}

int32_t orc_amalgamate(int32_t nsn, const int32_t *sn_begin, const int32_t *sn_end,
                       const int64_t *pat_ptr, const int32_t *pat, const int32_t *parent,
                       int32_t nemin, int32_t *begin_out, int32_t *end_out, int32_t *parent_out,
                       int64_t *rng_ptr, int32_t *rng_lo, int32_t *rng_hi, int64_t *pat_ptr_out,
                       int32_t *pat_out) {
    vector<SnInfo> s = amalgamate(input_structure(nsn, sn_begin, sn_end, pat_ptr, pat, parent), nemin);
    i64 rp = 0, pp = 0;
    rng_ptr[0] = 0;
    pat_ptr_out[0] = 0;
    for (size_t i = 0; i < s.size(); ++i) {
        begin_out[i] = s[i].col_begin;
        end_out[i] = s[i].col_end;
        parent_out[i] = s[i].parent;
        for (auto &r : s[i].owned) { rng_lo[rp] = r.lo; rng_hi[rp] = r.hi; ++rp; }
        rng_ptr[i + 1] = rp;
        for (int v : s[i].pattern) pat_out[pp++] = v;
        pat_ptr_out[i + 1] = pp;
    }
    return (int32_t)s.size();
}

}  // extern "C"
