// solve.hip -- supernodal forward / diagonal / backward solve on the device-resident factors
// (aptp_solve, sparse/src/symmetric/solve.rs:57-326), for nrhs right-hand sides.
//
// The factors are read in place from the fronts: L11 = F[0:ne, 0:ne] (unit lower, 2x2
// off-diagonals zeroed), L21 = F[ne:m, 0:ne], row/column indices from the solve index.
// Forward: levels bottom-up; a front first pulls the L21*x updates of its descendants into
// its own columns, in postorder (the order the reference subtracts them), then solves with
// L11 and forms its own update vector.  Backward: levels top-down, gather-only.
// Each right-hand side is one blockIdx.y slice.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "rv_numeric.h"

namespace rv {

namespace {

// x_perm[i] = rhs[perm_fwd[i]] (* s[i])  (solver.rs:396-410)
__global__ void k_perm_in(const double *__restrict__ rhs, long long ld, double *__restrict__ xp,
                          long long n, const int *__restrict__ pf, const double *__restrict__ s) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;
    if (i >= n) return;
    double v = rhs[pf[i] + q * ld];
    if (s) v *= s[i];
    xp[i + q * n] = v;
}

// rhs[perm_fwd[i]] = x_perm[i] (* s[i])  (solver.rs:416-425)
__global__ void k_perm_out(double *__restrict__ rhs, long long ld, const double *__restrict__ xp,
                           long long n, const int *__restrict__ pf, const double *__restrict__ s) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;
    if (i >= n) return;
    double v = xp[i + q * n];
    if (s) v *= s[i];
    rhs[pf[i] + q * ld] = v;
}

// Pull the descendants' updates into the eliminated columns of each front of the level.
__global__ void k_fwd_pull(const FrontDev *__restrict__ fr, const int *__restrict__ flist,
                           const int *__restrict__ ptr, const long long *__restrict__ ent,
                           const double *__restrict__ Y, long long ytotal, double *__restrict__ x,
                           long long n) {
    const FrontDev &F = fr[flist[blockIdx.x]];
    const int q = blockIdx.y;
    double *xq = x + q * n;
    const double *Yq = Y + q * ytotal;
    for (int i = threadIdx.x; i < F.ne; i += blockDim.x) {
        const int t = F.sidx[i];
        double v = xq[t];
        for (int e = ptr[t]; e < ptr[t + 1]; ++e) v -= Yq[ent[e]];
        xq[t] = v;
    }
}

// Same pull for columns that no front eliminated (unresolved delays at the roots).
__global__ void k_pull_rows(const int *__restrict__ rows, int nrows, const int *__restrict__ ptr,
                            const long long *__restrict__ ent, const double *__restrict__ Y,
                            long long ytotal, double *__restrict__ x, long long n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;
    if (i >= nrows) return;
    const int t = rows[i];
    double v = x[t + q * n];
    for (int e = ptr[t]; e < ptr[t + 1]; ++e) v -= Y[q * ytotal + ent[e]];
    x[t + q * n] = v;
}

// Forward substitution with L11 (unit lower) for one front, blocks of 64 columns:
// the diagonal block by one wavefront (register broadcast), the rows below by all waves.
// Per row the subtractions run in increasing column order.
__global__ __launch_bounds__(256) void k_fwd_trsv(const FrontDev *__restrict__ fr,
                                                  const int *__restrict__ flist,
                                                  const long long *__restrict__ wbase,
                                                  long long wtotal, double *__restrict__ W,
                                                  double *__restrict__ x, long long n) {
    const int li = blockIdx.x;
    const FrontDev &F = fr[flist[li]];
    const int q = blockIdx.y;
    const int ne = F.ne;
    const size_t lda = F.m;
    const double *L = F.F;
    double *w = W + q * wtotal + wbase[li];
    double *xq = x + q * n;
    const int tid = threadIdx.x;
    for (int i = tid; i < ne; i += 256) w[i] = xq[F.sidx[i]];
    __syncthreads();
    for (int jb = 0; jb < ne; jb += 64) {
        const int je = min(jb + 64, ne);
        if (tid < 64) {
            const int i = jb + tid;
            double wi = (i < je) ? w[i] : 0.0;
            for (int j = jb; j < je; ++j) {
                const double wj = __shfl(wi, j - jb);
                if (i > j && i < je) wi -= L[i + j * lda] * wj;
            }
            if (i < je) w[i] = wi;
        }
        __syncthreads();
        for (int i = je + tid; i < ne; i += 256) {
            double wi = w[i];
            for (int j = jb; j < je; ++j) wi -= L[i + j * lda] * w[j];
            w[i] = wi;
        }
        __syncthreads();
    }
    for (int i = tid; i < ne; i += 256) xq[F.sidx[i]] = w[i];
}

// y = L21 * w for a chunk of 256 rows of L21 (sum in increasing column order).
__global__ __launch_bounds__(256) void k_fwd_gemv(const FrontDev *__restrict__ fr,
                                                  const int *__restrict__ flist,
                                                  const int2 *__restrict__ items,
                                                  const long long *__restrict__ wbase,
                                                  const long long *__restrict__ ybase,
                                                  long long wtotal, long long ytotal,
                                                  const double *__restrict__ W,
                                                  double *__restrict__ Y) {
    const int2 it = items[blockIdx.x];
    const int li = it.x;
    const FrontDev &F = fr[flist[li]];
    const int q = blockIdx.y;
    const int ne = F.ne, r = F.m - ne;
    const int i = it.y * 256 + threadIdx.x;
    if (i >= r) return;
    const size_t lda = F.m;
    const double *L = F.F + ne + i;
    const double *w = W + q * wtotal + wbase[li];
    double acc = 0.0;
    for (int j = 0; j < ne; ++j) acc += L[j * lda] * w[j];
    Y[q * ytotal + ybase[li] + i] = acc;
}

// D solve (MixedDiagonal::solve_in_place, diagonal.rs:312-365).
__global__ void k_diag(const FrontDev *__restrict__ fr, const int *__restrict__ flist,
                       double *__restrict__ x, long long n) {
    const FrontDev &F = fr[flist[blockIdx.x]];
    const int q = blockIdx.y;
    double *xq = x + q * n;
    for (int c = threadIdx.x; c < F.ne; c += blockDim.x) {
        const int kd = F.dkind[c];
        if (kd == 1) {
            const double d = F.ddiag[c];
            const int t = F.sidx[c];
            xq[t] = (d == 0.0) ? 0.0 : xq[t] / d;
        } else if (kd == 2) {
            const double a = F.ddiag[c], b = F.doff[c], cc = F.ddiag[c + 1];
            const double det = a * cc - b * b;
            const int t1 = F.sidx[c], t2 = F.sidx[c + 1];
            if (det == 0.0) {
                xq[t1] = 0.0;
                xq[t2] = 0.0;
            } else {
                const double r1 = xq[t1], r2 = xq[t2];
                xq[t1] = (cc * r1 - b * r2) / det;
                xq[t2] = (a * r2 - b * r1) / det;
            }
        }
    }
}

// t_j = L21[:, j] . x[rows]: one wavefront per column.
__global__ __launch_bounds__(256) void k_bwd_gemv(const FrontDev *__restrict__ fr,
                                                  const int *__restrict__ flist,
                                                  const int2 *__restrict__ items,
                                                  const long long *__restrict__ wbase,
                                                  long long wtotal, double *__restrict__ W,
                                                  const double *__restrict__ x, long long n) {
    const int2 it = items[blockIdx.x];
    const int li = it.x;
    const FrontDev &F = fr[flist[li]];
    const int q = blockIdx.y;
    const int ne = F.ne, r = F.m - ne;
    const int lane = threadIdx.x & 63;
    const int j = it.y * 4 + (threadIdx.x >> 6);
    if (j >= ne) return;
    const size_t lda = F.m;
    const double *L = F.F + ne + j * lda;
    const int *ri = F.sidx + ne;
    const double *xq = x + q * n;
    double acc = 0.0;
    for (int i = lane; i < r; i += 64) acc += L[i] * xq[ri[i]];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if (lane == 0) W[q * wtotal + wbase[li] + j] = acc;
}

// Backward substitution with L11^T (unit upper) for one front, blocks of 64 from the end.
__global__ __launch_bounds__(256) void k_bwd_trsv(const FrontDev *__restrict__ fr,
                                                  const int *__restrict__ flist,
                                                  const long long *__restrict__ wbase,
                                                  long long wtotal, double *__restrict__ W,
                                                  double *__restrict__ x, long long n) {
    const int li = blockIdx.x;
    const FrontDev &F = fr[flist[li]];
    const int q = blockIdx.y;
    const int ne = F.ne;
    const size_t lda = F.m;
    const double *L = F.F;
    double *w = W + q * wtotal + wbase[li];
    double *xq = x + q * n;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int i = tid; i < ne; i += 256) w[i] = xq[F.sidx[i]] - w[i];
    __syncthreads();
    const int nblk = (ne + 63) / 64;
    for (int b = nblk - 1; b >= 0; --b) {
        const int jb = b * 64, je = min(jb + 64, ne);
        // contributions of the already solved rows below the block
        for (int j = jb + wv; j < je; j += 4) {
            double acc = 0.0;
            for (int i = je + lane; i < ne; i += 64) acc += L[i + j * lda] * w[i];
            for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
            if (lane == 0) w[j] -= acc;
        }
        __syncthreads();
        if (wv == 0) {
            const int i = jb + lane;
            double wi = (i < je) ? w[i] : 0.0;
            for (int j = je - 1; j >= jb; --j) {
                double c = (i > j && i < je) ? L[i + j * lda] * wi : 0.0;
                for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
                if (i == j) wi -= c;
            }
            if (i < je) w[i] = wi;
        }
        __syncthreads();
    }
    for (int i = tid; i < ne; i += 256) xq[F.sidx[i]] = w[i];
}

}  // namespace

void NumericGPU::prepare_solve() {
    const Prepared &P = *prep_;
    const int ns = (int)P.sns.size();
    const i64 n = P.n;
    // per-front bases
    std::vector<long long> wb(ns, 0), yb(ns, 0);
    long long wt = 0, yt = 0;
    for (int s = 0; s < ns; ++s) {
        wb[s] = wt;
        wt += ne_[s];
        yb[s] = yt;
        if (ne_[s] > 0) yt += m_[s] - ne_[s];
    }
    // pull lists: for each global index, the Y entries of the fronts (postorder) that update it
    std::vector<int> cnt(n + 1, 0);
    for (int s = 0; s < ns; ++s) {
        if (ne_[s] == 0) continue;
        for (int g : delayed_[s]) cnt[g + 1]++;
        for (int g : P.sns[s].pattern) cnt[g + 1]++;
    }
    for (i64 i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    std::vector<long long> ent(cnt[n]);
    {
        std::vector<int> pos(cnt.begin(), cnt.end() - 1);
        for (int s = 0; s < ns; ++s) {
            if (ne_[s] == 0) continue;
            long long p = yb[s];
            for (int g : delayed_[s]) ent[pos[g]++] = p++;
            for (int g : P.sns[s].pattern) ent[pos[g]++] = p++;
        }
    }
    // unresolved columns (delays left at the roots)
    std::vector<int> unres;
    for (int s = 0; s < ns; ++s)
        if (P.sns[s].parent < 0) unres.insert(unres.end(), delayed_[s].begin(), delayed_[s].end());
    // per level front lists and items
    const int nl = (int)P.levels.size();
    std::vector<int> flist;
    std::vector<long long> lwb, lyb;
    std::vector<int2> gf, gb;
    fwd_off_.assign(nl + 1, 0);
    fwd_cnt_.assign(nl, 0);
    bwd_off_.assign(nl + 1, 0);
    bwd_cnt_.assign(nl, 0);
    scat_off_.assign(nl + 1, 0);
    std::vector<size_t> gfo(nl + 1, 0), gbo(nl + 1, 0);
    for (int lv = 0; lv < nl; ++lv) {
        fwd_off_[lv] = flist.size();
        gfo[lv] = gf.size();
        gbo[lv] = gb.size();
        for (int s : P.levels[lv]) {
            if (ne_[s] == 0) continue;
            const int li = (int)(flist.size() - fwd_off_[lv]);
            flist.push_back(s);
            lwb.push_back(wb[s]);
            lyb.push_back(yb[s]);
            const int r = m_[s] - ne_[s];
            for (int c = 0; c < (r + 255) / 256; ++c) gf.push_back(make_int2(li, c));
            for (int c = 0; c < (ne_[s] + 3) / 4; ++c) gb.push_back(make_int2(li, c));
        }
        fwd_cnt_[lv] = (int)(flist.size() - fwd_off_[lv]);
    }
    fwd_off_[nl] = flist.size();
    gfo[nl] = gf.size();
    gbo[nl] = gb.size();
    // device metadata
    std::vector<char> &st = sol_meta_host_;
    st.clear();
    auto push = [&](const void *p, size_t bytes) {
        size_t off = (st.size() + 255) / 256 * 256;
        st.resize(off + bytes);
        if (bytes) std::memcpy(st.data() + off, p, bytes);
        return off;
    };
    const size_t o_ptr = push(cnt.data(), cnt.size() * sizeof(int));
    const size_t o_ent = push(ent.data(), ent.size() * sizeof(long long));
    const size_t o_fl = push(flist.data(), flist.size() * sizeof(int));
    const size_t o_wb = push(lwb.data(), lwb.size() * sizeof(long long));
    const size_t o_yb = push(lyb.data(), lyb.size() * sizeof(long long));
    const size_t o_gf = push(gf.data(), gf.size() * sizeof(int2));
    const size_t o_gb = push(gb.data(), gb.size() * sizeof(int2));
    const size_t o_un = push(unres.data(), unres.size() * sizeof(int));
    if (st.size() > sol_meta_bytes_) {
        if (d_sol_meta_) (void)hipFree(d_sol_meta_);
        RV_HIP(hipMalloc(&d_sol_meta_, st.size()));
        sol_meta_bytes_ = st.size();
    }
    RV_HIP(hipMemcpy(d_sol_meta_, st.data(), st.size(), hipMemcpyHostToDevice));
    char *b = (char *)d_sol_meta_;
    sm_.ptr = (const int *)(b + o_ptr);
    sm_.ent = (const long long *)(b + o_ent);
    sm_.flist = (const int *)(b + o_fl);
    sm_.wbase = (const long long *)(b + o_wb);
    sm_.ybase = (const long long *)(b + o_yb);
    sm_.gf = (const int2 *)(b + o_gf);
    sm_.gb = (const int2 *)(b + o_gb);
    sm_.unres = (const int *)(b + o_un);
    sm_.nunres = (int)unres.size();
    sm_.wtotal = std::max<long long>(wt, 1);
    sm_.ytotal = std::max<long long>(yt, 1);
    sm_.gfo = gfo;
    sm_.gbo = gbo;
    solve_ready_ = true;
}

void NumericGPU::solve(double *d_rhs, i64 ldrhs, int nrhs, const double *d_scaling) {
    if (!solve_ready_) prepare_solve();
    const Prepared &P = *prep_;
    const i64 n = P.n;
    if (n == 0 || nrhs <= 0) return;
    double *xp = (double *)sol_buf_.get((size_t)n * nrhs * sizeof(double));
    double *W = (double *)ybuf_.get(((size_t)sm_.wtotal + (size_t)sm_.ytotal) * nrhs * sizeof(double));
    double *Y = W + (size_t)sm_.wtotal * nrhs;
    const dim3 gp((unsigned)((n + 255) / 256), nrhs);
    hipLaunchKernelGGL(k_perm_in, gp, dim3(256), 0, stream_, d_rhs, (long long)ldrhs, xp, (long long)n,
                       P.d_perm_fwd, d_scaling);
    const int nl = (int)P.levels.size();
    // forward (solve.rs:89-92)
    for (int lv = 0; lv < nl; ++lv) {
        const int c = fwd_cnt_[lv];
        if (c == 0) continue;
        const int *fl = sm_.flist + fwd_off_[lv];
        const long long *wb = sm_.wbase + fwd_off_[lv];
        const long long *yb = sm_.ybase + fwd_off_[lv];
        hipLaunchKernelGGL(k_fwd_pull, dim3(c, nrhs), dim3(256), 0, stream_, d_fr_, fl, sm_.ptr, sm_.ent,
                           (const double *)Y, sm_.ytotal, xp, (long long)n);
        hipLaunchKernelGGL(k_fwd_trsv, dim3(c, nrhs), dim3(256), 0, stream_, d_fr_, fl, wb, sm_.wtotal,
                           W, xp, (long long)n);
        const size_t ng = sm_.gfo[lv + 1] - sm_.gfo[lv];
        if (ng)
            hipLaunchKernelGGL(k_fwd_gemv, dim3((unsigned)ng, nrhs), dim3(256), 0, stream_, d_fr_, fl,
                               sm_.gf + sm_.gfo[lv], wb, yb, sm_.wtotal, sm_.ytotal, (const double *)W, Y);
    }
    if (sm_.nunres > 0)
        hipLaunchKernelGGL(k_pull_rows, dim3((sm_.nunres + 255) / 256, nrhs), dim3(256), 0, stream_,
                           sm_.unres, sm_.nunres, sm_.ptr, sm_.ent, (const double *)Y, sm_.ytotal, xp,
                           (long long)n);
    // diagonal (solve.rs:97-135)
    for (int lv = 0; lv < nl; ++lv) {
        const int c = fwd_cnt_[lv];
        if (c == 0) continue;
        hipLaunchKernelGGL(k_diag, dim3(c, nrhs), dim3(64), 0, stream_, d_fr_, sm_.flist + fwd_off_[lv],
                           xp, (long long)n);
    }
    // backward (solve.rs:139-141)
    for (int lv = nl - 1; lv >= 0; --lv) {
        const int c = fwd_cnt_[lv];
        if (c == 0) continue;
        const int *fl = sm_.flist + fwd_off_[lv];
        const long long *wb = sm_.wbase + fwd_off_[lv];
        const size_t ng = sm_.gbo[lv + 1] - sm_.gbo[lv];
        if (ng)
            hipLaunchKernelGGL(k_bwd_gemv, dim3((unsigned)ng, nrhs), dim3(256), 0, stream_, d_fr_, fl,
                               sm_.gb + sm_.gbo[lv], wb, sm_.wtotal, W, (const double *)xp, (long long)n);
        hipLaunchKernelGGL(k_bwd_trsv, dim3(c, nrhs), dim3(256), 0, stream_, d_fr_, fl, wb, sm_.wtotal, W,
                           xp, (long long)n);
    }
    hipLaunchKernelGGL(k_perm_out, gp, dim3(256), 0, stream_, d_rhs, (long long)ldrhs, (const double *)xp,
                       (long long)n, P.d_perm_fwd, d_scaling);
}

}  // namespace rv
