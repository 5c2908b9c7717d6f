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

constexpr int SB = 128;  // diagonal block of the blocked triangular solves (two 64-column halves)
constexpr int GW = 256;  // columns of L21 per forward-GEMV work item

__device__ __forceinline__ double readlane_f64(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// sum over the wavefront (all 64 lanes active): DPP within rows of 16, then four readlanes
__device__ __forceinline__ double wave_sum(double v) {
    v = v + dpp_f64<0xB1>(v);   // quad_perm [1,0,3,2]
    v = v + dpp_f64<0x4E>(v);   // quad_perm [2,3,0,1]
    v = v + dpp_f64<0x141>(v);  // row_half_mirror
    v = v + dpp_f64<0x140>(v);  // row_mirror
    return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}

// The diagonal block [jb, jb + SB) of L11 (unit lower) as two 64 x 64 triangles A (rows and
// columns jb..jb+63) and B (jb+64..jb+127) and the tile C = L[jb+64 : jb+128, jb : jb+64] between
// them: thread (row i = tid & 63, column quarter jq = tid >> 6) loads entries (i, jq + 4u) of all
// three in one round trip (lanes over rows: coalesced); zeros outside the strictly lower parts.
struct DiagBlock {
    double a[16], b[16], c[16];
};

__device__ __forceinline__ void load_diag_block(DiagBlock &d, const gdouble *L, size_t lda, int jb, int n1,
                                                int n2) {
    const int i = threadIdx.x & 63, jq = threadIdx.x >> 6;
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int j = jq + 4 * u;
        d.a[u] = (i < n1 && j < i) ? L[(jb + i) + (size_t)(jb + j) * lda] : 0.0;
        d.b[u] = (i < n2 && j < i) ? L[(jb + 64 + i) + (size_t)(jb + 64 + j) * lda] : 0.0;
        d.c[u] = (i < n2) ? L[(jb + 64 + i) + (size_t)(jb + j) * lda] : 0.0;
    }
}

// Forward substitution with a 64 x 64 unit lower triangle staged as Ls[r][c] = L[r][c] (zero on
// and above the diagonal), by wavefront 0: lane i keeps w_i; x_j reaches the lanes by readlane,
// terms in ascending j.
__device__ __forceinline__ void tri_fwd_wave(const double (*Ls)[65], double *ws, int nn) {
    const int i = threadIdx.x;
    double wi = (i < nn) ? ws[i] : 0.0;
#pragma unroll
    for (int j = 0; j < 63; ++j) {
        const double wj = readlane_f64(wi, j);
        if (i > j) wi -= Ls[i][j] * wj;
    }
    if (i < nn) ws[i] = wi;
}

// Backward substitution with the transpose of such a triangle, by wavefront 0 (column axpy
// form: x_j = w_j, then w_i -= L[j][i] x_j for i < j, j descending).
__device__ __forceinline__ void tri_bwd_wave(const double (*Ls)[65], double *ws, int nn) {
    const int i = threadIdx.x;
    double wi = (i < nn) ? ws[i] : 0.0;
#pragma unroll
    for (int j = 63; j > 0; --j) {
        const double wj = readlane_f64(wi, j);
        if (i < j) wi -= Ls[j][i] * wj;
    }
    if (i < nn) ws[i] = wi;
}

// Pull the descendants' updates into the eliminated columns of a front (postorder within each
// list) and start the front's working vector: w[i] = x[t] - sum of updates, t = sidx[i].
// The list is summed in order, its loads issued four at a time.
__device__ __forceinline__ double pull_column(const FrontDev &F, int i, const int *__restrict__ ptr,
                                              const long long *__restrict__ ent, const double *__restrict__ Yq,
                                              const double *__restrict__ xq) {
    const int t = gbl(F.sidx)[i];
    const int e1 = ptr[t + 1];
    int e = ptr[t];
    double acc = 0.0;
    for (; e + 4 <= e1; e += 4) {
        const long long a0 = ent[e], a1 = ent[e + 1], a2 = ent[e + 2], a3 = ent[e + 3];
        const double y0 = Yq[a0], y1 = Yq[a1], y2 = Yq[a2], y3 = Yq[a3];
        acc = (((acc + y0) + y1) + y2) + y3;
    }
    for (; e < e1; ++e) acc += Yq[ent[e]];
    return xq[t] - acc;
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

// Forward substitution of one diagonal block (all threads of the workgroup): ws holds the block's
// w on entry and its x on return; wavefront 0 solves with A, all threads apply C to the second
// half (column quarters combined in a fixed order), wavefront 0 solves with B.
__device__ __forceinline__ void fwd_block(const DiagBlock &d, double (*Ls)[65], double *ws, double (*part)[64],
                                          int n1, int n2) {
    const int tid = threadIdx.x, i = tid & 63, jq = tid >> 6;
#pragma unroll
    for (int u = 0; u < 16; ++u) Ls[i][jq + 4 * u] = d.a[u];
    __syncthreads();
    if (tid < 64) tri_fwd_wave(Ls, ws, n1);
    __syncthreads();
    if (n2 > 0) {
        double s = 0.0;
#pragma unroll
        for (int u = 0; u < 16; ++u) s += d.c[u] * ws[jq + 4 * u];
        part[jq][i] = s;
#pragma unroll
        for (int u = 0; u < 16; ++u) Ls[i][jq + 4 * u] = d.b[u];
        __syncthreads();
        if (tid < 64) {
            if (i < n2) ws[64 + i] -= ((part[0][i] + part[1][i]) + part[2][i]) + part[3][i];
            tri_fwd_wave(Ls, ws + 64, n2);
        }
        __syncthreads();
    }
}

// Start of a level's forward sweep, one launch: workgroups [0, npull) pull the descendants'
// updates into the front columns beyond the first diagonal block (one thread per column);
// workgroup npull + li pulls the first SB columns of front li itself and solves them as the first
// diagonal block of L11 (the later blocks are solved by k_fwd_rows).  The block's values are
// final: they go to w and to x.
__global__ __launch_bounds__(256) void k_fwd_first(const FrontDev *__restrict__ fr,
                                                   const int *__restrict__ flist,
                                                   const int2 *__restrict__ items, int npull,
                                                   const int *__restrict__ ptr,
                                                   const long long *__restrict__ ent,
                                                   const double *__restrict__ Y, long long ytotal,
                                                   const long long *__restrict__ wbase,
                                                   long long wtotal, double *__restrict__ W,
                                                   double *__restrict__ x, long long n) {
    const int q = blockIdx.y;
    const double *Yq = Y + q * ytotal;
    const double *xq = x + q * n;
    if ((int)blockIdx.x < npull) {
        const int2 it = items[blockIdx.x];
        const FrontDev &F = fr[flist[it.x]];
        const int i = it.y * 256 + threadIdx.x;
        if (i < SB || i >= F.ne) return;
        W[q * wtotal + wbase[it.x] + i] = pull_column(F, i, ptr, ent, Yq, xq);
        return;
    }
    const int li = blockIdx.x - npull;
    const FrontDev &F = fr[flist[li]];
    const int ne = F.ne;
    if (ne <= 0) return;
    const int nb = min(SB, ne), n1 = min(64, nb), n2 = nb - n1;
    const gint *sidx = gbl(F.sidx);
    double *w = W + q * wtotal + wbase[li];
    __shared__ double Ls[64][65];
    __shared__ double ws[SB];
    __shared__ double part[4][64];
    const int tid = threadIdx.x;
    DiagBlock d;
    load_diag_block(d, gbl(F.F), (size_t)F.m, 0, n1, n2);
    const int st = (tid < nb) ? sidx[tid] : 0;
    if (tid < nb) ws[tid] = pull_column(F, tid, ptr, ent, Yq, xq);
    fwd_block(d, Ls, ws, part, n1, n2);
    if (tid < nb) {
        const double v = ws[tid];
        w[tid] = v;
        x[st + q * n] = v;
    }
}

// Rows below the block [jb, jb + SB): w[i] -= sum_{j in block} L[i,j] x_j.  Work item of 128 rows,
// two threads per row (one per 64 columns, all 64 loads in flight together); the row's two
// partial sums are subtracted in order.  The first work item holds the next diagonal block's rows
// once they are updated, so it solves that block as well (its loads issued first, solved as by
// k_fwd_first): one launch per block.
__global__ __launch_bounds__(256) void k_fwd_rows(const FrontDev *__restrict__ fr,
                                                  const int *__restrict__ flist,
                                                  const int2 *__restrict__ items, int jb,
                                                  const long long *__restrict__ wbase,
                                                  long long wtotal, double *__restrict__ W,
                                                  double *__restrict__ x, long long n) {
    const int2 it = items[blockIdx.x];
    const FrontDev &F = fr[flist[it.x]];
    const int ne = F.ne;
    const int je = jb + SB, r0 = je + it.y * 128;
    if (je >= ne || r0 >= ne) return;
    const int q = blockIdx.y;
    const size_t lda = F.m;
    double *w = W + q * wtotal + wbase[it.x];
    __shared__ double ws[SB], wn[SB];
    __shared__ double Ls[64][65];
    __shared__ double part[4][64];
    __shared__ double ps[2][128];
    const int tid = threadIdx.x, ri = tid & 127, ch = tid >> 7;
    const bool next = it.y == 0;  // workgroup-uniform
    const int nb = min(SB, ne - je), n1 = min(64, nb), n2 = nb - n1;
    const bool rin = r0 + ri < ne;
    DiagBlock d;
    int st = 0;
    if (next) {
        load_diag_block(d, gbl(F.F), lda, je, n1, n2);
        if (tid < nb) st = gbl(F.sidx)[je + tid];
    }
    double v[64];
    {
        const gdouble *L = gbl(F.F) + (r0 + ri) + (size_t)(jb + 64 * ch) * lda;
#pragma unroll
        for (int k = 0; k < 64; ++k) v[k] = rin ? L[(size_t)k * lda] : 0.0;
    }
    const double w0 = (tid < 128 && rin) ? w[r0 + ri] : 0.0;
    if (tid < SB) ws[tid] = w[jb + tid];
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 64; ++k) s += v[k] * ws[64 * ch + k];
    ps[ch][ri] = s;
    __syncthreads();
    if (tid < 128 && rin) {
        const double wi = (w0 - ps[0][tid]) - ps[1][tid];
        if (next) wn[tid] = wi;
        else w[r0 + tid] = wi;
    }
    if (!next) return;
    fwd_block(d, Ls, wn, part, n1, n2);
    if (tid < nb) {
        const double v2 = wn[tid];
        w[je + tid] = v2;
        x[st + q * n] = v2;
    }
}

// y = L21 * w.  Work item (front, 64 rows, GW columns): lanes take rows, each wavefront a
// quarter of the columns; a front with more than GW columns leaves one partial sum per column
// chunk in P, added up in chunk order by k_fwd_gemv_sum.
__global__ __launch_bounds__(256) void k_fwd_gemv(const FrontDev *__restrict__ fr,
                                                  const int *__restrict__ flist,
                                                  const int4 *__restrict__ items,
                                                  const long long *__restrict__ wbase,
                                                  const long long *__restrict__ ybase,
                                                  const long long *__restrict__ pbase,
                                                  long long wtotal, long long ytotal,
                                                  long long ptotal, const double *__restrict__ W,
                                                  double *__restrict__ Y, double *__restrict__ P) {
    const int4 it = items[blockIdx.x];
    const int li = it.x;
    const FrontDev &F = fr[flist[li]];
    const int q = blockIdx.y;
    const int ne = F.ne, r = F.m - ne;
    const int c0 = it.z * GW, c1 = min(c0 + GW, ne);
    const size_t lda = F.m;
    __shared__ double ws[GW];
    __shared__ double part[4][64];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const double *w = W + q * wtotal + wbase[li];
    if (c0 + tid < c1) ws[tid] = w[c0 + tid];
    __syncthreads();
    const int i = it.y * 64 + lane;
    double acc = 0.0;
    if (i < r) {
        const gdouble *L = gbl(F.F) + ne + i;
        const int jb = c0 + wv * (GW / 4), je = min(jb + GW / 4, c1);
        for (int j = jb; j < je; j += 32) {
            double lv[32];
#pragma unroll
            for (int u = 0; u < 32; ++u) lv[u] = (j + u < je) ? L[(size_t)(j + u) * lda] : 0.0;
#pragma unroll
            for (int u = 0; u < 32; ++u)
                if (j + u < je) acc += lv[u] * ws[j + u - c0];
        }
    }
    part[wv][lane] = acc;
    __syncthreads();
    if (wv == 0 && i < r) {
        const double sum = ((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane];
        if (ne <= GW) Y[q * ytotal + ybase[li] + i] = sum;
        else P[q * ptotal + pbase[li] + (size_t)it.z * r + i] = sum;
    }
}

// Y = sum of the column-chunk partials of k_fwd_gemv, in chunk order (fronts with ne > GW).
__global__ __launch_bounds__(256) void k_fwd_gemv_sum(const FrontDev *__restrict__ fr,
                                                      const int *__restrict__ flist,
                                                      const int2 *__restrict__ items,
                                                      const long long *__restrict__ ybase,
                                                      const long long *__restrict__ pbase,
                                                      long long ytotal, long long ptotal,
                                                      const double *__restrict__ P,
                                                      double *__restrict__ Y) {
    const int2 it = items[blockIdx.x];
    const FrontDev &F = fr[flist[it.x]];
    const int q = blockIdx.y;
    const int ne = F.ne, r = F.m - ne;
    const int i = it.y * 256 + threadIdx.x;
    if (i >= r) return;
    const int nc = (ne + GW - 1) / GW;
    const double *p = P + q * ptotal + pbase[it.x] + i;
    double s = 0.0;
    for (int c = 0; c < nc; ++c) s += p[(size_t)c * r];
    Y[q * ytotal + ybase[it.x] + i] = s;
}

// D solve (MixedDiagonal::solve_in_place, diagonal.rs:312-365).
// Work item (front, chunk of 256 columns); a 2x2 pivot is taken by the thread of its first column.
__global__ __launch_bounds__(256) void k_diag(const FrontDev *__restrict__ fr, const int *__restrict__ flist,
                                              const int2 *__restrict__ items, double *__restrict__ x,
                                              long long n) {
    const int2 it = items[blockIdx.x];
    const FrontDev &F = fr[flist[it.x]];
    const int q = blockIdx.y;
    double *xq = x + q * n;
    {
        const int c = it.y * 256 + threadIdx.x;
        if (c >= F.ne) return;
        const int kd = gbl(F.dkind)[c];
        if (kd == 1) {
            const double d = gbl(F.ddiag)[c];
            const int t = gbl(F.sidx)[c];
            xq[t] = (d == 0.0) ? 0.0 : xq[t] / d;
        } else if (kd == 2) {
            const double a = gbl(F.ddiag)[c], b = gbl(F.doff)[c], cc = gbl(F.ddiag)[c + 1];
            const double det = a * cc - b * b;
            const int t1 = gbl(F.sidx)[c], t2 = gbl(F.sidx)[c + 1];
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

// Start of a front's backward sweep: w[j] = x[sidx[j]] - t_j, t_j = L21[:, j] . x[rows], one
// wavefront per column (the work items cover every column of every front of the level, those
// without L21 rows too).
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
    const gdouble *L = gbl(F.F) + ne + j * lda;
    const gint *ri = gbl(F.sidx) + ne;
    const double *xq = x + q * n;
    double acc = 0.0;
    // eight rows per lane in flight: the index and factor loads, then the gathers
    for (int i0 = lane; i0 < r; i0 += 512) {
        int idx[8];
        double lv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int i = i0 + 64 * u;
            idx[u] = (i < r) ? ri[i] : -1;
            lv[u] = (i < r) ? L[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += lv[u] * ((idx[u] >= 0) ? xq[idx[u]] : 0.0);
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if (lane == 0) W[q * wtotal + wbase[li] + j] = xq[gbl(F.sidx)[j]] - acc;
}

// Backward substitution of one diagonal block with L11^T (all threads of the workgroup): ws holds
// the block's w, from which every later block has been removed, on entry and its x on return.
// Wavefront 0 solves with B^T, all threads form C^T x_B (products through LDS, summed over rows
// in ascending order), wavefront 0 solves with A^T.
__device__ __forceinline__ void bwd_block(const DiagBlock &d, double (*Ls)[65], double *ws, int n1, int n2) {
    const int tid = threadIdx.x, i = tid & 63, jq = tid >> 6;
    if (n2 > 0) {
#pragma unroll
        for (int u = 0; u < 16; ++u) Ls[i][jq + 4 * u] = d.b[u];
        __syncthreads();
        if (tid < 64) tri_bwd_wave(Ls, ws + 64, n2);
        __syncthreads();
        const double xb = (i < n2) ? ws[64 + i] : 0.0;
#pragma unroll
        for (int u = 0; u < 16; ++u) Ls[jq + 4 * u][i] = d.c[u] * xb;
        __syncthreads();
        if (tid < 64) {
            double s = 0.0;
#pragma unroll
            for (int r = 0; r < 64; ++r) s += Ls[tid][r];
            ws[tid] -= s;
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) Ls[i][jq + 4 * u] = d.a[u];
    __syncthreads();
    if (tid < 64) tri_bwd_wave(Ls, ws, n1);
    __syncthreads();
}

// Backward step b of a front with L11^T, right-looking: work item (front, column tile t) with
// t <= b.  When block b + 1 exists, its x (final) is applied to the tile's SB columns:
// w[j] -= sum_{i in block b+1} L[i][j] x_i (lanes over rows, column sums through LDS).  The tile t = b is then final and is solved as a
// diagonal block (loaded at the start); its values go to w and to x.  One launch per block.
// UPD = false: a level whose fronts are single blocks (diagonal blocks only).
template <bool UPD>
__global__ __launch_bounds__(256) void k_bwd_step(const FrontDev *__restrict__ fr,
                                                  const int *__restrict__ flist,
                                                  const int2 *__restrict__ items, int b,
                                                  const long long *__restrict__ wbase,
                                                  long long wtotal, double *__restrict__ W,
                                                  double *__restrict__ x, long long n) {
    const int2 it = items[blockIdx.x];
    const FrontDev &F = fr[flist[it.x]];
    const int ne = F.ne;
    const int nblk = (ne + SB - 1) / SB, t = it.y;
    if (b >= nblk || t > b) return;
    const bool upd = UPD && b + 1 < nblk, diag = t == b;  // workgroup-uniform
    if (!upd && !diag) return;
    const int q = blockIdx.y;
    const size_t lda = F.m;
    const gdouble *L = gbl(F.F);
    double *w = W + q * wtotal + wbase[it.x];
    __shared__ double Ls[64][65];
    __shared__ double ws[SB];
    __shared__ double pbuf[32 * 129];
    __shared__ double part8[8][32];
    __shared__ double colsum[SB];
    const int tid = threadIdx.x;
    const int jb = b * SB, nb = min(SB, ne - jb), n1 = min(64, nb), n2 = nb - n1;
    DiagBlock d;
    int st = 0;
    if (diag && !upd) {
        load_diag_block(d, L, lda, jb, n1, n2);
        if (tid < nb) st = gbl(F.sidx)[jb + tid];
    }
    const int c0 = t * SB;
    if (upd) {
        // rows r0 + ri of block b + 1, columns c0 + 64 ch + k (the tile is full: t <= b < nblk - 1);
        // all 64 loads in flight together, then the products' column sums through LDS, 32
        // columns at a time: partial sums over 8 groups of 16 rows, added in order
        const int r0 = (b + 1) * SB, nr = min(SB, ne - r0);
        const int ri = tid & 127, ch = tid >> 7;
        const bool rin = ri < nr;
        double v[64];
        {
            const gdouble *Lr = L + (r0 + ri) + (size_t)(c0 + 64 * ch) * lda;
#pragma unroll
            for (int k = 0; k < 64; ++k) v[k] = rin ? Lr[(size_t)k * lda] : 0.0;
        }
        const double xi = rin ? w[r0 + ri] : 0.0;
        const double wc = (tid < SB) ? w[c0 + tid] : 0.0;
        double (*prod)[129] = (double (*)[129])pbuf;
        const int cc = tid & 31, pp = tid >> 5;
#pragma unroll
        for (int rd = 0; rd < 4; ++rd) {
            if (rd == 2 && diag) {
                // the diagonal block's loads, once half of the tile's registers are free
                load_diag_block(d, L, lda, jb, n1, n2);
                if (tid < nb) st = gbl(F.sidx)[jb + tid];
            }
            // columns 16 rd + k of each half: prod[16 ch + k][ri]
#pragma unroll
            for (int k = 0; k < 16; ++k) prod[16 * ch + k][ri] = v[16 * rd + k] * xi;
            __syncthreads();
            double s = 0.0;
#pragma unroll
            for (int r = 0; r < 16; ++r) s += prod[cc][16 * pp + r];
            part8[pp][cc] = s;
            __syncthreads();
            if (tid < 32) {
                double c = 0.0;
#pragma unroll
                for (int g = 0; g < 8; ++g) c += part8[g][tid];
                // prod row 16 ch + k <-> column 64 ch + 16 rd + k
                colsum[64 * (tid >> 4) + 16 * rd + (tid & 15)] = c;
            }
        }
        __syncthreads();
        if (tid < SB) {
            const double nv = wc - colsum[tid];
            if (diag) ws[tid] = nv;
            else w[c0 + tid] = nv;
        }
    } else if (tid < nb) {
        ws[tid] = w[jb + tid];
    }
    if (!diag) return;
    bwd_block(d, Ls, ws, n1, n2);
    if (tid < nb) {
        const double v = ws[tid];
        w[jb + tid] = v;
        x[st + q * n] = v;
    }
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
    // per level front lists and work items
    const int nl = (int)P.levels.size();
    std::vector<int> flist;
    std::vector<long long> lwb, lyb, lpb;
    std::vector<int4> gf;
    std::vector<int2> gb, gp, gr, gc, gs, gd;
    long long pt = 0;
    fwd_off_.assign(nl + 1, 0);
    fwd_cnt_.assign(nl, 0);
    sm_.gfo.assign(nl + 1, 0);
    sm_.gbo.assign(nl + 1, 0);
    sm_.gpo.assign(nl + 1, 0);
    sm_.gro.assign(nl + 1, 0);
    sm_.gco.assign(nl + 1, 0);
    sm_.gso.assign(nl + 1, 0);
    sm_.maxne.assign(nl, 0);
    for (int lv = 0; lv < nl; ++lv) {
        fwd_off_[lv] = flist.size();
        sm_.gfo[lv] = gf.size();
        sm_.gbo[lv] = gb.size();
        sm_.gpo[lv] = gp.size();
        sm_.gro[lv] = gr.size();
        sm_.gco[lv] = gc.size();
        sm_.gso[lv] = gs.size();
        for (int s : P.levels[lv]) {
            const int ne = ne_[s];
            if (ne == 0) continue;
            const int li = (int)(flist.size() - fwd_off_[lv]);
            flist.push_back(s);
            lwb.push_back(wb[s]);
            lyb.push_back(yb[s]);
            lpb.push_back(pt);
            sm_.maxne[lv] = std::max(sm_.maxne[lv], ne);
            const int r = m_[s] - ne;
            const int ncc = (ne + GW - 1) / GW;
            for (int rc = 0; rc < (r + 63) / 64; ++rc)
                for (int cc = 0; cc < ncc; ++cc) gf.push_back(make_int4(li, rc, cc, 0));
            if (ncc > 1) {
                pt += (long long)ncc * r;
                for (int c = 0; c < (r + 255) / 256; ++c) gs.push_back(make_int2(li, c));
            }
            for (int c = 0; c < (ne + 3) / 4; ++c) gb.push_back(make_int2(li, c));
            for (int c = 0; c < (ne + 255) / 256; ++c) gp.push_back(make_int2(li, c));
            for (int c = 0; c < (ne + 255) / 256; ++c) gd.push_back(make_int2((int)flist.size() - 1, c));
            for (int t = 0; t < (ne + SB - 1) / SB; ++t) gc.push_back(make_int2(li, t));
            if (ne > SB)
                for (int c = 0; c < (ne - SB + 127) / 128; ++c) gr.push_back(make_int2(li, c));
        }
        fwd_cnt_[lv] = (int)(flist.size() - fwd_off_[lv]);
    }
    fwd_off_[nl] = flist.size();
    sm_.gfo[nl] = gf.size();
    sm_.gbo[nl] = gb.size();
    sm_.gpo[nl] = gp.size();
    sm_.gro[nl] = gr.size();
    sm_.gco[nl] = gc.size();
    sm_.gso[nl] = gs.size();
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
    const size_t o_pb = push(lpb.data(), lpb.size() * sizeof(long long));
    const size_t o_gf = push(gf.data(), gf.size() * sizeof(int4));
    const size_t o_gs = push(gs.data(), gs.size() * sizeof(int2));
    const size_t o_gb = push(gb.data(), gb.size() * sizeof(int2));
    const size_t o_gp = push(gp.data(), gp.size() * sizeof(int2));
    const size_t o_gd = push(gd.data(), gd.size() * sizeof(int2));
    const size_t o_gr = push(gr.data(), gr.size() * sizeof(int2));
    const size_t o_gc = push(gc.data(), gc.size() * sizeof(int2));
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
    sm_.pbase = (const long long *)(b + o_pb);
    sm_.gf = (const int4 *)(b + o_gf);
    sm_.gs = (const int2 *)(b + o_gs);
    sm_.gb = (const int2 *)(b + o_gb);
    sm_.gp = (const int2 *)(b + o_gp);
    sm_.gd = (const int2 *)(b + o_gd);
    sm_.ngd = (int)gd.size();
    sm_.gr = (const int2 *)(b + o_gr);
    sm_.gc = (const int2 *)(b + o_gc);
    sm_.unres = (const int *)(b + o_un);
    sm_.nunres = (int)unres.size();
    sm_.wtotal = std::max<long long>(wt, 1);
    sm_.ytotal = std::max<long long>(yt, 1);
    sm_.ptotal = std::max<long long>(pt, 1);
    solve_ready_ = true;
}

void NumericGPU::solve(double *d_rhs, i64 ldrhs, int nrhs, const double *d_scaling) {
    if (!solve_ready_) prepare_solve();
    const Prepared &P = *prep_;
    const i64 n = P.n;
    if (n == 0 || nrhs <= 0) return;
    double *xp = (double *)sol_buf_.get((size_t)n * nrhs * sizeof(double));
    double *W = (double *)ybuf_.get(((size_t)sm_.wtotal + (size_t)sm_.ytotal + (size_t)sm_.ptotal) * nrhs *
                                     sizeof(double));
    double *Y = W + (size_t)sm_.wtotal * nrhs;
    double *Pp = Y + (size_t)sm_.ytotal * nrhs;
    const long long wt = sm_.wtotal, yt = sm_.ytotal, pt = sm_.ptotal, nn = n;
    const dim3 gp((unsigned)((n + 255) / 256), nrhs);
    hipLaunchKernelGGL(k_perm_in, gp, dim3(256), 0, stream_, d_rhs, (long long)ldrhs, xp, nn,
                       P.d_perm_fwd, d_scaling);
    const int nl = (int)P.levels.size();
    auto cnt_of = [](const std::vector<size_t> &o, int lv) { return (unsigned)(o[lv + 1] - o[lv]); };
    // forward (solve.rs:89-92)
    for (int lv = 0; lv < nl; ++lv) {
        const int c = fwd_cnt_[lv];
        if (c == 0) continue;
        const int *fl = sm_.flist + fwd_off_[lv];
        const long long *wb = sm_.wbase + fwd_off_[lv];
        const long long *yb = sm_.ybase + fwd_off_[lv];
        const int npull = (int)cnt_of(sm_.gpo, lv);
        hipLaunchKernelGGL(k_fwd_first, dim3((unsigned)(npull + c), nrhs), dim3(256), 0, stream_, d_fr_, fl,
                           sm_.gp + sm_.gpo[lv], npull, sm_.ptr, sm_.ent, (const double *)Y, yt, wb, wt, W, xp,
                           nn);
        if (cnt_of(sm_.gro, lv))
            for (int jb = 0; jb + SB < sm_.maxne[lv]; jb += SB)
                hipLaunchKernelGGL(k_fwd_rows, dim3(cnt_of(sm_.gro, lv), nrhs), dim3(256), 0, stream_, d_fr_, fl,
                                   sm_.gr + sm_.gro[lv], jb, wb, wt, W, xp, nn);
        const long long *pb = sm_.pbase + fwd_off_[lv];
        if (cnt_of(sm_.gfo, lv))
            hipLaunchKernelGGL(k_fwd_gemv, dim3(cnt_of(sm_.gfo, lv), nrhs), dim3(256), 0, stream_, d_fr_, fl,
                               sm_.gf + sm_.gfo[lv], wb, yb, pb, wt, yt, pt, (const double *)W, Y, Pp);
        if (cnt_of(sm_.gso, lv))
            hipLaunchKernelGGL(k_fwd_gemv_sum, dim3(cnt_of(sm_.gso, lv), nrhs), dim3(256), 0, stream_, d_fr_,
                               fl, sm_.gs + sm_.gso[lv], yb, pb, yt, pt, (const double *)Pp, Y);
    }
    if (sm_.nunres > 0)
        hipLaunchKernelGGL(k_pull_rows, dim3((sm_.nunres + 255) / 256, nrhs), dim3(256), 0, stream_,
                           sm_.unres, sm_.nunres, sm_.ptr, sm_.ent, (const double *)Y, yt, xp, nn);
    // diagonal (solve.rs:97-135): every front at once (the fronts' columns are disjoint)
    if (sm_.ngd > 0)
        hipLaunchKernelGGL(k_diag, dim3((unsigned)sm_.ngd, nrhs), dim3(256), 0, stream_, d_fr_, sm_.flist, sm_.gd,
                           xp, nn);
    // backward (solve.rs:139-141)
    for (int lv = nl - 1; lv >= 0; --lv) {
        const int c = fwd_cnt_[lv];
        if (c == 0) continue;
        const int *fl = sm_.flist + fwd_off_[lv];
        const long long *wb = sm_.wbase + fwd_off_[lv];
        if (cnt_of(sm_.gbo, lv))
            hipLaunchKernelGGL(k_bwd_gemv, dim3(cnt_of(sm_.gbo, lv), nrhs), dim3(256), 0, stream_, d_fr_, fl,
                               sm_.gb + sm_.gbo[lv], wb, wt, W, (const double *)xp, nn);
        const int nblk = (sm_.maxne[lv] + SB - 1) / SB;
        for (int b = nblk - 1; b >= 0; --b) {
            if (nblk > 1)
                hipLaunchKernelGGL(k_bwd_step<true>, dim3(cnt_of(sm_.gco, lv), nrhs), dim3(256), 0, stream_, d_fr_,
                                   fl, sm_.gc + sm_.gco[lv], b, wb, wt, W, xp, nn);
            else
                hipLaunchKernelGGL(k_bwd_step<false>, dim3(cnt_of(sm_.gco, lv), nrhs), dim3(256), 0, stream_,
                                   d_fr_, fl, sm_.gc + sm_.gco[lv], b, wb, wt, W, xp, nn);
        }
    }
    hipLaunchKernelGGL(k_perm_out, gp, dim3(256), 0, stream_, d_rhs, (long long)ldrhs, (const double *)xp, nn,
                       P.d_perm_fwd, d_scaling);
}

}  // namespace rv
