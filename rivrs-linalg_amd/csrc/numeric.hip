// numeric.hip -- MI355X (gfx950) numeric factorization of the multifrontal APTP LDL^T.
//
// Fronts are processed level by level of the assembly tree (leaves first), all fronts of a
// level in the same launches.  Per level:
//   1. assembly      k_scatter_orig (original entries) + k_extend_add (one launch per child
//                    ordinal, so the children of a parent are summed in reference order)
//   2. dense factor  fronts with k < ib or inside a small-leaf subtree: k_tpp (threshold
//                    partial pivoting, one wavefront per front);
//                    other fronts: APTP with ib x ib complete pivoting blocks
//                    (k_block_factor -> k_panel_apply -> k_panel_update -> k_delay_swap per
//                    inner block, the trailing update restricted to the outer block, and the
//                    rank-nb outer update on MFMA), then k_tpp on the columns APTP left.
//   3. contribution  k_schur: C -= (L21 D) L21^T on the NFS x NFS block, FP64 MFMA.
//   4. k_collect     eliminated counts and delayed columns back to the host, which builds
//                    the row lists and extend-add maps of the next level.
// The scalar pivoting arithmetic is compiled without FMA contraction so that every decision
// step reproduces the reference's operation sequence (sparse/src/symmetric/factor.rs).
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>

#include "rv_numeric.h"

namespace rv {

// ===========================================================================
// Arena / buffers
// ===========================================================================
void *DevArena::alloc(size_t bytes, size_t align) {
    bytes = (bytes + align - 1) / align * align;
    if (bytes == 0) bytes = align;
    for (Chunk &c : chunks_) {
        size_t off = (c.used + align - 1) / align * align;
        if (off + bytes <= c.size) {
            c.used = off + bytes;
            return c.base + off;
        }
    }
    size_t sz = std::max(bytes, (size_t)256 << 20);
    void *p = nullptr;
    RV_HIP(hipMalloc(&p, sz));
    chunks_.push_back(Chunk{(char *)p, sz, bytes});
    reserved_ += sz;
    return p;
}

void DevArena::reset_keep() {
    for (Chunk &c : chunks_) c.used = 0;
}

void DevArena::release() {
    for (Chunk &c : chunks_) (void)hipFree(c.base);
    chunks_.clear();
    reserved_ = 0;
}

void *DevBuf::get(size_t bytes) {
    if (bytes <= cap && p) return p;
    if (p) (void)hipFree(p);
    cap = std::max(bytes, cap * 3 / 2);
    cap = std::max(cap, (size_t)4096);
    RV_HIP(hipMalloc(&p, cap));
    return p;
}

DevBuf::~DevBuf() {
    if (p) (void)hipFree(p);
}

// ===========================================================================
// Device helpers
// ===========================================================================
__device__ __forceinline__ bool aptp_active(int p, int nb, int ib, int ob, int it) {
    const bool two = p > nb;
    const int nouter = two ? (p + nb - 1) / nb : 1;
    if (ob >= nouter) return false;
    const int obcols = two ? min(nb, p - ob * nb) : p;
    return it < (obcols + ib - 1) / ib;
}

__device__ __forceinline__ double wave_fmax(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

// larger value wins; equal values -> smaller index
__device__ __forceinline__ void wave_argmax(double &v, int &idx) {
    for (int o = 32; o > 0; o >>= 1) {
        const double v2 = __shfl_xor(v, o);
        const int i2 = __shfl_xor(idx, o);
        if (v2 > v || (v2 == v && i2 < idx)) { v = v2; idx = i2; }
    }
}

__device__ __forceinline__ int wave_min_int(int v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return v;
}

__device__ __forceinline__ void atomic_max_nonneg(unsigned long long *p, double v) {
    if (v > 0.0) atomicMax(p, (unsigned long long)__double_as_longlong(v));
}

// effective nelim of an inner block: first failing pivot, never splitting a 2x2
// (apply_and_check + adjust_for_2x2_boundary, factor.rs:1160-1172, 1295-1408)
__device__ __forceinline__ int effective_nelim(const AptpState &S) {
    int e = min(S.fail_min, S.block_nelim);
    if (e > 0 && S.bkind[e - 1] == 2) e -= 1;
    return e;
}

// W = L * D entry (compute_ld_into, factor.rs:1536-1563) for row values of the pivot columns.
// kind: 1 = 1x1, 2 = first of 2x2, 3 = second of 2x2.
__device__ __forceinline__ double ld_entry(const double *rowbase, size_t lda, int c, int kind,
                                           const double *diag, const double *off) {
    if (kind == 1) return rowbase[c * lda] * diag[c];
    if (kind == 2) {
        const double l1 = rowbase[c * lda], l2 = rowbase[(c + 1) * lda];
        return l1 * diag[c] + l2 * off[c];
    }
    if (kind == 3) {
        const double l1 = rowbase[(c - 1) * lda], l2 = rowbase[c * lda];
        return l1 * off[c - 1] + l2 * diag[c];
    }
    return 0.0;
}

// ===========================================================================
// Assembly
// ===========================================================================
struct ScatterChunk { int front; int begin; int end; int pad; };

__global__ void k_set_fronts(FrontDev *__restrict__ fr, const int *__restrict__ ids,
                             const FrontDev *__restrict__ src, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) fr[ids[i]] = src[i];
}

// perm = identity for every front of the level
__global__ void k_init_perm(const FrontDev *__restrict__ fr, const int *__restrict__ ids, int n) {
    const FrontDev &F = fr[ids[blockIdx.x]];
    for (int i = threadIdx.x; i < F.m; i += blockDim.x) F.perm[i] = i;
}

// Original entries (numeric.rs:1306-1344; fallback scatter numeric.rs:2376-2435 when delayed
// columns shift the pattern rows).  The two reference paths scale as (v*s_r)*s_c and
// v*(s_r*s_c) respectively; both are reproduced.
__global__ __launch_bounds__(256) void k_scatter_orig(const FrontDev *__restrict__ fr,
                                                      const ScatterChunk *__restrict__ ch,
                                                      const AmapEntry *__restrict__ amap,
                                                      const double *__restrict__ val,
                                                      const double *__restrict__ sc) {
    const ScatterChunk c = ch[blockIdx.x];
    const int e = c.begin + threadIdx.x;
    if (e >= c.end) return;
    const FrontDev &F = fr[c.front];
    const AmapEntry a = amap[e];
    int r = (int)a.row, col = (int)a.col;
    if (r >= F.n_owned) r += F.ndelay_in;
    if (col >= F.n_owned) col += F.ndelay_in;
    double v = val[a.src];
    if (sc) v = (F.ndelay_in == 0) ? (v * sc[a.sr]) * sc[a.sc] : v * (sc[a.sr] * sc[a.sc]);
    F.F[r + (size_t)col * F.m] += v;
}

// Extend-add of one 64x64 lower tile of a child's contribution block (the trailing
// (m-ne)^2 block of the child front) into the parent, with both index maps staged in LDS
// (extend_add numeric.rs:2445-2485, extend_add_mapped 2507-2549).
__global__ __launch_bounds__(256) void k_extend_add(const FrontDev *__restrict__ fr,
                                                    const EaItem *__restrict__ items) {
    const EaItem it = items[blockIdx.x];
    const FrontDev &C = fr[it.child];
    const FrontDev &P = fr[it.parent];
    const int ne = C.ne;
    const int sz = C.m - ne;
    const size_t ldc = C.m;
    const double *CB = C.F + ne + (size_t)ne * ldc;
    const int i0 = it.ti * 64, j0 = it.tj * 64;
    __shared__ unsigned mi[64], mj[64];
    const int t = threadIdx.x;
    if (t < 64) mi[t] = (i0 + t < sz) ? it.map[i0 + t] : 0u;
    else if (t < 128) mj[t - 64] = (j0 + t - 64 < sz) ? it.map[j0 + t - 64] : 0u;
    __syncthreads();
    const int ii = t & 63;
    const int i = i0 + ii;
    if (i >= sz) return;
    double *PF = P.F;
    const size_t ldp = P.m;
    for (int jj = t >> 6; jj < 64; jj += 4) {
        const int j = j0 + jj;
        if (j >= sz || i < j) continue;
        const double v = CB[i + (size_t)j * ldc];
        if (it.skip_zero && v == 0.0) continue;
        const unsigned li = mi[ii], lj = mj[jj];
        if (li >= lj) PF[li + (size_t)lj * ldp] += v;
        else PF[lj + (size_t)li * ldp] += v;
    }
}

// ===========================================================================
// APTP: inner block of complete pivoting (factor_block_diagonal, factor.rs:891-1153)
// One wavefront per front; the block lives in LDS.
// ===========================================================================
__device__ __forceinline__ void lds_block_swap(double (*b)[MAX_IB + 1], int i, int j, int bs,
                                               int lane) {
    if (i > j) { const int t = i; i = j; j = t; }
    for (int q = lane; q < bs; q += 64) {
        if (q < i) { const double t = b[i][q]; b[i][q] = b[j][q]; b[j][q] = t; }
        else if (q > i && q < j) { const double t = b[q][i]; b[q][i] = b[j][q]; b[j][q] = t; }
        else if (q > j) { const double t = b[q][i]; b[q][i] = b[q][j]; b[q][j] = t; }
    }
    if (lane == 0) { const double t = b[i][i]; b[i][i] = b[j][j]; b[j][j] = t; }
}

__global__ __launch_bounds__(64) void k_block_factor(FrontDev *__restrict__ fr,
                                                     AptpState *__restrict__ st,
                                                     const int *__restrict__ flist, int ob, int it,
                                                     int nb, int ib, double small) {
    const int li = blockIdx.x;
    FrontDev &F = fr[flist[li]];
    AptpState &S = st[li];
    if (!aptp_active(F.k, nb, ib, ob, it)) return;
    const int lane = threadIdx.x;
    const size_t lda = F.m;
    double *A = F.F;
    const int k = S.kpos;
    const int bs = min(S.end_pos - k, ib);
    __shared__ double blk[MAX_IB][MAX_IB + 1];
    __shared__ double bkp[MAX_IB][MAX_IB + 1];
    __shared__ int bp[MAX_IB];
    __shared__ int8_t bkind[MAX_IB];
    __shared__ double bdiag[MAX_IB], boff[MAX_IB];
    __shared__ int ptmp[MAX_IB];
    for (int e = lane; e < bs * bs; e += 64) {
        const int i = e % bs, j = e / bs;
        if (i >= j) {
            const double v = A[(k + i) + (k + j) * lda];
            blk[i][j] = v;
            bkp[i][j] = v;
            S.backup[i + j * lda] = v;
        }
    }
    if (lane < MAX_IB) { bp[lane] = lane; bkind[lane] = 0; bdiag[lane] = 0.0; boff[lane] = 0.0; }
    __syncthreads();
    int kk = 0;
    double maxl = 0.0;
    while (kk < bs) {
        const int cur = kk;
        const int w = bs - cur;
        double bv = 0.0;
        int bpos = INT_MAX;
        for (int e = lane; e < w * w; e += 64) {
            const int jj = cur + e / w, ii = cur + e % w;
            if (ii >= jj) {
                const double v = fabs(blk[ii][jj]);
                if (v > bv) { bv = v; bpos = jj * 64 + ii; }
            }
        }
        wave_argmax(bv, bpos);
        if (bv < small) break;
        const int mr = (bpos == INT_MAX) ? cur : bpos % 64;
        const int mc = (bpos == INT_MAX) ? cur : bpos / 64;
        int piv2 = 0;  // 0: 1x1 at cur, 1: 2x2 at cur
        if (mr == mc) {
            if (mr != cur) {
                lds_block_swap(blk, cur, mr, bs, lane);
                if (lane == 0) { const int t = bp[cur]; bp[cur] = bp[mr]; bp[mr] = t; }
            }
        } else {
            const int t = mr, mi = mc;
            const double amm = blk[mi][mi], att = blk[t][t], atm = blk[t][mi];
            const double delta = amm * att - atm * atm;
            if (fabs(delta) >= 0.5 * atm * atm) {
                piv2 = 1;
                if (mi != cur) {
                    lds_block_swap(blk, cur, mi, bs, lane);
                    if (lane == 0) { const int q = bp[cur]; bp[cur] = bp[mi]; bp[mi] = q; }
                }
                __syncthreads();
                const int nt = (t == cur) ? mi : t;
                if (nt != cur + 1) {
                    lds_block_swap(blk, cur + 1, nt, bs, lane);
                    if (lane == 0) { const int q = bp[cur + 1]; bp[cur + 1] = bp[nt]; bp[nt] = q; }
                }
            } else {
                const int pos = (fabs(amm) >= fabs(att)) ? mi : t;
                if (pos != cur) {
                    lds_block_swap(blk, cur, pos, bs, lane);
                    if (lane == 0) { const int q = bp[cur]; bp[cur] = bp[pos]; bp[pos] = q; }
                }
            }
        }
        __syncthreads();
        if (!piv2) {
            const double dkk = blk[cur][cur];
            const double inv = 1.0 / dkk;
            double ml = 0.0;
            for (int i = cur + 1 + lane; i < bs; i += 64) {
                const double l = blk[i][cur] * inv;
                blk[i][cur] = l;
                if (fabs(l) > ml) ml = fabs(l);
            }
            ml = wave_fmax(ml);
            __syncthreads();
            const int w1 = bs - cur - 1;
            for (int e = lane; e < w1 * w1; e += 64) {
                const int jj = cur + 1 + e / w1, ii = cur + 1 + e % w1;
                if (ii >= jj) {
                    const double ldl = blk[jj][cur] * dkk;
                    blk[ii][jj] -= blk[ii][cur] * ldl;
                }
            }
            __syncthreads();
            if (lane == 0) { bkind[cur] = 1; bdiag[cur] = dkk; }
            if (ml > maxl) maxl = ml;
            kk += 1;
        } else {
            const double a11 = blk[cur][cur], a22 = blk[cur + 1][cur + 1], a21 = blk[cur + 1][cur];
            const double det = a11 * a22 - a21 * a21;
            const double inv_det = 1.0 / det;
            double ml = 0.0;
            for (int i = cur + 2 + lane; i < bs; i += 64) {
                const double ai1 = blk[i][cur], ai2 = blk[i][cur + 1];
                const double l1 = (ai1 * a22 - ai2 * a21) * inv_det;
                const double l2 = (ai2 * a11 - ai1 * a21) * inv_det;
                blk[i][cur] = l1;
                blk[i][cur + 1] = l2;
                if (fabs(l1) > ml) ml = fabs(l1);
                if (fabs(l2) > ml) ml = fabs(l2);
            }
            ml = wave_fmax(ml);
            __syncthreads();
            const int w2 = bs - cur - 2;
            for (int e = lane; e < w2 * w2; e += 64) {
                const int jj = cur + 2 + e / w2, ii = cur + 2 + e % w2;
                if (ii >= jj) {
                    const double l1j = blk[jj][cur], l2j = blk[jj][cur + 1];
                    const double wj1 = l1j * a11 + l2j * a21;
                    const double wj2 = l1j * a21 + l2j * a22;
                    blk[ii][jj] -= blk[ii][cur] * wj1 + blk[ii][cur + 1] * wj2;
                }
            }
            __syncthreads();
            if (lane == 0) {
                bkind[cur] = 2; bkind[cur + 1] = 3;
                bdiag[cur] = a11; bdiag[cur + 1] = a22; boff[cur] = a21;
            }
            if (ml > maxl) maxl = ml;
            kk += 2;
        }
        __syncthreads();
    }
    const int bn = kk;
    if (bn == 0) {
        // whole block singular: restore from the backup (factor.rs:1866-1869)
        for (int e = lane; e < bs * bs; e += 64) {
            const int i = e % bs, j = e / bs;
            if (i >= j) A[(k + i) + (k + j) * lda] = bkp[i][j];
        }
    } else {
        for (int e = lane; e < bs * bs; e += 64) {
            const int i = e % bs, j = e / bs;
            if (i >= j) A[(k + i) + (k + j) * lda] = blk[i][j];
        }
        __syncthreads();
        // zero D off-diagonals of 2x2 pivots so the TRSM reads them as L11 zeros (factor.rs:1898-1912)
        if (lane < bn && bkind[lane] == 2) A[(k + lane + 1) + (k + lane) * lda] = 0.0;
        // column order by the block permutation (factor.rs:1926-1930)
        if (lane < bs) ptmp[lane] = F.perm[k + lane];
        __syncthreads();
        if (lane < bs) F.perm[k + lane] = ptmp[bp[lane]];
    }
    if (lane < MAX_IB) {
        S.bperm[lane] = bp[lane];
        S.bkind[lane] = (lane < bn) ? bkind[lane] : 0;
        S.bdiag[lane] = bdiag[lane];
        S.boff[lane] = boff[lane];
    }
    if (lane == 0) {
        S.bs = bs;
        S.block_nelim = bn;
        S.fail_min = bn;
        atomic_max_nonneg(&F.max_l_bits, maxl);
    }
}

// ===========================================================================
// APTP: panel rows below the block -- backup, column permutation, TRSM with L11, D^-1
// scaling and the a-posteriori threshold check (apply_and_check, factor.rs:1295-1408);
// plus the block row permutation applied to the already-factored columns (factor.rs:1914-1924).
// ===========================================================================
__global__ __launch_bounds__(256) void k_panel_apply(FrontDev *__restrict__ fr,
                                                     AptpState *__restrict__ st,
                                                     const int *__restrict__ flist,
                                                     const int2 *__restrict__ items, int ob,
                                                     int it, int nb, int ib, double threshold) {
    const int2 itm = items[blockIdx.x];
    const int li = itm.x, chunk = itm.y;
    FrontDev &F = fr[flist[li]];
    AptpState &S = st[li];
    if (!aptp_active(F.k, nb, ib, ob, it)) return;
    const int bn = S.block_nelim;
    if (bn == 0) return;
    const int k = S.kpos, bs = S.bs, be = k + bs, m = F.m;
    const size_t lda = m;
    double *A = F.F;
    __shared__ double L11[MAX_IB][MAX_IB + 1];
    __shared__ int bp[MAX_IB];
    __shared__ int8_t bkind[MAX_IB];
    __shared__ double dinv[MAX_IB], boff[MAX_IB], bdiag[MAX_IB];
    __shared__ int wgmin;
    const int tid = threadIdx.x;
    for (int e = tid; e < bs * bs; e += 256) {
        const int i = e % bs, j = e / bs;
        L11[i][j] = (i > j && j < bn && i < bn) ? A[(k + i) + (k + j) * lda] : 0.0;
    }
    if (tid < MAX_IB) {
        bp[tid] = S.bperm[tid];
        bkind[tid] = S.bkind[tid];
        bdiag[tid] = S.bdiag[tid];
        boff[tid] = S.boff[tid];
    }
    if (tid == 0) wgmin = INT_MAX;
    __syncthreads();
    if (tid < bn) {
        const int kd = bkind[tid];
        if (kd == 1) dinv[tid] = 1.0 / bdiag[tid];
        else if (kd == 2) {
            const double det = bdiag[tid] * bdiag[tid + 1] - boff[tid] * boff[tid];
            dinv[tid] = 1.0 / det;
        }
    }
    __syncthreads();
    const int r = be + chunk * 256 + tid;
    if (r < m) {
        double *row = A + r;
        double *bk = S.backup + (r - k);
        double y[MAX_IB];
#pragma unroll
        for (int j = 0; j < MAX_IB; ++j)
            if (j < bs) bk[j * lda] = row[(k + j) * lda];
#pragma unroll
        for (int j = 0; j < MAX_IB; ++j) y[j] = (j < bs) ? row[(k + bp[j]) * lda] : 0.0;
#pragma unroll
        for (int j = 0; j < MAX_IB; ++j) {
            if (j < bn) {
                double s = y[j];
#pragma unroll
                for (int i = 0; i < j; ++i) s -= L11[j][i] * y[i];
                y[j] = s;
            }
        }
#pragma unroll
        for (int c = 0; c < MAX_IB; ++c) {
            if (c < bn) {
                const int kd = bkind[c];
                if (kd == 1) {
                    y[c] *= dinv[c];
                } else if (kd == 2 && c + 1 < MAX_IB) {
                    const double x1 = y[c], x2 = y[c + 1];
                    y[c] = (x1 * bdiag[c + 1] - x2 * boff[c]) * dinv[c];
                    y[c + 1] = (x2 * bdiag[c] - x1 * boff[c]) * dinv[c];
                }
            }
        }
#pragma unroll
        for (int j = 0; j < MAX_IB; ++j)
            if (j < bs) row[(k + j) * lda] = y[j];
        const double bound = 1.0 / threshold;
        int first = INT_MAX;
#pragma unroll
        for (int c = MAX_IB - 1; c >= 0; --c)
            if (c < bn && fabs(y[c]) >= bound) first = c;
        if (first != INT_MAX) {
            const int start = (bkind[first] == 3) ? first - 1 : first;
            atomicMin(&wgmin, start);
        }
    }
    // row permutation of the block rows in the previously factored columns
    const int cc = chunk * 256 + tid;
    if (cc < k) {
        double t[MAX_IB];
#pragma unroll
        for (int i = 0; i < MAX_IB; ++i)
            if (i < bs) t[i] = A[(k + bp[i]) + cc * lda];
#pragma unroll
        for (int i = 0; i < MAX_IB; ++i)
            if (i < bs) A[(k + i) + cc * lda] = t[i];
    }
    __syncthreads();
    if (tid == 0 && wgmin != INT_MAX) atomicMin(&S.fail_min, wgmin);
}

// ===========================================================================
// APTP: partial restore of failed columns, Schur updates of the trailing columns of the
// outer block and cross terms for failed columns, max|L| statistics
// (factor.rs:1945-2008: restore_diagonal_with_perm, update_trailing, update_cross_terms).
// Tile: 64 rows (relative to kpos) x 32 columns; column tile 0 = the current block.
// ===========================================================================
__global__ __launch_bounds__(256) void k_panel_update(FrontDev *__restrict__ fr,
                                                      AptpState *__restrict__ st,
                                                      const int *__restrict__ flist,
                                                      const int4 *__restrict__ items, int ob,
                                                      int it, int nb, int ib) {
    const int4 itm = items[blockIdx.x];
    const int li = itm.x, rt = itm.y, q = itm.z;
    FrontDev &F = fr[flist[li]];
    AptpState &S = st[li];
    if (!aptp_active(F.k, nb, ib, ob, it)) return;
    const int bn = S.block_nelim;
    if (bn == 0) return;
    const int k = S.kpos, bs = S.bs, be = k + bs, m = F.m;
    const int obe = S.ob_start + S.ob_cols;
    const size_t lda = m;
    double *A = F.F;
    const int nelim = effective_nelim(S);
    int c_lo, c_hi;
    if (q == 0) { c_lo = k; c_hi = be; }
    else {
        c_lo = be + 32 * (q - 1);
        c_hi = min(c_lo + 32, obe);
        if (c_lo >= obe || nelim == 0) return;
    }
    const int r_lo = k + 64 * rt;
    if (r_lo >= m) return;
    const int r_hi = min(r_lo + 64, m);
    if (q >= 1 && r_hi <= c_lo) return;
    __shared__ double Wr[64][MAX_IB + 1];
    __shared__ double Lc[32][MAX_IB + 1];
    __shared__ int bp[MAX_IB];
    __shared__ int kind[MAX_IB];
    __shared__ double dg[MAX_IB], of[MAX_IB];
    const int tid = threadIdx.x;
    if (tid < MAX_IB) {
        bp[tid] = S.bperm[tid];
        kind[tid] = S.bkind[tid];
        dg[tid] = S.bdiag[tid];
        of[tid] = S.boff[tid];
    }
    __syncthreads();
    const int ncol = c_hi - c_lo;
    // W rows of the tile (only passed columns)
    for (int e = tid; e < 64 * nelim; e += 256) {
        const int rr = e % 64, j = e / 64;
        const int r = r_lo + rr;
        Wr[rr][j] = (r < r_hi) ? ld_entry(A + r + k * lda, lda, j, kind[j], dg, of) : 0.0;
    }
    if (q >= 1) {
        for (int e = tid; e < ncol * nelim; e += 256) {
            const int cc = e % ncol, j = e / ncol;
            Lc[cc][j] = A[(c_lo + cc) + (k + j) * lda];
        }
    } else {
        // L rows of the failed block rows: Lb[jf][c] = A(k+nelim+jf, k+c)
        const int nf = bs - nelim;
        for (int e = tid; e < nf * nelim; e += 256) {
            const int jf = e % nf, c = e / nf;
            Lc[jf][c] = A[(k + nelim + jf) + (k + c) * lda];
        }
    }
    __syncthreads();
    const int rr = tid & 63;
    const int r = r_lo + rr;
    const bool valid = r < r_hi;
    double *row = A + r;
    if (q >= 1) {
        if (!valid) return;
        for (int cc = tid >> 6; cc < ncol; cc += 4) {
            const int c = c_lo + cc;
            if (r < c) continue;
            double s = 0.0;
            for (int j = 0; j < nelim; ++j) s += Wr[rr][j] * Lc[cc][j];
            row[c * lda] -= s;
        }
        return;
    }
    // current block columns
    double ml = 0.0;
    for (int j = tid >> 6; valid && j < bs; j += 4) {
        const int c = k + j;
        if (j >= nelim) {
            // partial restore (only when the check removed pivots)
            if (nelim < bn) {
                if (r < be) {
                    const int i = r - k;
                    if (i >= j) {
                        const int br = bp[i], bc = bp[j];
                        row[c * lda] = (br >= bc) ? S.backup[br + bc * lda] : S.backup[bc + br * lda];
                    }
                } else {
                    row[c * lda] = S.backup[(r - k) + bp[j] * lda];
                }
            }
            // cross terms from the passed columns
            if (nelim > 0) {
                const int jf = j - nelim;
                if (r < be) {
                    const int i = r - k - nelim;
                    if (i >= jf) {
                        double s = 0.0;
                        for (int cc2 = 0; cc2 < nelim; ++cc2) s += Wr[rr][cc2] * Lc[jf][cc2];
                        row[c * lda] -= s;
                    }
                } else {
                    double s = 0.0;
                    for (int cc2 = 0; cc2 < nelim; ++cc2) s += Wr[rr][cc2] * Lc[jf][cc2];
                    row[c * lda] -= s;
                }
            }
        } else if (r >= be) {
            const double v = fabs(row[c * lda]);
            if (v > ml) ml = v;
        }
    }
    ml = wave_fmax(ml);
    if ((tid & 63) == 0) atomic_max_nonneg(&F.max_l_bits, ml);
}

// ===========================================================================
// Symmetric line permutation for a sequence of swap_symmetric calls (factor.rs:452-505):
// the composite is a symmetric permutation touching at most 2*ib lines; the touched lines
// are snapshot and rewritten from the snapshot.
// ===========================================================================
__device__ void apply_swaps(FrontDev &F, AptpState &S, const int *sa, const int *sb, int nsw,
                            int *U, int *src, int *cnt, int *ptmp) {
    const int tid = threadIdx.x;
    const int m = F.m;
    const size_t lda = m;
    double *A = F.F;
    if (tid == 0) {
        int nU = 0;
        for (int s = 0; s < nsw; ++s) {
            const int x = sa[s], y = sb[s];
            int ix = -1, iy = -1;
            for (int u = 0; u < nU; ++u) {
                if (U[u] == x) ix = u;
                if (U[u] == y) iy = u;
            }
            if (ix < 0) { U[nU] = x; src[nU] = x; ix = nU++; }
            if (iy < 0) { U[nU] = y; src[nU] = y; iy = nU++; }
            const int t = src[ix]; src[ix] = src[iy]; src[iy] = t;
        }
        *cnt = nU;
    }
    __syncthreads();
    const int nU = *cnt;
    if (nU == 0) return;
    int *mark = S.mark;
    double *buf = S.swapbuf;
    for (int u = tid; u < nU; u += blockDim.x) mark[U[u]] = u + 1;
    for (int e = tid; e < nU * m; e += blockDim.x) {
        const int ui = e / m, q = e % m;
        const int u = U[ui];
        buf[e] = (u >= q) ? A[u + q * lda] : A[q + u * lda];
    }
    if (tid < nU) ptmp[tid] = F.perm[U[tid]];
    __syncthreads();
    for (int pi = 0; pi < nU; ++pi) {
        const int p = U[pi];
        const int sp = src[pi];
        if (sp == p) continue;
        const int sidx = mark[sp] - 1;
        const double *line = buf + (size_t)sidx * m;
        for (int q = tid; q < m; q += blockDim.x) {
            const int mq = mark[q];
            const int qs = mq ? src[mq - 1] : q;
            const double v = line[qs];
            if (p >= q) A[p + q * lda] = v; else A[q + p * lda] = v;
        }
    }
    if (tid < nU) {
        // new perm[p] = old perm[src[p]]
        const int sp = src[tid];
        F.perm[U[tid]] = ptmp[mark[sp] - 1];
    }
    __syncthreads();
    for (int u = tid; u < nU; u += blockDim.x) mark[U[u]] = 0;
    __syncthreads();
}

// End of an inner block: accumulate D and counts for the passed columns, then delay the
// failed columns to the end of the outer block (factor.rs:1962-2077).
__global__ __launch_bounds__(256) void k_delay_swap(FrontDev *__restrict__ fr,
                                                    AptpState *__restrict__ st,
                                                    const int *__restrict__ flist, int ob, int it,
                                                    int nb, int ib) {
    const int li = blockIdx.x;
    FrontDev &F = fr[flist[li]];
    AptpState &S = st[li];
    if (!aptp_active(F.k, nb, ib, ob, it)) return;
    __shared__ int sa[MAX_IB], sb[MAX_IB], U[2 * MAX_IB], src[2 * MAX_IB], ptmp[2 * MAX_IB];
    __shared__ int nsw, cnt, s_nelim, s_end;
    const int tid = threadIdx.x;
    if (tid == 0) {
        const int k = S.kpos, bs = S.bs, bn = S.block_nelim;
        int end_pos = S.end_pos;
        int n = 0;
        int nelim;
        if (bn == 0) {
            nelim = 0;
            for (int off = 0; off < bs; ++off) {
                end_pos -= 1;
                if (k + off < end_pos) { sa[n] = k + off; sb[n] = end_pos; ++n; }
            }
        } else {
            nelim = effective_nelim(S);
            int col = 0, n1 = 0, n2 = 0;
            while (col < nelim) {
                const int kd = S.bkind[col];
                if (kd == 1) {
                    F.dkind[k + col] = 1; F.ddiag[k + col] = S.bdiag[col]; F.doff[k + col] = 0.0;
                    ++n1; col += 1;
                } else if (kd == 2 && col + 1 < nelim) {
                    F.dkind[k + col] = 2; F.dkind[k + col + 1] = 3;
                    F.ddiag[k + col] = S.bdiag[col]; F.ddiag[k + col + 1] = S.bdiag[col + 1];
                    F.doff[k + col] = S.boff[col]; F.doff[k + col + 1] = 0.0;
                    ++n2; col += 2;
                } else {
                    col += 1;
                }
            }
            F.n1 += n1;
            F.n2 += n2;
            const int nd = bs - nelim;
            for (int i = nd - 1; i >= 0; --i) {
                const int dp = k + nelim + i;
                end_pos -= 1;
                if (dp < end_pos) { sa[n] = dp; sb[n] = end_pos; ++n; }
            }
        }
        nsw = n;
        s_nelim = nelim;
        s_end = end_pos;
    }
    __syncthreads();
    apply_swaps(F, S, sa, sb, nsw, U, src, &cnt, ptmp);
    if (tid == 0) {
        S.kpos += s_nelim;
        S.end_pos = s_end;
    }
}

// Start of an outer block (two_level_factor, factor.rs:2126-2131); ob == 0 also initialises.
__global__ void k_outer_begin(FrontDev *__restrict__ fr, AptpState *__restrict__ st,
                              const int *__restrict__ flist, int nf, int ob, int nb) {
    const int li = blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= nf) return;
    FrontDev &F = fr[flist[li]];
    AptpState &S = st[li];
    const int p = F.k;
    const bool two = p > nb;
    const int nouter = two ? (p + nb - 1) / nb : 1;
    if (ob >= nouter) return;
    if (ob == 0) { S.gne = 0; S.remaining = p; }
    S.ob_start = S.gne;
    S.ob_cols = two ? min(S.remaining, nb) : p;
    S.kpos = S.ob_start;
    S.end_pos = S.ob_start + S.ob_cols;
}

// End of an outer block: failed columns move to the end of the remaining fully-summed
// region (factor.rs:2195-2206), counters advance; the last block records ne.
__global__ __launch_bounds__(256) void k_outer_end(FrontDev *__restrict__ fr,
                                                   AptpState *__restrict__ st,
                                                   const int *__restrict__ flist, int ob, int nb) {
    const int li = blockIdx.x;
    FrontDev &F = fr[flist[li]];
    AptpState &S = st[li];
    const int p = F.k;
    const bool two = p > nb;
    const int nouter = two ? (p + nb - 1) / nb : 1;
    if (ob >= nouter) return;
    __shared__ int sa[MAX_IB * 8], sb[MAX_IB * 8];
    __shared__ int U[2 * MAX_IB], src[2 * MAX_IB], ptmp[2 * MAX_IB];
    __shared__ int nsw, cnt;
    const int tid = threadIdx.x;
    if (!two) {
        if (tid == 0) F.ne = S.kpos;
        return;
    }
    const int obne = S.kpos - S.ob_start;
    const int nfail = S.ob_cols - obne;
    // process the swap sequence in pieces of at most MAX_IB swaps (<= 2*MAX_IB lines)
    int done = 0;
    while (done < nfail) {
        const int cnt_sw = min(MAX_IB, nfail - done);
        if (tid == 0) {
            int n = 0;
            for (int i = done; i < done + cnt_sw; ++i) {
                const int fp = S.ob_start + obne + i;
                const int end = S.ob_start + S.remaining - 1 - i;
                if (fp < end) { sa[n] = fp; sb[n] = end; ++n; }
            }
            nsw = n;
        }
        __syncthreads();
        apply_swaps(F, S, sa, sb, nsw, U, src, &cnt, ptmp);
        done += cnt_sw;
    }
    __syncthreads();
    if (tid == 0) {
        S.gne += obne;
        S.remaining -= S.ob_cols;
        if (ob == nouter - 1) F.ne = S.gne;
    }
}

// ===========================================================================
// Threshold partial pivoting (tpp_factor / tpp_factor_rect, factor.rs:2268-2947).
// One wavefront per front; every lane follows the same control flow.
// mode 0: whole front from column 0; mode 1: fallback from F.ne.
// ===========================================================================
struct TppCtx {
    double *A;
    size_t lda;
    int m, n;  // rows, fully-summed columns (square storage: all columns stored)
    int lane;
};

__device__ __forceinline__ double tpp_sym(const TppCtx &c, int r, int q) {
    return (r >= q) ? c.A[r + q * c.lda] : c.A[q + r * c.lda];
}

__device__ bool tpp_col_small(const TppCtx &c, int idx, int from, double small) {
    bool big = false;
    for (int q = from + c.lane; q < idx; q += 64)
        if (fabs(c.A[idx + q * c.lda]) >= small) big = true;
    for (int r = idx + c.lane; r < c.m; r += 64)
        if (fabs(c.A[r + idx * c.lda]) >= small) big = true;
    return !__any(big);
}

__device__ int tpp_row_abs_max(const TppCtx &c, int p, int from, int to) {
    if (from >= to) return from;
    const double f0 = fabs(tpp_sym(c, p, from));
    double bv = -1.0;
    int bi = INT_MAX;
    for (int q = from + 1 + c.lane; q < to; q += 64) {
        const double v = fabs(tpp_sym(c, p, q));
        if (v > bv) { bv = v; bi = q; }
    }
    wave_argmax(bv, bi);
    if (f0 != f0) return from;
    return (bi != INT_MAX && bv > f0) ? bi : from;
}

__device__ double tpp_rc_max_excl(const TppCtx &c, int col, int nelim, int exclude) {
    double best = 0.0;
    for (int q = nelim + c.lane; q < col; q += 64)
        if (q != exclude) best = fmax(best, fabs(c.A[col + q * c.lda]));
    for (int r = col + 1 + c.lane; r < c.m; r += 64)
        if (r != exclude) best = fmax(best, fabs(c.A[r + col * c.lda]));
    return wave_fmax(best);
}

__device__ bool tpp_test_2x2(const TppCtx &c, int t, int p, double maxt, double maxp, double u,
                             double small) {
    const double a11 = c.A[t + t * c.lda], a21 = c.A[p + t * c.lda], a22 = c.A[p + p * c.lda];
    const double maxpiv = fmax(fmax(fabs(a11), fabs(a21)), fabs(a22));
    if (maxpiv < small) return false;
    const double detscale = 1.0 / maxpiv;
    const double d0 = (a11 * detscale) * a22;
    const double d1 = (a21 * detscale) * a21;
    const double det = d0 - d1;
    if (fabs(det) < fmax(fmax(small, fabs(d0 / 2.0)), fabs(d1 / 2.0))) return false;
    const double i11 = (a22 * detscale) / det;
    const double i12 = (-a21 * detscale) / det;
    const double i22 = (a11 * detscale) / det;
    if (fmax(maxt, maxp) < small) return true;
    const double x1 = fabs(i11) * maxt + fabs(i12) * maxp;
    const double x2 = fabs(i12) * maxt + fabs(i22) * maxp;
    return u * fmax(x1, x2) < 1.0;
}

__device__ void tpp_swap(const TppCtx &c, int *perm, int i, int j) {
    if (i == j) return;
    if (i > j) { const int t = i; i = j; j = t; }
    double *A = c.A;
    const size_t lda = c.lda;
    for (int q = c.lane; q < c.m; q += 64) {
        if (q < i) { const double t = A[i + q * lda]; A[i + q * lda] = A[j + q * lda]; A[j + q * lda] = t; }
        else if (q > i && q < j) { const double t = A[q + i * lda]; A[q + i * lda] = A[j + q * lda]; A[j + q * lda] = t; }
        else if (q > j) { const double t = A[q + i * lda]; A[q + i * lda] = A[q + j * lda]; A[q + j * lda] = t; }
    }
    if (c.lane == 0) {
        const double t = A[i + i * lda]; A[i + i * lda] = A[j + j * lda]; A[j + j * lda] = t;
        const int pt = perm[i]; perm[i] = perm[j]; perm[j] = pt;
    }
    __syncthreads();
}

__device__ double tpp_apply_1x1(const TppCtx &c, int ne) {
    double *A = c.A;
    const size_t lda = c.lda;
    const double d = A[ne + ne * lda];
    const double inv = 1.0 / d;
    double ml = 0.0;
    for (int i = ne + 1 + c.lane; i < c.m; i += 64) {
        const double l = A[i + ne * lda] * inv;
        A[i + ne * lda] = l;
        ml = fmax(ml, fabs(l));
    }
    ml = wave_fmax(ml);
    __syncthreads();
    for (int j = ne + 1; j < c.n; ++j) {
        const double ldlj = A[j + ne * lda] * d;
        for (int i = j + c.lane; i < c.m; i += 64) A[i + j * lda] -= A[i + ne * lda] * ldlj;
    }
    __syncthreads();
    return ml;
}

__device__ double tpp_apply_2x2(const TppCtx &c, int ne) {
    double *A = c.A;
    const size_t lda = c.lda;
    const double a11 = A[ne + ne * lda], a21 = A[(ne + 1) + ne * lda], a22 = A[(ne + 1) + (ne + 1) * lda];
    const double det = a11 * a22 - a21 * a21;
    const double inv_det = 1.0 / det;
    double ml = 0.0;
    for (int i = ne + 2 + c.lane; i < c.m; i += 64) {
        const double ai1 = A[i + ne * lda], ai2 = A[i + (ne + 1) * lda];
        const double l1 = (ai1 * a22 - ai2 * a21) * inv_det;
        const double l2 = (ai2 * a11 - ai1 * a21) * inv_det;
        A[i + ne * lda] = l1;
        A[i + (ne + 1) * lda] = l2;
        ml = fmax(fmax(ml, fabs(l1)), fabs(l2));
    }
    ml = wave_fmax(ml);
    __syncthreads();
    for (int j = ne + 2; j < c.n; ++j) {
        const double wj1 = A[j + ne * lda] * a11 + A[j + (ne + 1) * lda] * a21;
        const double wj2 = A[j + ne * lda] * a21 + A[j + (ne + 1) * lda] * a22;
        for (int i = j + c.lane; i < c.m; i += 64)
            A[i + j * lda] -= A[i + ne * lda] * wj1 + A[i + (ne + 1) * lda] * wj2;
    }
    __syncthreads();
    if (c.lane == 0) A[(ne + 1) + ne * lda] = 0.0;
    __syncthreads();
    return ml;
}

__global__ __launch_bounds__(64) void k_tpp(FrontDev *__restrict__ fr, const int *__restrict__ flist,
                                            int mode, double u, double small) {
    FrontDev &F = fr[flist[blockIdx.x]];
    const int n = F.k;
    int nelim = (mode == 0) ? 0 : F.ne;
    if (nelim >= n) return;
    TppCtx c{F.F, (size_t)F.m, F.m, n, (int)threadIdx.x};
    int *perm = F.perm;
    int n1 = 0, n2 = 0;
    double maxl = 0.0;
    auto set1 = [&](int col, double v) {
        if (c.lane == 0) { F.dkind[col] = 1; F.ddiag[col] = v; F.doff[col] = 0.0; }
    };
    while (nelim < n) {
        if (tpp_col_small(c, nelim, nelim, small)) {
            set1(nelim, 0.0);
            ++n1;
            for (int r = nelim + 1 + c.lane; r < c.m; r += 64) c.A[r + nelim * c.lda] = 0.0;
            __syncthreads();
            nelim += 1;
            continue;
        }
        bool found = false;
        for (int p = nelim + 1; p < n; ++p) {
            if (tpp_col_small(c, p, nelim, small)) {
                tpp_swap(c, perm, p, nelim);
                set1(nelim, 0.0);
                ++n1;
                for (int r = nelim + 1 + c.lane; r < c.m; r += 64) c.A[r + nelim * c.lda] = 0.0;
                __syncthreads();
                nelim += 1;
                found = true;
                break;
            }
            const int t = tpp_row_abs_max(c, p, nelim, p);
            const double maxt = tpp_rc_max_excl(c, t, nelim, p);
            const double maxp = tpp_rc_max_excl(c, p, nelim, t);
            if (tpp_test_2x2(c, t, p, maxt, maxp, u, small)) {
                tpp_swap(c, perm, t, nelim);
                const int np = (p == nelim) ? t : p;
                tpp_swap(c, perm, np, nelim + 1);
                const double d11 = c.A[nelim + nelim * c.lda];
                const double d21 = c.A[(nelim + 1) + nelim * c.lda];
                const double d22 = c.A[(nelim + 1) + (nelim + 1) * c.lda];
                __syncthreads();
                const double ml = tpp_apply_2x2(c, nelim);
                if (c.lane == 0) {
                    F.dkind[nelim] = 2; F.dkind[nelim + 1] = 3;
                    F.ddiag[nelim] = d11; F.ddiag[nelim + 1] = d22;
                    F.doff[nelim] = d21; F.doff[nelim + 1] = 0.0;
                }
                ++n2;
                maxl = fmax(maxl, ml);
                nelim += 2;
                found = true;
                break;
            }
            const double maxpt = fmax(maxp, fabs(tpp_sym(c, p, t)));
            if (fabs(c.A[p + p * c.lda]) >= u * maxpt) {
                tpp_swap(c, perm, p, nelim);
                const double ml = tpp_apply_1x1(c, nelim);
                set1(nelim, c.A[nelim + nelim * c.lda]);
                ++n1;
                maxl = fmax(maxl, ml);
                nelim += 1;
                found = true;
                break;
            }
        }
        if (!found) {
            const double mp = tpp_rc_max_excl(c, nelim, nelim, -1);
            if (fabs(c.A[nelim + nelim * c.lda]) >= u * mp) {
                const double ml = tpp_apply_1x1(c, nelim);
                set1(nelim, c.A[nelim + nelim * c.lda]);
                ++n1;
                maxl = fmax(maxl, ml);
                nelim += 1;
            } else {
                break;
            }
        }
    }
    __syncthreads();
    if (c.lane == 0) {
        F.ne = nelim;
        F.n1 += n1;
        F.n2 += n2;
        atomic_max_nonneg(&F.max_l_bits, maxl);
    }
}

// ===========================================================================
// Schur update on FP64 MFMA: C -= (L_r D) L_c^T, lower part, 64x64 output tile per
// workgroup (4 waves x 32x32), K staged through LDS in chunks of 16.
//   mode 0 (contribution, compute_contribution_gemm factor.rs:1590-1652):
//           rows/cols [k, m), K = ne columns from 0.
//   mode 1 (outer-block update of the deferred trailing columns):
//           rows [obe, m), cols [obe, k), K = columns eliminated in the outer block.
// W = L D is formed while staging the A operand (1x1 and 2x2 pivots).
// ===========================================================================
typedef double d4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_schur(const FrontDev *__restrict__ fr,
                                               const AptpState *__restrict__ st,
                                               const TileItem *__restrict__ tiles, int mode) {
    const TileItem t = tiles[blockIdx.x];
    const FrontDev &F = fr[t.front];
    int r0, nr, nc, kc0, K;
    if (mode == 0) {
        r0 = F.k; nr = F.m - F.k; nc = nr; kc0 = 0; K = F.ne;
    } else {
        const AptpState &S = st[t.aux];
        r0 = S.ob_start + S.ob_cols; nr = F.m - r0; nc = F.k - r0; kc0 = S.ob_start;
        K = S.kpos - S.ob_start;
    }
    if (K <= 0 || nr <= 0 || nc <= 0) return;
    const int bi = t.ti * 64, bj = t.tj * 64;
    if (bi >= nr || bj >= nc || bi + 63 < bj) return;
    const size_t lda = F.m;
    const double *A = F.F;
    const int8_t *dk = F.dkind;
    const double *dg = F.ddiag, *of = F.doff;
    __shared__ double As[16][65];
    __shared__ double Bs[16][65];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    const int wr = wv >> 1, wc = wv & 1;
    d4 acc00 = {0, 0, 0, 0}, acc01 = {0, 0, 0, 0}, acc10 = {0, 0, 0, 0}, acc11 = {0, 0, 0, 0};
    for (int kc = 0; kc < K; kc += 16) {
        for (int e = tid; e < 1024; e += 256) {
            const int rr = e & 63, cc = e >> 6;
            const int c = kc + cc;
            double wa = 0.0, lb = 0.0;
            if (c < K) {
                const int gc = kc0 + c;
                if (bi + rr < nr) {
                    const double *rowp = A + (r0 + bi + rr);
                    wa = ld_entry(rowp, lda, gc, dk[gc], dg, of);
                }
                if (bj + rr < nc) lb = A[(r0 + bj + rr) + gc * lda];
            }
            As[cc][rr] = wa;
            Bs[cc][rr] = lb;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; kk += 4) {
            const int kr = kk + (lane >> 4);
            const double a0 = As[kr][wr * 32 + (lane & 15)];
            const double a1 = As[kr][wr * 32 + 16 + (lane & 15)];
            const double b0 = Bs[kr][wc * 32 + (lane & 15)];
            const double b1 = Bs[kr][wc * 32 + 16 + (lane & 15)];
            acc00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc00, 0, 0, 0);
            acc01 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc01, 0, 0, 0);
            acc10 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc10, 0, 0, 0);
            acc11 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc11, 0, 0, 0);
        }
        __syncthreads();
    }
    double *C = const_cast<double *>(A);
    // v_mfma_f64_16x16x4f64 result layout: lane l, element q holds row (l/16) + 4q, column l%16
    const int rbase = bi + wr * 32 + (lane >> 4);
    const int cbase = bj + wc * 32 + (lane & 15);
    d4 accs[2][2] = {{acc00, acc01}, {acc10, acc11}};
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int rl = rbase + ti * 16 + 4 * q;
                const int cl = cbase + tj * 16;
                if (rl < nr && cl < nc && rl >= cl)
                    C[(r0 + rl) + (size_t)(r0 + cl) * lda] -= accs[ti][tj][q];
            }
}

// Per-front results for the host: ne, pivot counts, delayed columns (global indices) and
// the solve index (global index of every local row after pivoting).
__global__ __launch_bounds__(64) void k_collect(FrontDev *__restrict__ fr,
                                                const int *__restrict__ flist,
                                                const long long *__restrict__ doffs,
                                                int *__restrict__ res, int *__restrict__ delayed) {
    const int li = blockIdx.x;
    const int f = flist[li];
    FrontDev &F = fr[f];
    if (threadIdx.x == 0) {
        res[li * 4 + 0] = F.ne;
        res[li * 4 + 1] = F.n1;
        res[li * 4 + 2] = F.n2;
        int z = 0;
        for (int c = 0; c < F.ne; ++c)
            if (F.dkind[c] == 1 && F.ddiag[c] == 0.0) ++z;
        res[li * 4 + 3] = z;
        ((unsigned long long *)(res + 4 * gridDim.x))[li] = F.max_l_bits;
    }
    int *dl = delayed + doffs[li];
    for (int i = threadIdx.x; i < F.m; i += blockDim.x) {
        const int g = (i < F.k) ? F.rows[F.perm[i]] : F.rows[i];
        F.sidx[i] = g;
        if (i >= F.ne && i < F.k) dl[i - F.ne] = g;
    }
}

// ===========================================================================
// Host side
// ===========================================================================
namespace {

template <class T>
size_t stage_push(std::vector<char> &st, const T *data, size_t count) {
    size_t off = (st.size() + 255) / 256 * 256;
    st.resize(off + count * sizeof(T));
    if (count) std::memcpy(st.data() + off, data, count * sizeof(T));
    return off;
}

inline int cdiv(int a, int b) { return (a + b - 1) / b; }

}  // namespace

void Prepared::build(const Symbolic &sym, const i64 *col_ptr, const i64 *row_idx, int nemin,
                     int small_leaf_threshold) {
    n = sym.n;
    nnz = col_ptr[n];
    perm_fwd = sym.perm_fwd;
    perm_inv = sym.perm_inv;
    sn_before = (long long)sym.sn.size();
    sns = amalgamate(sym.sn, nemin);
    merges = sn_before - (long long)sns.size();
    children = children_map(sns);
    int nodes = 0;
    small_leaf_subtrees = classify_small_leaf(sns, children, small_leaf_threshold, &nodes);
    small_leaf_nodes = nodes;
    const int ns = (int)sns.size();
    // height levels: a front's children are all in lower levels
    level_of.assign(ns, 0);
    int maxlv = 0;
    for (int s = 0; s < ns; ++s) {
        int lv = 0;
        for (int c : children[s]) lv = std::max(lv, level_of[c] + 1);
        level_of[s] = lv;
        maxlv = std::max(maxlv, lv);
    }
    levels.assign(ns ? maxlv + 1 : 0, {});
    for (int s = 0; s < ns; ++s) levels[level_of[s]].push_back(s);
    // amap (build_assembly_maps, numeric.rs:264-351), positions assuming no delays
    std::vector<int> g2l(n, -1);
    amap.clear();
    amap_off.assign(ns + 1, 0);
    for (int s = 0; s < ns; ++s) {
        const Supernode &S = sns[s];
        std::vector<int> rows;
        for (const ColRange &r : S.owned)
            for (int j = r.lo; j < r.hi; ++j) rows.push_back(j);
        const int total_owned = (int)rows.size();
        rows.insert(rows.end(), S.pattern.begin(), S.pattern.end());
        for (size_t i = 0; i < rows.size(); ++i) g2l[rows[i]] = (int)i;
        for (const ColRange &r : S.owned)
            for (int pj = r.lo; pj < r.hi; ++pj) {
                const int oc = perm_fwd[pj];
                const int lc = g2l[pj];
                for (i64 idx = col_ptr[oc]; idx < col_ptr[oc + 1]; ++idx) {
                    const int orow = (int)row_idx[idx];
                    if (orow < oc) {
                        const int peer = g2l[perm_inv[orow]];
                        if (peer >= 0 && peer < total_owned) continue;
                    }
                    const int gr = perm_inv[orow];
                    const int lr = g2l[gr];
                    if (lr < 0) continue;
                    AmapEntry e;
                    e.src = (unsigned)idx;
                    e.row = (unsigned)std::max(lr, lc);
                    e.col = (unsigned)std::min(lr, lc);
                    e.sr = (unsigned)gr;
                    e.sc = (unsigned)pj;
                    amap.push_back(e);
                }
            }
        for (int g : rows) g2l[g] = -1;
        amap_off[s + 1] = (i64)amap.size();
    }
    if (d_amap) { (void)hipFree(d_amap); d_amap = nullptr; }
    if (d_perm_fwd) { (void)hipFree(d_perm_fwd); d_perm_fwd = nullptr; }
    RV_HIP(hipMalloc(&d_amap, std::max<size_t>(1, amap.size()) * sizeof(AmapEntry)));
    if (!amap.empty())
        RV_HIP(hipMemcpy(d_amap, amap.data(), amap.size() * sizeof(AmapEntry), hipMemcpyHostToDevice));
    RV_HIP(hipMalloc(&d_perm_fwd, std::max<i64>(1, n) * sizeof(int)));
    if (n) RV_HIP(hipMemcpy(d_perm_fwd, perm_fwd.data(), n * sizeof(int), hipMemcpyHostToDevice));
}

Prepared::~Prepared() {
    if (d_amap) (void)hipFree(d_amap);
    if (d_perm_fwd) (void)hipFree(d_perm_fwd);
}

NumericGPU::NumericGPU(const Prepared *prep, const FactorOptions &opts, hipStream_t stream)
    : prep_(prep), o_(opts), stream_(stream) {
    const int ns = (int)prep->sns.size();
    RV_HIP(hipMalloc(&d_fr_, std::max(1, ns) * sizeof(FrontDev)));
    h_fr_.assign(ns, FrontDev());
}

NumericGPU::~NumericGPU() {
    if (d_fr_) (void)hipFree(d_fr_);
    if (pinned_) (void)hipHostFree(pinned_);
    if (d_sol_meta_) (void)hipFree(d_sol_meta_);
}

char *NumericGPU::pinned(size_t bytes) {
    if (bytes > pinned_cap_) {
        if (pinned_) (void)hipHostFree(pinned_);
        pinned_cap_ = std::max(bytes, pinned_cap_ * 2);
        RV_HIP(hipHostMalloc((void **)&pinned_, pinned_cap_, hipHostMallocDefault));
    }
    return pinned_;
}

void NumericGPU::factor(const double *d_val, const double *d_scaling) {
    const int ns = (int)prep_->sns.size();
    arena_.reset_keep();
    solve_ready_ = false;
    ne_.assign(ns, 0); m_.assign(ns, 0); k_.assign(ns, 0); n1_.assign(ns, 0); n2_.assign(ns, 0);
    zp_.assign(ns, 0);
    maxl_.assign(ns, 0.0);
    delayed_.assign(ns, {});
    rows_.assign(ns, {});
    times_ = PhaseTimes();
    for (size_t lv = 0; lv < prep_->levels.size(); ++lv) factor_level((int)lv, d_val, d_scaling);
    // aggregate statistics (numeric.rs:929-1009)
    stats_ = FactorStats();
    stats_.sn_before = prep_->sn_before;
    stats_.sn_after = ns;
    stats_.merges = prep_->merges;
    stats_.small_leaf_subtrees = prep_->small_leaf_subtrees;
    stats_.small_leaf_nodes = prep_->small_leaf_nodes;
    for (int s = 0; s < ns; ++s) {
        stats_.total_1x1 += n1_[s];
        stats_.total_2x2 += n2_[s];
        stats_.total_delayed += k_[s] - ne_[s];
        stats_.max_front_size = std::max<long long>(stats_.max_front_size, m_[s]);
        if (prep_->sns[s].parent < 0 && ne_[s] < m_[s]) stats_.zero_pivots += k_[s] - ne_[s];
        stats_.zero_pivots += zp_[s];
    }
}

void NumericGPU::factor_level(int lv, const double *d_val, const double *d_scaling) {
    const Prepared &P = *prep_;
    const std::vector<int> &fronts = P.levels[lv];
    const int nf = (int)fronts.size();
    const int nb = o_.outer_block_size, ib = o_.inner_block_size;
    // ---- host: front structure (numeric.rs:1238-1262) ----
    std::vector<int> g2l_dummy;
    size_t totF = 0, totM = 0, totK = 0;
    std::vector<int> ndel_in(nf);
    for (int li = 0; li < nf; ++li) {
        const int s = fronts[li];
        const Supernode &S = P.sns[s];
        std::vector<int> &rows = rows_[s];
        rows.clear();
        for (const ColRange &r : S.owned)
            for (int j = r.lo; j < r.hi; ++j) rows.push_back(j);
        const int nown = (int)rows.size();
        for (int c : P.children[s]) rows.insert(rows.end(), delayed_[c].begin(), delayed_[c].end());
        ndel_in[li] = (int)rows.size() - nown;
        const int k = (int)rows.size();
        rows.insert(rows.end(), S.pattern.begin(), S.pattern.end());
        m_[s] = (int)rows.size();
        k_[s] = k;
        totF += (size_t)m_[s] * m_[s];
        totM += m_[s];
        totK += std::max(k, 1);
    }
    // ---- device memory for the level ----
    double *Fbase = (double *)arena_.alloc(std::max<size_t>(totF, 1) * sizeof(double));
    int *ibase = (int *)arena_.alloc(std::max<size_t>(3 * totM, 1) * sizeof(int));
    int8_t *kbase = (int8_t *)arena_.alloc(std::max<size_t>(totK, 1));
    double *dbase = (double *)arena_.alloc(std::max<size_t>(2 * totK, 1) * sizeof(double));
    RV_HIP(hipMemsetAsync(Fbase, 0, totF * sizeof(double), stream_));
    std::vector<FrontDev> lfr(nf);
    {
        size_t fo = 0, mo = 0, ko = 0;
        for (int li = 0; li < nf; ++li) {
            const int s = fronts[li];
            FrontDev d;
            std::memset(&d, 0, sizeof(d));
            d.F = Fbase + fo;
            d.rows = ibase + mo;
            d.perm = ibase + totM + mo;
            d.sidx = ibase + 2 * totM + mo;
            d.dkind = kbase + ko;
            d.ddiag = dbase + ko;
            d.doff = dbase + totK + ko;
            d.m = m_[s];
            d.k = k_[s];
            d.n_owned = P.sns[s].n_owned();
            d.ndelay_in = ndel_in[li];
            d.ne = 0;
            const bool tpp_whole = P.sns[s].in_small_leaf || d.k < ib;
            d.flags = tpp_whole ? FRONT_TPP : 0;
            fo += (size_t)d.m * d.m;
            mo += d.m;
            ko += std::max(d.k, 1);
            lfr[li] = d;
            h_fr_[s] = d;
        }
    }
    // ---- work lists ----
    std::vector<ScatterChunk> sch;
    std::vector<int> tpp_list, aptp_list, all_list(fronts.begin(), fronts.end());
    int max_children = 0;
    for (int li = 0; li < nf; ++li) {
        const int s = fronts[li];
        for (i64 b = P.amap_off[s]; b < P.amap_off[s + 1]; b += 256)
            sch.push_back(ScatterChunk{s, (int)b, (int)std::min<i64>(b + 256, P.amap_off[s + 1]), 0});
        if (lfr[li].flags & FRONT_TPP) tpp_list.push_back(s);
        else aptp_list.push_back(s);
        max_children = std::max(max_children, (int)P.children[s].size());
    }
    // extend-add maps: contribution rows of each child -> parent local rows
    std::vector<unsigned> maps;
    struct EaPending { int child, parent, round, size; size_t mapoff; int skip; };
    std::vector<EaPending> eap;
    {
        std::vector<int> g2l;
        g2l.assign(P.n, -1);
        for (int li = 0; li < nf; ++li) {
            const int s = fronts[li];
            const std::vector<int> &prow = rows_[s];
            for (size_t i = 0; i < prow.size(); ++i) g2l[prow[i]] = (int)i;
            const std::vector<int> &ch = P.children[s];
            for (size_t ci = 0; ci < ch.size(); ++ci) {
                const int c = ch[ci];
                const int sz = m_[c] - ne_[c];
                if (sz == 0) continue;  // fully eliminated child: no contribution
                EaPending e;
                e.child = c; e.parent = s; e.round = (int)ci; e.size = sz;
                e.mapoff = maps.size();
                for (int g : delayed_[c]) maps.push_back((unsigned)g2l[g]);
                for (int g : P.sns[c].pattern) maps.push_back((unsigned)g2l[g]);
                for (size_t q = e.mapoff; q < maps.size(); ++q)
                    if ((int)maps[q] < 0 || (int)maps[q] >= (int)prow.size())
                        throw std::runtime_error("contribution row of supernode " + std::to_string(c) +
                                                 " is not in its parent's front");
                const bool g2l_path = !delayed_[c].empty() || ndel_in[li] > 0;
                e.skip = (P.sns[s].in_small_leaf || g2l_path) ? 1 : 0;
                eap.push_back(e);
            }
            for (int g : prow) g2l[g] = -1;
        }
    }
    // ---- staging ----
    std::vector<char> &stg = stage_host_;
    stg.clear();
    std::vector<int> ids(fronts.begin(), fronts.end());
    const size_t o_ids = stage_push(stg, ids.data(), ids.size());
    const size_t o_lfr = stage_push(stg, lfr.data(), lfr.size());
    std::vector<int> rowsflat;
    rowsflat.reserve(totM);
    for (int s : fronts) rowsflat.insert(rowsflat.end(), rows_[s].begin(), rows_[s].end());
    const size_t o_rows = stage_push(stg, rowsflat.data(), rowsflat.size());
    const size_t o_sch = stage_push(stg, sch.data(), sch.size());
    const size_t o_maps = stage_push(stg, maps.data(), maps.size());
    const size_t o_tpp = stage_push(stg, tpp_list.data(), tpp_list.size());
    const size_t o_aptp = stage_push(stg, aptp_list.data(), aptp_list.size());
    // extend-add items per round (need the device map address -> patched after upload)
    std::vector<std::vector<EaItem>> ea_rounds(max_children);
    for (const EaPending &e : eap) {
        const int nt = cdiv(e.size, 64);
        for (int ti = 0; ti < nt; ++ti)
            for (int tj = 0; tj <= ti; ++tj) {
                EaItem it;
                it.child = e.child; it.parent = e.parent; it.ti = ti; it.tj = tj;
                it.map = (const unsigned *)(uintptr_t)e.mapoff;  // patched below
                it.skip_zero = e.skip; it.pad = 0;
                ea_rounds[e.round].push_back(it);
            }
    }
    std::vector<size_t> o_ea(max_children);
    // device staging buffer
    const size_t ea_bytes_est = [&] {
        size_t b = 0;
        for (auto &r : ea_rounds) b += r.size() * sizeof(EaItem) + 256;
        return b;
    }();
    // APTP item lists
    const int na = (int)aptp_list.size();
    std::vector<int2> k2items;
    std::vector<int4> k3items;
    int max_outer = 0;
    for (int li = 0; li < na; ++li) {
        const int s = aptp_list[li];
        const int m = m_[s], p = k_[s];
        for (int c = 0; c < cdiv(m, 256); ++c) k2items.push_back(make_int2(li, c));
        const int qmax = 1 + cdiv(std::min(p, nb), 32);
        for (int rt = 0; rt < cdiv(m, 64); ++rt)
            for (int q = 0; q < qmax; ++q) k3items.push_back(make_int4(li, rt, q, 0));
        const int nouter = p > nb ? cdiv(p, nb) : 1;
        max_outer = std::max(max_outer, nouter);
    }
    const size_t o_k2 = stage_push(stg, k2items.data(), k2items.size());
    const size_t o_k3 = stage_push(stg, k3items.data(), k3items.size());
    // contribution GEMM tiles (mode 0)
    std::vector<TileItem> ctiles;
    for (int li = 0; li < nf; ++li) {
        const int s = fronts[li];
        if (P.sns[s].parent < 0) continue;
        const int nfs = m_[s] - k_[s];
        if (nfs <= 0) continue;
        const int nt = cdiv(nfs, 64);
        for (int ti = 0; ti < nt; ++ti)
            for (int tj = 0; tj <= ti; ++tj) ctiles.push_back(TileItem{s, ti, tj, li});
    }
    const size_t o_ct = stage_push(stg, ctiles.data(), ctiles.size());
    // outer-update tiles (mode 1) per outer block index
    std::vector<std::vector<TileItem>> otiles(max_outer);
    for (int li = 0; li < na; ++li) {
        const int s = aptp_list[li];
        const int m = m_[s], p = k_[s];
        if (p <= nb) continue;
        const int nouter = cdiv(p, nb);
        // every outer block: columns beyond it hold later blocks and earlier blocks' delays
        for (int ob = 0; ob < nouter; ++ob) {
            const int obcols = std::min(nb, p - ob * nb);
            const int nr = m - obcols, nc = p - obcols;
            if (nc <= 0) continue;
            const int ntr = cdiv(nr, 64), ntc = cdiv(nc, 64);
            for (int ti = 0; ti < ntr; ++ti)
                for (int tj = 0; tj <= std::min(ti, ntc - 1); ++tj)
                    otiles[ob].push_back(TileItem{s, ti, tj, li});
        }
    }
    std::vector<size_t> o_ot(max_outer);
    for (int ob = 0; ob < max_outer; ++ob) o_ot[ob] = stage_push(stg, otiles[ob].data(), otiles[ob].size());
    std::vector<long long> doffs(nf + 1, 0);
    for (int li = 0; li < nf; ++li) doffs[li + 1] = doffs[li] + std::max(k_[fronts[li]], 1);
    const size_t o_doffs = stage_push(stg, doffs.data(), doffs.size());
    // extend-add items last: patch map pointers relative to the device base
    size_t stage_total_est = stg.size() + ea_bytes_est + 4096;
    char *dstage = (char *)stage_buf_.get(stage_total_est);
    for (int r = 0; r < max_children; ++r) {
        for (EaItem &it : ea_rounds[r])
            it.map = (const unsigned *)(dstage + o_maps + (size_t)(uintptr_t)it.map * sizeof(unsigned));
        o_ea[r] = stage_push(stg, ea_rounds[r].data(), ea_rounds[r].size());
    }
    if (stg.size() > stage_total_est) {
        // should not happen (estimate covers all pushes); fall back to a bigger buffer
        throw std::runtime_error("staging estimate exceeded");
    }
    const size_t rbytes = (size_t)nf * 4 * sizeof(int) + (size_t)nf * 8;
    const size_t dbytes = (size_t)doffs[nf] * sizeof(int);
    const size_t tot = (rbytes + 255) / 256 * 256 + dbytes;
    char *hp = pinned(std::max(stg.size(), tot));
    std::memcpy(hp, stg.data(), stg.size());
    RV_HIP(hipMemcpyAsync(dstage, hp, stg.size(), hipMemcpyHostToDevice, stream_));
    auto D = [&](size_t off) { return dstage + off; };
    // ---- per-front descriptors + rows/perm ----
    hipLaunchKernelGGL(k_set_fronts, dim3(cdiv(nf, 256)), dim3(256), 0, stream_, d_fr_,
                       (const int *)D(o_ids), (const FrontDev *)D(o_lfr), nf);
    RV_HIP(hipMemcpyAsync(ibase, D(o_rows), totM * sizeof(int), hipMemcpyDeviceToDevice, stream_));
    hipLaunchKernelGGL(k_init_perm, dim3(nf), dim3(256), 0, stream_, d_fr_, (const int *)D(o_ids), nf);
    // ---- assembly ----
    if (!sch.empty())
        hipLaunchKernelGGL(k_scatter_orig, dim3((unsigned)sch.size()), dim3(256), 0, stream_,
                           d_fr_, (const ScatterChunk *)D(o_sch), P.d_amap, d_val, d_scaling);
    for (int r = 0; r < max_children; ++r)
        if (!ea_rounds[r].empty())
            hipLaunchKernelGGL(k_extend_add, dim3((unsigned)ea_rounds[r].size()), dim3(256), 0,
                               stream_, d_fr_, (const EaItem *)D(o_ea[r]));
    // ---- TPP fronts ----
    if (!tpp_list.empty())
        hipLaunchKernelGGL(k_tpp, dim3((unsigned)tpp_list.size()), dim3(64), 0, stream_, d_fr_,
                           (const int *)D(o_tpp), 0, o_.threshold, o_.small);
    // ---- APTP fronts ----
    if (na > 0) {
        size_t sbytes = (size_t)na * sizeof(AptpState);
        size_t wbytes = 0;
        for (int s : aptp_list)
            wbytes += (size_t)m_[s] * ib * sizeof(double) + ((size_t)2 * MAX_IB * m_[s]) * sizeof(double) + 512;
        size_t markbytes = 0;
        for (int s : aptp_list) markbytes += (size_t)m_[s] * sizeof(int) + 256;
        char *wb = (char *)work_buf_.get(sbytes + wbytes + markbytes + 4096);
        std::vector<AptpState> hs(na);
        size_t woff = (sbytes + 255) / 256 * 256;
        for (int li = 0; li < na; ++li) {
            AptpState &S = hs[li];
            std::memset(&S, 0, sizeof(S));
            const int m = m_[aptp_list[li]];
            S.backup = (double *)(wb + woff);
            woff += (size_t)m * ib * sizeof(double);
            woff = (woff + 255) / 256 * 256;
            S.swapbuf = (double *)(wb + woff);
            woff += (size_t)2 * MAX_IB * m * sizeof(double);
            woff = (woff + 255) / 256 * 256;
        }
        const size_t mark0 = woff;
        for (int li = 0; li < na; ++li) {
            hs[li].mark = (int *)(wb + woff);
            woff += (size_t)m_[aptp_list[li]] * sizeof(int);
            woff = (woff + 255) / 256 * 256;
        }
        hs_stage_.assign((char *)hs.data(), (char *)hs.data() + sbytes);
        RV_HIP(hipMemcpyAsync(wb, hs_stage_.data(), sbytes, hipMemcpyHostToDevice, stream_));
        RV_HIP(hipMemsetAsync(wb + mark0, 0, woff - mark0, stream_));
        AptpState *dS = (AptpState *)wb;
        const int *dl = (const int *)D(o_aptp);
        for (int ob = 0; ob < max_outer; ++ob) {
            hipLaunchKernelGGL(k_outer_begin, dim3(cdiv(na, 64)), dim3(64), 0, stream_, d_fr_, dS, dl,
                               na, ob, nb);
            int ninner = 0;
            for (int s : aptp_list) {
                const int p = k_[s];
                const bool two = p > nb;
                const int nouter = two ? cdiv(p, nb) : 1;
                if (ob >= nouter) continue;
                const int obcols = two ? std::min(nb, p - ob * nb) : p;
                ninner = std::max(ninner, cdiv(obcols, ib));
            }
            for (int it = 0; it < ninner; ++it) {
                hipLaunchKernelGGL(k_block_factor, dim3(na), dim3(64), 0, stream_, d_fr_, dS, dl, ob,
                                   it, nb, ib, o_.small);
                hipLaunchKernelGGL(k_panel_apply, dim3((unsigned)k2items.size()), dim3(256), 0, stream_,
                                   d_fr_, dS, dl, (const int2 *)D(o_k2), ob, it, nb, ib, o_.threshold);
                hipLaunchKernelGGL(k_panel_update, dim3((unsigned)k3items.size()), dim3(256), 0,
                                   stream_, d_fr_, dS, dl, (const int4 *)D(o_k3), ob, it, nb, ib);
                hipLaunchKernelGGL(k_delay_swap, dim3(na), dim3(256), 0, stream_, d_fr_, dS, dl, ob,
                                   it, nb, ib);
            }
            if (!otiles[ob].empty())
                hipLaunchKernelGGL(k_schur, dim3((unsigned)otiles[ob].size()), dim3(256), 0, stream_,
                                   d_fr_, dS, (const TileItem *)D(o_ot[ob]), 1);
            hipLaunchKernelGGL(k_outer_end, dim3(na), dim3(256), 0, stream_, d_fr_, dS, dl, ob, nb);
        }
        if (o_.failed_pivot_tpp)
            hipLaunchKernelGGL(k_tpp, dim3(na), dim3(64), 0, stream_, d_fr_, dl, 1, o_.threshold,
                               o_.small);
    }
    // ---- contribution blocks ----
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (!ctiles.empty()) {
        if (timing) {
            RV_HIP(hipEventCreate(&ev0));
            RV_HIP(hipEventCreate(&ev1));
            RV_HIP(hipEventRecord(ev0, stream_));
        }
        hipLaunchKernelGGL(k_schur, dim3((unsigned)ctiles.size()), dim3(256), 0, stream_, d_fr_,
                           (const AptpState *)nullptr, (const TileItem *)D(o_ct), 0);
        if (timing) RV_HIP(hipEventRecord(ev1, stream_));
    }
    // ---- collect ----
    char *rb = (char *)res_buf_.get(rbytes + dbytes + 512);
    int *dres = (int *)rb;
    int *ddel = (int *)(rb + (rbytes + 255) / 256 * 256);
    hipLaunchKernelGGL(k_collect, dim3(nf), dim3(64), 0, stream_, d_fr_, (const int *)D(o_ids),
                       (const long long *)D(o_doffs), dres, ddel);
    char *hres = hp;
    RV_HIP(hipMemcpyAsync(hres, rb, tot, hipMemcpyDeviceToHost, stream_));
    RV_HIP(hipStreamSynchronize(stream_));
    const int *hr = (const int *)hres;
    const unsigned long long *hml = (const unsigned long long *)(hres + (size_t)nf * 4 * sizeof(int));
    const int *hdel = (const int *)(hres + (rbytes + 255) / 256 * 256);
    for (int li = 0; li < nf; ++li) {
        const int s = fronts[li];
        ne_[s] = hr[li * 4 + 0];
        n1_[s] = hr[li * 4 + 1];
        n2_[s] = hr[li * 4 + 2];
        zp_[s] = hr[li * 4 + 3];
        unsigned long long bits = hml[li];
        double v;
        std::memcpy(&v, &bits, sizeof(v));
        maxl_[s] = v;
        h_fr_[s].ne = ne_[s];
        delayed_[s].assign(hdel + doffs[li], hdel + doffs[li] + (k_[s] - ne_[s]));
    }
    if (ev0) {
        float ms = 0.f;
        RV_HIP(hipEventElapsedTime(&ms, ev0, ev1));
        times_.schur_ms += ms;
        times_.schur_launches += 1;
        double fl = 0;
        for (const TileItem &t : ctiles) (void)t;
        for (int li = 0; li < nf; ++li) {
            const int s = fronts[li];
            if (P.sns[s].parent < 0) continue;
            const double nfs = m_[s] - k_[s];
            fl += (double)ne_[s] * nfs * (nfs + 1.0);
        }
        times_.schur_flops += fl;
        (void)hipEventDestroy(ev0);
        (void)hipEventDestroy(ev1);
    }
}

void NumericGPU::export_raw(int s, double *Fh, int *perm, int *rows) const {
    const FrontDev &F = h_fr_[s];
    const int m = F.m;
    if (Fh) RV_HIP(hipMemcpy(Fh, F.F, (size_t)m * m * sizeof(double), hipMemcpyDeviceToHost));
    if (perm) RV_HIP(hipMemcpy(perm, F.perm, m * sizeof(int), hipMemcpyDeviceToHost));
    if (rows) RV_HIP(hipMemcpy(rows, F.rows, m * sizeof(int), hipMemcpyDeviceToHost));
}

void NumericGPU::front_d(int s, signed char *kind, double *diag, double *off) const {
    const FrontDev &F = h_fr_[s];
    const int ne = ne_[s];
    if (ne == 0) return;
    std::vector<int8_t> kd(ne);
    RV_HIP(hipMemcpy(kd.data(), F.dkind, ne, hipMemcpyDeviceToHost));
    if (diag) RV_HIP(hipMemcpy(diag, F.ddiag, ne * sizeof(double), hipMemcpyDeviceToHost));
    if (off) RV_HIP(hipMemcpy(off, F.doff, ne * sizeof(double), hipMemcpyDeviceToHost));
    for (int c = 0; c < ne; ++c) {
        if (kind) kind[c] = (signed char)(kd[c] == 3 ? 2 : kd[c]);
        if (off && kd[c] != 2) off[c] = 0.0;
    }
}

void NumericGPU::supernode_info(int s, long long out[8], double *max_l) const {
    out[0] = m_[s]; out[1] = k_[s]; out[2] = ne_[s]; out[3] = k_[s] - ne_[s];
    out[4] = n1_[s]; out[5] = n2_[s]; out[6] = prep_->sns[s].parent;
    out[7] = prep_->sns[s].in_small_leaf ? 1 : 0;
    if (max_l) *max_l = maxl_[s];
}

// inertia from the D factors (diagonal.rs:391-461), summed over fronts (solver.rs:504-519)
void NumericGPU::inertia(long long out[3]) const {
    out[0] = out[1] = out[2] = 0;
    const int ns = (int)prep_->sns.size();
    std::vector<int8_t> kind;
    std::vector<double> dg, of;
    for (int s = 0; s < ns; ++s) {
        const int ne = ne_[s];
        if (ne == 0) continue;
        kind.resize(ne); dg.resize(ne); of.resize(ne);
        RV_HIP(hipMemcpy(kind.data(), h_fr_[s].dkind, ne, hipMemcpyDeviceToHost));
        RV_HIP(hipMemcpy(dg.data(), h_fr_[s].ddiag, ne * sizeof(double), hipMemcpyDeviceToHost));
        RV_HIP(hipMemcpy(of.data(), h_fr_[s].doff, ne * sizeof(double), hipMemcpyDeviceToHost));
        for (int c = 0; c < ne; ++c) {
            if (kind[c] == 1) {
                if (dg[c] > 0.0) out[0]++; else if (dg[c] < 0.0) out[1]++; else out[2]++;
            } else if (kind[c] == 2) {
                const double a = dg[c], b = of[c], cc = dg[c + 1];
                const double det = a * cc - b * b, tr = a + cc;
                if (det > 0.0) { if (tr > 0.0) out[0] += 2; else out[1] += 2; }
                else if (det < 0.0) { out[0]++; out[1]++; }
                else {
                    if (tr > 0.0) { out[0]++; out[2]++; }
                    else if (tr < 0.0) { out[1]++; out[2]++; }
                    else out[2] += 2;
                }
            }
        }
    }
}

void NumericGPU::export_front(int s, double *l11, double *l21, signed char *dkind, double *ddiag,
                              double *doff, int *col_indices, int *row_indices,
                              int *local_perm) const {
    const FrontDev &F = h_fr_[s];
    const int m = F.m, k = F.k, ne = ne_[s];
    std::vector<double> A((size_t)m * m);
    RV_HIP(hipMemcpy(A.data(), F.F, A.size() * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<int> perm(m), rows(m);
    RV_HIP(hipMemcpy(perm.data(), F.perm, m * sizeof(int), hipMemcpyDeviceToHost));
    RV_HIP(hipMemcpy(rows.data(), F.rows, m * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<int8_t> kind(std::max(ne, 1));
    std::vector<double> dg(std::max(ne, 1)), of(std::max(ne, 1));
    if (ne) {
        RV_HIP(hipMemcpy(kind.data(), F.dkind, ne, hipMemcpyDeviceToHost));
        RV_HIP(hipMemcpy(dg.data(), F.ddiag, ne * sizeof(double), hipMemcpyDeviceToHost));
        RV_HIP(hipMemcpy(of.data(), F.doff, ne * sizeof(double), hipMemcpyDeviceToHost));
    }
    // extract_front_factors (numeric.rs:2579-2713)
    if (l11) {
        for (int j = 0; j < ne; ++j)
            for (int i = 0; i < ne; ++i) l11[i + (size_t)j * ne] = 0.0;
        int col = 0;
        while (col < ne) {
            l11[col + (size_t)col * ne] = 1.0;
            if (kind[col] == 1) {
                for (int i = col + 1; i < ne; ++i) l11[i + (size_t)col * ne] = A[i + (size_t)col * m];
                col += 1;
            } else {
                l11[(col + 1) + (size_t)(col + 1) * ne] = 1.0;
                for (int i = col + 2; i < ne; ++i) {
                    l11[i + (size_t)col * ne] = A[i + (size_t)col * m];
                    l11[i + (size_t)(col + 1) * ne] = A[i + (size_t)(col + 1) * m];
                }
                col += 2;
            }
        }
    }
    const int r = m - ne;
    if (l21)
        for (int j = 0; j < ne; ++j)
            for (int i = 0; i < r; ++i) l21[i + (size_t)j * r] = A[(ne + i) + (size_t)j * m];
    for (int c = 0; c < ne; ++c) {
        if (dkind) dkind[c] = (signed char)(kind[c] == 3 ? 2 : kind[c]);
        if (ddiag) ddiag[c] = dg[c];
        if (doff) doff[c] = (kind[c] == 2) ? of[c] : 0.0;
        if (col_indices) col_indices[c] = rows[perm[c]];
    }
    if (row_indices) {
        int q = 0;
        for (int i = ne; i < k; ++i) row_indices[q++] = rows[perm[i]];
        for (int i = k; i < m; ++i) row_indices[q++] = rows[i];
    }
    if (local_perm)
        for (int i = 0; i < k; ++i) local_perm[i] = perm[i];
}

}  // namespace rv
