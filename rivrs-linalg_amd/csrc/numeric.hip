// numeric.hip -- MI355X (gfx950) numeric factorization of the multifrontal APTP LDL^T.
//
// Fronts are processed level by level of the assembly tree (leaves first), all fronts of a
// level in the same launches.  Per level:
//   1. assembly      k_scatter_orig (original entries) + k_extend_add (one launch per child
//                    ordinal, so the children of a parent are summed in reference order)
//   2. dense factor  fronts with k < ib or inside a small-leaf subtree: k_tpp (threshold
//                    partial pivoting, one wavefront per front);
//                    other fronts: APTP with ib x ib complete pivoting blocks
//                    (k_block_apply -> k_block_update -> k_block_finish per inner block, the
//                    trailing update restricted to the outer block, and the rank-nb outer
//                    update k_form_ld + k_schur on MFMA), then k_tpp on the columns APTP left.
//   3. contribution  k_form_ld + k_schur: C -= (L21 D) L21^T on the NFS x NFS block, FP64 MFMA.
//   4. k_collect     eliminated counts, D inertia and delayed columns for the host.
// The host plans every level from the children's results the previous factorization
// produced and enqueues all levels without waiting; levels above one whose delays differ
// from that plan are planned again and rerun (NumericGPU::factor).
// The scalar pivoting arithmetic is compiled without FMA contraction so that every decision
// step reproduces the reference's operation sequence (sparse/src/symmetric/factor.rs).
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "rv_numeric.h"

namespace rv {

// ===========================================================================
// Arena / buffers
// ===========================================================================
void *DevArena::alloc(size_t bytes, size_t align) {
    bytes = (bytes + align - 1) / align * align;
    if (bytes == 0) bytes = align;
    while (cur_ < chunks_.size()) {
        Chunk &c = chunks_[cur_];
        const size_t off = (c.used + align - 1) / align * align;
        if (off + bytes <= c.size) {
            c.used = off + bytes;
            return c.base + off;
        }
        if (cur_ + 1 >= chunks_.size()) break;
        chunks_[++cur_].used = 0;
    }
    const size_t sz = std::max(bytes, (size_t)256 << 20);
    void *p = nullptr;
    RV_HIP(hipMalloc(&p, sz));
    chunks_.push_back(Chunk{(char *)p, sz, bytes});
    cur_ = chunks_.size() - 1;
    reserved_ += sz;
    return p;
}

void DevArena::reset_keep() {
    cur_ = 0;
    if (!chunks_.empty()) chunks_[0].used = 0;
}

DevArena::Mark DevArena::mark() const {
    Mark m;
    m.chunk = cur_;
    m.used = cur_ < chunks_.size() ? chunks_[cur_].used : 0;
    return m;
}

void DevArena::reset_to(const Mark &m) {
    cur_ = m.chunk;
    if (cur_ < chunks_.size()) chunks_[cur_].used = m.used;
}

void DevArena::release() {
    for (Chunk &c : chunks_) (void)hipFree(c.base);
    chunks_.clear();
    cur_ = 0;
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

void FrontPool::reserve(size_t bytes) {
    if (virtual_ || !chunks_.empty() || bytes == 0) return;
    bytes = (bytes + 255) / 256 * 256;
    void *p = nullptr;
    RV_HIP(hipMalloc(&p, bytes));
    chunks_.push_back(Chunk{(char *)p, bytes});
    reserved_ += bytes;
}

void FrontPool::drop_from(int lv) {
    size_t n = 0;
    for (const Block &b : blocks_)
        if (b.lv < lv) blocks_[n++] = b;
    blocks_.resize(n);
}

int FrontPool::place(int lv, int death, size_t bytes) {
    bytes = std::max<size_t>((bytes + 255) / 256 * 256, 256);
    if (virtual_ && chunks_.empty()) chunks_.push_back(Chunk{nullptr, SIZE_MAX});
    std::vector<std::pair<size_t, size_t>> iv;
    for (size_t c = 0; c <= chunks_.size(); ++c) {
        if (c == chunks_.size()) {
            // no chunk has room: a new one (its blocks are placed from its start)
            const size_t sz = std::max(bytes, (size_t)256 << 20);
            void *p = nullptr;
            RV_HIP(hipMalloc(&p, sz));
            chunks_.push_back(Chunk{(char *)p, sz});
            reserved_ += sz;
        }
        iv.clear();
        for (const Block &b : blocks_)
            if (b.chunk == (int)c && (b.lv == lv || (b.lv < lv && b.death >= lv)))
                iv.emplace_back(b.off, b.off + b.size);
        std::sort(iv.begin(), iv.end());
        size_t pos = 0;
        for (const auto &r : iv) {
            if (r.first >= pos + bytes) break;
            pos = std::max(pos, r.second);
        }
        if (pos + bytes > chunks_[c].size) continue;
        for (Block &b : blocks_)
            if (b.chunk == (int)c && b.lv < lv && b.death < lv && b.off < pos + bytes && pos < b.off + b.size)
                b.clobbered = true;
        blocks_.push_back(Block{(int)c, pos, bytes, lv, death, false});
        top_ = std::max(top_, pos + bytes);
        return (int)blocks_.size() - 1;
    }
    return -1;  // not reached
}

void FrontPool::release() {
    if (!virtual_)
        for (Chunk &c : chunks_) (void)hipFree(c.base);
    chunks_.clear();
    blocks_.clear();
    top_ = 0;
    reserved_ = 0;
}

// Phase timing of the inner-block kernels for a probe build (-DRV_PHASE_TIMING): workgroup 0,
// thread 0 of every launch adds the cycles since kernel start at each phase mark.
#ifdef RV_PHASE_TIMING
__device__ unsigned long long g_phase[2][16];
__device__ unsigned long long g_phase_launches[2];
#define RV_PHASE_START() const unsigned long long rv_t0_ = clock64()
#define RV_PHASE(kern, i)                                                              \
    do {                                                                               \
        if (blockIdx.x == 0 && threadIdx.x == 0)                                       \
            atomicAdd(&g_phase[kern][i], (unsigned long long)(clock64() - rv_t0_));     \
    } while (0)
#define RV_PHASE_COUNT(kern)                                                           \
    do {                                                                               \
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&g_phase_launches[kern], 1ull); \
    } while (0)
#else
#define RV_PHASE_START() do {} while (0)
#define RV_PHASE(kern, i) do {} while (0)
#define RV_PHASE_COUNT(kern) do {} while (0)
#endif

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
template <class P>
__device__ __forceinline__ double ld_entry(P rowbase, size_t lda, int c, int kind, const double *diag,
                                           const double *off) {
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

// Fronts are zeroed in their lower triangle only (diagonal included): every kernel reads and
// writes only that part of a front.  Work item (front, ZCOLS columns), one column at a time.
constexpr int ZCOLS = 16;

// A level's set-up in one launch (the parts are independent; descriptors are read from the
// level's staged copy lfr): workgroups [0, nzl) zero the fronts' lower triangles (ZCOLS columns
// each), the next cdiv(nf, 256) install the descriptors, the next nf set perm = identity, the
// rest copy the level's row lists (1024 entries each).
constexpr int ROWCOPY = 1024;
__global__ __launch_bounds__(256) void k_level_setup(FrontDev *__restrict__ fr, const int *__restrict__ ids,
                                                     const FrontDev *__restrict__ lfr, int nf,
                                                     const int2 *__restrict__ zitems, int nzl,
                                                     int *__restrict__ rows_dst, const int *__restrict__ rows_src,
                                                     long long nrows) {
    int b = blockIdx.x;
    const int tid = threadIdx.x;
    if (b < nzl) {
        const int2 it = zitems[b];
        const FrontDev &F = lfr[it.x];
        const int m = F.m;
        gdouble *A = gbl(F.F);
        const int c1 = min(it.y + ZCOLS, m);
        for (int c = it.y; c < c1; ++c)
            for (int r = c + tid; r < m; r += 256) A[r + (size_t)c * m] = 0.0;
        return;
    }
    b -= nzl;
    const int nset = (nf + 255) / 256;
    if (b < nset) {
        const int i = b * 256 + tid;
        if (i < nf) fr[ids[i]] = lfr[i];
        return;
    }
    b -= nset;
    if (b < nf) {
        const FrontDev &F = lfr[b];
        for (int i = tid; i < F.m; i += 256) gbl(F.perm)[i] = i;
        return;
    }
    b -= nf;
    const long long i0 = (long long)b * ROWCOPY;
    for (int u = tid; u < ROWCOPY; u += 256)
        if (i0 + u < nrows) rows_dst[i0 + u] = rows_src[i0 + u];
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
    gbl(F.F)[r + (size_t)col * F.m] += v;
}

// Extend-add of one 64x64 lower tile of a child's contribution block (the trailing
// (m-ne)^2 block of the child front) into the parent, with both index maps staged in LDS
// (extend_add numeric.rs:2445-2485, extend_add_mapped 2507-2549).
__global__ __launch_bounds__(256) void k_extend_add(const FrontDev *__restrict__ fr,
                                                    const EaItem *__restrict__ items) {
    const EaItem it = items[blockIdx.x];
    const FrontDev &C = fr[it.child];
    const FrontDev &P = fr[it.parent];
    const int ne = it.cne;  // the child's eliminated count the parent's map was built for
    const int sz = C.m - ne;
    const size_t ldc = C.m;
    const gdouble *CB = gbl(C.F) + ne + (size_t)ne * ldc;
    const int i0 = it.ti * 64, j0 = it.tj * 64;
    __shared__ unsigned mi[64], mj[64];
    const int t = threadIdx.x;
    if (t < 64) mi[t] = (i0 + t < sz) ? gbl(it.map)[i0 + t] : 0u;
    else if (t < 128) mj[t - 64] = (j0 + t - 64 < sz) ? gbl(it.map)[j0 + t - 64] : 0u;
    __syncthreads();
    const int ii = t & 63;
    const int i = i0 + ii;
    if (i >= sz) return;
    gdouble *PF = gbl(P.F);
    const size_t ldp = P.m;
    // all loads first (the 16 targets of a thread are distinct), then the stores
    double v[16], pv[16];
    size_t dst[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int jj = (t >> 6) + 4 * u;
        const int j = j0 + jj;
        const bool ok = j < sz && i >= j;
        v[u] = ok ? CB[i + (size_t)j * ldc] : 0.0;
        const unsigned li = mi[ii], lj = mj[jj];
        dst[u] = (li >= lj) ? li + (size_t)lj * ldp : lj + (size_t)li * ldp;
        pv[u] = ok ? PF[dst[u]] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int j = j0 + (t >> 6) + 4 * u;
        if (j >= sz || i < j) continue;
        if (it.skip_zero && v[u] == 0.0) continue;
        PF[dst[u]] = pv[u] + v[u];
    }
}

// ===========================================================================
// Wavefront primitives.  DPP moves within rows of 16 lanes plus four readlanes reduce a
// wavefront without the LDS crossbar; callers keep all 64 lanes active.
// ===========================================================================
// Workgroup barrier that orders LDS only: global loads and stores in flight are not waited for.
__device__ __forceinline__ void lds_barrier() {
#ifdef RV_DBG_FULLBAR
    __syncthreads();
#else
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
#endif
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const long long b = __double_as_longlong(v);
    // every lane is written (full row and bank masks), so no old value is needed
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ double readlane_f64(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// fmax over the wavefront (NaN inputs are ignored, as by fmax)
__device__ __forceinline__ double wave_max_dpp(double v) {
    v = fmax(v, dpp_f64<0xB1>(v));   // quad_perm [1,0,3,2]
    v = fmax(v, dpp_f64<0x4E>(v));   // quad_perm [2,3,0,1]
    v = fmax(v, dpp_f64<0x141>(v));  // row_half_mirror
    v = fmax(v, dpp_f64<0x140>(v));  // row_mirror
    return fmax(fmax(readlane_f64(v, 0), readlane_f64(v, 16)),
                fmax(readlane_f64(v, 32), readlane_f64(v, 48)));
}

// Position of the wavefront maximum: larger value wins, equal values -> smaller position.
// vmax is the wavefront maximum of v; lanes without a candidate hold pos == INT_MAX.
__device__ __forceinline__ int wave_argmax_pos(double v, int pos, double vmax) {
    const unsigned long long mask = __ballot(pos != INT_MAX && v == vmax);
    if (mask == 0) return INT_MAX;
    if ((mask & (mask - 1)) == 0) return __builtin_amdgcn_readlane(pos, __ffsll((long long)mask) - 1);
    return wave_min_int((v == vmax) ? pos : INT_MAX);
}

// Copy of the lower diagonal block [kpos, kpos + bs) of the next inner block (all threads of
// the calling workgroup).  Every workgroup of k_block_apply stages the block from this copy,
// because the first of them writes the factored block back into the front.
__device__ __forceinline__ void copy_next_block(const FrontDev &F, const AptpState &S, int ib) {
    const int k = S.kpos, bs = min(S.end_pos - k, ib);
    const size_t lda = F.m;
    for (int e = threadIdx.x; e < MAX_IB * MAX_IB; e += blockDim.x) {
        const int i = e & (MAX_IB - 1), j = e / MAX_IB;
        if (i < bs && j <= i) gbl(S.blkcopy)[e] = gbl(F.F)[(k + i) + (size_t)(k + j) * lda];
    }
}

// ===========================================================================
// APTP: complete pivoting on the inner block (factor_block_diagonal, factor.rs:891-1153).
// Returns the eliminated count; bp/bkind/bdiag/boff describe the pivots and *maxl the largest
// |l| of the block.
// ===========================================================================
// The same complete pivoting by a whole workgroup of 256 threads: thread (row r = tid & 31,
// column group cg = tid >> 5) owns the entries j = cg + 8u of row r and keeps them in registers.
// One barrier per pivot step: scan of the own entries, argmax over the workgroup (DPP within
// wavefronts, per-wavefront partials with the winner's signed value through LDS), barrier, then
// the pivot decision and the step's symmetric swaps are folded into where each thread reads its
// new entries: the block lives in two LDS buffers, each step reads the current one at the
// swapped positions (new(r, c) = old(s(r), s(c)), s the step's swap sequence) and writes the
// updated entries to the other.  Partials, diagonals and the row order are double-buffered the
// same way.  Multipliers are kept by row identity in lbuf and placed into the block at the end.
__device__ int block_factor_wg(double (*blk)[MAX_IB + 1], double (*blk2)[MAX_IB + 1], int *bp, int *bp2,
                               int8_t *bkind, double *bdiag, double *boff, int bs, double small,
                               double *maxl, double *wbv, int *wbp, double *wbs,
                               double (*lbuf)[MAX_IB + 1], bool probe = false) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int r = tid & 31, cg = tid >> 5;
#ifdef RV_PHASE_TIMING
    unsigned long long tp[6] = {0, 0, 0, 0, 0, 0}, tq = clock64();
#define RV_STEP_MARK(i) do { if (probe) { const unsigned long long tn_ = clock64(); tp[i] += tn_ - tq; tq = tn_; } } while (0)
#else
#define RV_STEP_MARK(i) do {} while (0)
#endif
    if (tid < MAX_IB) { bp[tid] = tid; bkind[tid] = 0; bdiag[tid] = 0.0; boff[tid] = 0.0; }
    double e[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) e[u] = blk[r][cg + 8 * u];
    // the diagonal of the row at position r, kept by every thread of the row (recomputed each
    // step with the arithmetic of the thread that owns the entry); the diagonal at column j is
    // then held by the lane of row j with this thread's column group, in the same wavefront
    double dreg = blk[r][r];
    double (*X)[MAX_IB + 1] = blk;
    double (*Y)[MAX_IB + 1] = blk2;
    int *bpX = bp, *bpY = bp2;
    int kk = 0, par = 0;
    double ml = 0.0;
    while (kk < bs) {
        const int cur = kk;
        // argmax |a_ij| over the remaining lower block, first in column-major scan order
        double bv = 0.0, sv = 0.0;
        int bpos = INT_MAX;
        {
            const bool row_live = r >= cur && r < bs;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int j = cg + 8 * u;
                const double a = fabs(e[u]);
                const bool take = row_live && j >= cur && j <= r && a > bv;
                bv = take ? a : bv;
                sv = take ? e[u] : sv;
                bpos = take ? j * 64 + r : bpos;
            }
        }
        // per-wavefront partials: value, position, and the winner's signed value and the two
        // diagonals of its 2x2 test
        double *pv_ = wbv + 4 * par, *ps_ = wbs + 4 * par, *pr_ = wbs + 8 + 4 * par, *pc_ = wbs + 16 + 4 * par;
        int *pp_ = wbp + 4 * par;
        const int jb = (bpos == INT_MAX) ? r : (bpos >> 6);
        const double dc_ = __shfl(dreg, jb + 32 * (cg & 1));
        RV_STEP_MARK(0);
        {
            const double vmax = wave_max_dpp(bv);
            const int pos = wave_argmax_pos(bv, bpos, vmax);
            if (pos != INT_MAX && bpos == pos) { ps_[wv] = sv; pr_[wv] = dreg; pc_[wv] = dc_; }
            if (lane == 0) { pv_[wv] = vmax; pp_[wv] = pos; }
        }
        RV_STEP_MARK(1);
        lds_barrier();
        RV_STEP_MARK(2);
        // all partials in one round trip, then the combination in registers
        double qv[4], qs[4], qr[4], qc[4];
        int qp[4];
#pragma unroll
        for (int w = 0; w < 4; ++w) { qv[w] = pv_[w]; qp[w] = pp_[w]; qs[w] = ps_[w]; qr[w] = pr_[w]; qc[w] = pc_[w]; }
        // larger value wins, an equal value with the smaller position wins: a strict order over
        // the partials' distinct positions, combined as a tree (two levels instead of three)
        auto wins = [](double va, int pa, double vb, int pb) { return vb > va || (vb == va && pb < pa); };
        const int w01 = wins(qv[0], qp[0], qv[1], qp[1]) ? 1 : 0;
        const int w23 = wins(qv[2], qp[2], qv[3], qp[3]) ? 3 : 2;
        const double v01 = w01 ? qv[1] : qv[0], v23 = (w23 == 3) ? qv[3] : qv[2];
        const int p01 = w01 ? qp[1] : qp[0], p23 = (w23 == 3) ? qp[3] : qp[2];
        const bool hi = wins(v01, p01, v23, p23);
        const double gv = hi ? v23 : v01;
        const int gp = hi ? p23 : p01;
        const int gw = hi ? w23 : w01;
        const double atm = (gw == 0) ? qs[0] : (gw == 1) ? qs[1] : (gw == 2) ? qs[2] : qs[3];
        const double att = (gw == 0) ? qr[0] : (gw == 1) ? qr[1] : (gw == 2) ? qr[2] : qr[3];
        const double amm = (gw == 0) ? qc[0] : (gw == 1) ? qc[1] : (gw == 2) ? qc[2] : qc[3];
        if (gv < small) break;
        const int mr = (gp == INT_MAX) ? cur : (gp & 63);
        const int mc = (gp == INT_MAX) ? cur : (gp >> 6);
        // the step's swaps: first (a1 b1), then (a2 b2); s(x) = t1(t2(x)) maps a new position to
        // the current one.  p11 (1x1) or p11, p22, p21 (2x2) is the pivot (block) as it will
        // stand at position cur (and cur + 1): the winner's entry and the diagonals it carried.
        int a1 = cur, b1 = cur, a2 = -1, b2 = -1, piv2 = 0;
        double p11, p22 = 0.0, p21 = 0.0;
        if (gp == INT_MAX) {
            p11 = (&X[0][0])[__umul24((unsigned)cur, MAX_IB + 2)];
        } else if (mr == mc) {
            b1 = mr;
            p11 = atm;
        } else {
            const int t = mr, mi = mc;
            const double delta = amm * att - atm * atm;
            if (fabs(delta) >= 0.5 * atm * atm) {
                piv2 = 1;
                b1 = mi;
                a2 = cur + 1;
                b2 = (t == cur) ? mi : t;
                p11 = amm; p22 = att; p21 = atm;
            } else {
                const bool first = fabs(amm) >= fabs(att);
                b1 = first ? mi : t;
                p11 = first ? amm : att;
            }
        }
        auto sg = [&](int x) {
            x = (x == a2) ? b2 : (x == b2 ? a2 : x);
            return (x == a1) ? b1 : (x == b1 ? a1 : x);
        };
        // entry (x, y) of the lower block in the current buffer (24-bit multiplies: full rate)
        const double *Xf = &X[0][0];
        auto at = [&](int x, int y) { return Xf[__umul24((unsigned)x, MAX_IB + 1) + y]; };
        auto val = [&](int x, int y) { return x >= y ? at(x, y) : at(y, x); };
        // the step's permutation as a table in the lanes (lane l holds s(l & 31)), read by shuffle
        const int sgl = sg(lane & 31);
        const int sr = __shfl(sgl, r);
        int sj[4];
        double eo[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            sj[u] = __shfl(sgl, cg + 8 * u);
            eo[u] = val(sr, sj[u]);
        }
        const double drr = at(sr, sr);  // old diagonal of the row now at position r
        const int rid = bpX[sr];  // identity of the row now at position r
        RV_STEP_MARK(3);
        const double pinv = 1.0 / (piv2 ? p11 * p22 - p21 * p21 : p11);
        if (!piv2) {
            const int sc = sg(cur);
            const double dkk = p11;
            const double ar = val(sr, sc);
            double aj[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) aj[u] = val(sj[u], sc);
            const double inv = pinv;
            const bool row_live = r > cur && r < bs;
            const double lr = ar * inv;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int j = cg + 8 * u;
                const double ldl = (aj[u] * inv) * dkk;
                const double nv = eo[u] - lr * ldl;
                e[u] = (row_live && j > cur && j <= r) ? nv : eo[u];
            }
            // the new diagonal of the row (entry (r, r): eo = drr, aj = ar)
            dreg = row_live ? drr - lr * ((ar * inv) * dkk) : drr;
            if (cg == 0 && row_live) {
                lbuf[rid][cur] = lr;
                ml = fmax(ml, fabs(lr));
            }
            if (tid == 0) { bkind[cur] = 1; bdiag[cur] = dkk; }
            kk += 1;
        } else {
            const int s1 = sg(cur), s2 = sg(cur + 1);
            const double a11 = p11, a22 = p22, a21 = p21;
            const double ai1 = val(sr, s1), ai2 = val(sr, s2);
            double aj1[4], aj2[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                aj1[u] = val(sj[u], s1);
                aj2[u] = val(sj[u], s2);
            }
            const double inv_det = pinv;
            const bool row_live = r > cur + 1 && r < bs;
            const double l1r = (ai1 * a22 - ai2 * a21) * inv_det;
            const double l2r = (ai2 * a11 - ai1 * a21) * inv_det;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int j = cg + 8 * u;
                const double l1j = (aj1[u] * a22 - aj2[u] * a21) * inv_det;
                const double l2j = (aj2[u] * a11 - aj1[u] * a21) * inv_det;
                const double wj1 = l1j * a11 + l2j * a21;
                const double wj2 = l1j * a21 + l2j * a22;
                const double nv = eo[u] - (l1r * wj1 + l2r * wj2);
                e[u] = (row_live && j > cur + 1 && j <= r) ? nv : eo[u];
            }
            {
                // the new diagonal of the row (entry (r, r): eo = drr, aj1 = ai1, aj2 = ai2)
                const double wr1 = l1r * a11 + l2r * a21;
                const double wr2 = l1r * a21 + l2r * a22;
                dreg = row_live ? drr - (l1r * wr1 + l2r * wr2) : drr;
            }
            if (cg == 0 && row_live) {
                lbuf[rid][cur] = l1r;
                lbuf[rid][cur + 1] = l2r;
                ml = fmax(ml, fmax(fabs(l1r), fabs(l2r)));
            }
            if (tid == 0) {
                bkind[cur] = 2; bkind[cur + 1] = 3;
                bdiag[cur] = a11; bdiag[cur + 1] = a22; boff[cur] = a21;
            }
            kk += 2;
        }
        // the next buffer; the next step's barrier orders these stores before its reads
#pragma unroll
        for (int u = 0; u < 4; ++u) (&Y[0][0])[__umul24((unsigned)r, MAX_IB + 1) + cg + 8 * u] = e[u];
        if (cg == 0) bpY[r] = rid;
        double (*tX)[MAX_IB + 1] = X; X = Y; Y = tX;
        int *tb = bpX; bpX = bpY; bpY = tb;
        par ^= 1;
        RV_STEP_MARK(4);
    }
#ifdef RV_PHASE_TIMING
    if (probe)
        for (int i = 0; i < 5; ++i) atomicAdd(&g_phase[1][8 + i], tp[i]);
#endif
#undef RV_STEP_MARK
    lds_barrier();
    // the final block and row order into blk / bp
    if (X != blk) {
        for (int x = tid; x < MAX_IB * MAX_IB; x += 256) {
            const int p = x & (MAX_IB - 1), q = x >> 5;
            blk[p][q] = X[p][q];
        }
        if (tid < MAX_IB) bp[tid] = bpX[tid];
        lds_barrier();
    }
    // multipliers into the block: the row now at position p is row bp[p] of the steps that
    // formed them; the D off-diagonal of a 2x2 stays at (q+1, q)
    for (int x = tid; x < MAX_IB * MAX_IB; x += 256) {
        const int p = x & (MAX_IB - 1), q = x >> 5;
        if (q < kk && p > q && p < bs && !(bkind[q] == 2 && p == q + 1)) blk[p][q] = lbuf[bp[p]][q];
    }
    // largest |l| of the block
    const double wm = wave_max_dpp(ml);
    if (lane == 0) wbv[wv] = wm;
    lds_barrier();
    *maxl = fmax(fmax(wbv[0], wbv[1]), fmax(wbv[2], wbv[3]));
    return kk;
}

// Substitution with L11 for a block that eliminated only bn < MAX_IB columns: columns and rows
// from bn on take no part.  Kept out of line: its guards are uniform predicates that, inlined,
// the compiler keeps live in registers across the common full-block path.
__device__ __attribute__((noinline)) void panel_trsm_partial(double *y, const double (*blk)[MAX_IB + 1], int bn) {
    const int lc = threadIdx.x & 31;
    double v[MAX_IB];
#pragma unroll
    for (int j = 0; j < MAX_IB; ++j) v[j] = y[j];
#pragma unroll
    for (int i = 0; i < MAX_IB - 1; ++i) {
        const double colv = blk[lc][i];
        if (i < bn) {
            const double yi = v[i];
#pragma unroll
            for (int j = i + 1; j < MAX_IB; ++j) {
                const double t = v[j] - readlane_f64(colv, j) * yi;
                v[j] = (j < bn) ? t : v[j];
            }
        }
    }
#pragma unroll
    for (int j = 0; j < MAX_IB; ++j) y[j] = v[j];
}

// Forward substitution with a full L11 in groups: group (I, J) subtracts the terms of column I
// from rows J .. min(J + 8, MAX_IB) - 1; the groups run column by column, so every row receives
// its terms in ascending column order.  A group's eight products are formed before any of its
// subtractions (a scheduling barrier between them), so the products' latencies overlap instead
// of each subtraction waiting for its own product.
template <int I, int J>
struct TrsmNext {
    static constexpr bool more = J + 8 < MAX_IB;
    static constexpr int i = more ? I : I + 1;
    static constexpr int j = more ? J + 8 : I + 2;
    static constexpr bool valid = i < MAX_IB - 1;
};
template <int I, int J>
__device__ __forceinline__ void trsm_group_load(double (&c)[8], const double (*blk)[MAX_IB + 1]) {
#pragma unroll
    for (int u = 0; u < 8; ++u) c[u] = (J + u < MAX_IB) ? blk[J + u < MAX_IB ? J + u : 0][I] : 0.0;
}
template <int I, int J>
__device__ __forceinline__ void trsm_groups(double (&y)[MAX_IB], const double (*blk)[MAX_IB + 1],
                                            const double (&c0)[8], const double (&c1)[8]) {
    using N1 = TrsmNext<I, J>;
    using N2 = TrsmNext<N1::i, N1::j>;
    double c2[8] = {};
    if constexpr (N1::valid && N2::valid) trsm_group_load<N2::i, N2::j>(c2, blk);
    const double yi = y[I];
    double t[8];
#pragma unroll
    for (int u = 0; u < 8; ++u)
        if (J + u < MAX_IB) t[u] = c0[u] * yi;
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < 8; ++u)
        if (J + u < MAX_IB) y[J + u] = y[J + u] - t[u];
    if constexpr (N1::valid) trsm_groups<N1::i, N1::j>(y, blk, c1, c2);
}

// Forward substitution with L11 and D^-1 for panel rows, one row per thread (factor.rs:1295-1408):
// y holds the row's block entries in pivot order on entry and its L entries on return.  blk is L11
// (2x2 D off-diagonals read as zeros), bkind / dinv / bdiag / boff the pivots (kind 0 beyond the
// eliminated columns).  L11's columns and the pivots' D sit in the lanes (lane c & 31 holds column
// / pivot c) and are taken by readlane: no LDS round trip per entry.  Every lane of the wavefront
// runs it (rows beyond the front hold zeros and store nothing), so the lanes read are active.
__device__ __forceinline__ void panel_trsm_d(double (&y)[MAX_IB], const double (*blk)[MAX_IB + 1], int bn,
                                             const int8_t *bkind, const double *dinv, const double *bdiag,
                                             const double *boff) {
    const int lc = threadIdx.x & 31;
    const int kdl = bkind[lc];
    const double dil = dinv[lc], bdl = bdiag[lc], bol = boff[lc];
    // forward substitution with L11 by columns: y[j] receives the terms of i = 0, 1, ...
    // in ascending order, as in the row form, while the rows stay independent
    if (bn == MAX_IB) {
        // every column eliminated (the common case): no guards, L11 entries read from LDS at
        // the same address by all lanes (broadcast), in groups of eight loaded two groups ahead
        double c0[8], c1[8];
        trsm_group_load<0, 1>(c0, blk);
        trsm_group_load<TrsmNext<0, 1>::i, TrsmNext<0, 1>::j>(c1, blk);
        trsm_groups<0, 1>(y, blk, c0, c1);
    } else {
        // through a copy: only the copy's address leaves the function, y stays in registers
        double t[MAX_IB];
#pragma unroll
        for (int j = 0; j < MAX_IB; ++j) t[j] = y[j];
        panel_trsm_partial(t, blk, bn);
#pragma unroll
        for (int j = 0; j < MAX_IB; ++j) y[j] = t[j];
    }
    // D^-1 (1x1: y / d; 2x2: Cramer's rule with the stored inverse determinant); columns beyond
    // the eliminated ones have kind 0 and stay as they are
#pragma unroll
    for (int c = 0; c < MAX_IB; ++c) {
        const int kd = __builtin_amdgcn_readlane(kdl, c);
        if (kd == 1) {
            y[c] *= readlane_f64(dil, c);
        } else if (kd == 2 && c + 1 < MAX_IB) {
            const double x1 = y[c], x2 = y[c + 1];
            const double di = readlane_f64(dil, c), bo = readlane_f64(bol, c);
            y[c] = (x1 * readlane_f64(bdl, c + 1) - x2 * bo) * di;
            y[c + 1] = (x2 * readlane_f64(bdl, c) - x1 * bo) * di;
        }
    }
}

// ===========================================================================
// APTP inner block, part 1 (one workgroup per 256 panel rows of a front): every workgroup
// factors the diagonal block itself (same arithmetic, same result), the first one records it --
// unless the previous update kernel already factored it (look-ahead, AptpState::pref); then each
// applies it to its panel rows -- backup, column permutation, TRSM with L11, D^-1 scaling and the
// a-posteriori threshold check (apply_and_check, factor.rs:1295-1408).
// ===========================================================================
__global__ __launch_bounds__(256) void k_block_apply(FrontDev *__restrict__ fr,
                                                     AptpState *__restrict__ st,
                                                     const int *__restrict__ flist,
                                                     const int4 *__restrict__ items, int ob,
                                                     int it, int nb, int ib, double small,
                                                     double threshold) {
    RV_PHASE_START();
    const int4 itm = items[blockIdx.x];
    const int li = itm.x, chunk = itm.y;
    FrontDev &F = fr[itm.z];
    AptpState &S = st[li];
    if (!aptp_active(F.k, nb, ib, ob, it)) return;
    const int k = S.kpos, m = F.m;
    const int bs = min(S.end_pos - k, ib), be = k + bs;
    const size_t lda = m;
    gdouble *A = gbl(F.F);
    // read once: a store through a front pointer may alias the state, so fields read after
    // stores would be reloaded before every one of them
    gdouble *const bk = gbl(S.backup);
    RV_PHASE_COUNT(0);
    RV_PHASE(0, 0);
    __shared__ double blk[MAX_IB][MAX_IB + 1];
    __shared__ int bp[MAX_IB], operm[MAX_IB];
    __shared__ int8_t bkind[MAX_IB];
    __shared__ double bdiag[MAX_IB], boff[MAX_IB], dinv[MAX_IB];
    __shared__ double wbv[8], wbs[24];
    __shared__ int wbp[8], bp2[MAX_IB];
    __shared__ double lbuf[MAX_IB][MAX_IB + 1], blk2[MAX_IB][MAX_IB + 1];
    const int tid = threadIdx.x;
    const int r = be + chunk * 256 + tid;
    double y[MAX_IB];
    int bn;
    if (S.pref) {
        // the block was factored by the previous update kernel (look-ahead): its pivots, D and
        // L11 (2x2 D off-diagonals already zero) come from the state and the front; the panel
        // rows load directly in pivot order and are backed up in the original column order
        bn = S.block_nelim;
#ifdef RV_PHASE_TIMING
        if (blockIdx.x == 0 && tid == 0) atomicAdd(&g_phase[0][10], 1ull);
#endif
        if (tid < MAX_IB) {
            bp[tid] = S.bperm[tid];
            bkind[tid] = S.bkind[tid];
            bdiag[tid] = S.bdiag[tid];
            boff[tid] = S.boff[tid];
        }
        for (int e = tid; e < MAX_IB * MAX_IB; e += 256) {
            const int i = e & (MAX_IB - 1), j = e / MAX_IB;
            if (i < bs && j < i) blk[i][j] = A[(k + i) + (size_t)(k + j) * lda];
        }
        int pj[MAX_IB];
#pragma unroll
        for (int j = 0; j < MAX_IB; ++j) pj[j] = S.bperm[j];
        if (r < m) {
            gdouble *row = A + r;
#pragma unroll
            for (int j = 0; j < MAX_IB; ++j) y[j] = (j < bs) ? row[(size_t)(k + pj[j]) * lda] : 0.0;
#pragma unroll
            for (int j = 0; j < MAX_IB; ++j)
                if (j < bs) bk[(r - k) + (size_t)pj[j] * lda] = y[j];
        }
        lds_barrier();
        if (tid < bn) {
            const int kd = bkind[tid];
            if (kd == 1) dinv[tid] = 1.0 / bdiag[tid];
            else if (kd == 2) dinv[tid] = 1.0 / (bdiag[tid] * bdiag[tid + 1] - boff[tid] * boff[tid]);
        }
        lds_barrier();
        RV_PHASE(0, 3);
        if (bn == 0) return;
    } else {
    // every global load of the kernel's first part is issued before any is used: the panel
    // rows (backed up for restore_diagonal_with_perm), the lower block, the column order
    double pv[MAX_IB];
#pragma unroll
    for (int j = 0; j < MAX_IB; ++j) pv[j] = (r < m && j < bs) ? A[r + (size_t)(k + j) * lda] : 0.0;
    const int bi_ = tid & 31, bj0 = tid >> 5;
    double bv[MAX_IB / 8];
    const gdouble *const bc = gbl(S.blkcopy);
#pragma unroll
    for (int u = 0; u < MAX_IB / 8; ++u) {
        const int j = bj0 + 8 * u;
        const double v = bc[bi_ + j * MAX_IB];  // always in bounds (MAX_IB x MAX_IB)
        bv[u] = (bi_ < bs && j <= bi_) ? v : 0.0;
    }
    const int op = (chunk == 0 && tid < bs) ? gbl(F.perm)[k + tid] : 0;
#pragma unroll
    for (int u = 0; u < MAX_IB / 8; ++u) {
        const int j = bj0 + 8 * u;
        if (bi_ < bs && j <= bi_) {
            blk[bi_][j] = bv[u];
            if (chunk == 0) bk[bi_ + (size_t)j * lda] = bv[u];
        }
    }
    if (tid < MAX_IB) operm[tid] = op;
    lds_barrier();
    RV_PHASE(0, 1);
    double bml;
    bn = block_factor_wg(blk, blk2, bp, bp2, bkind, bdiag, boff, bs, small, &bml, wbv, wbp, wbs, lbuf);
    RV_PHASE(0, 2);
    // D off-diagonals of 2x2 pivots read as zeros of L11 (factor.rs:1898-1912)
    if (tid < bn) {
        const int kd = bkind[tid];
        if (kd == 2) blk[tid + 1][tid] = 0.0;
        if (kd == 1) dinv[tid] = 1.0 / bdiag[tid];
        else if (kd == 2) dinv[tid] = 1.0 / (bdiag[tid] * bdiag[tid + 1] - boff[tid] * boff[tid]);
    }
    if (r < m) {
#pragma unroll
        for (int j = 0; j < MAX_IB; ++j)
            if (j < bs) bk[(r - k) + (size_t)j * lda] = pv[j];
    }
    lds_barrier();
    if (chunk == 0) {
        if (bn > 0) {
            const int i = tid & 31;
            for (int j = tid >> 5; j < bs; j += 8)
                if (i < bs && j <= i) A[(k + i) + (k + j) * lda] = blk[i][j];
            // column order by the block permutation (factor.rs:1926-1930)
            if (tid < bs) gbl(F.perm)[k + tid] = operm[bp[tid]];
        }
        if (tid < MAX_IB) {
            S.bperm[tid] = bp[tid];
            S.bkind[tid] = (tid < bn) ? bkind[tid] : 0;
            S.bdiag[tid] = bdiag[tid];
            S.boff[tid] = boff[tid];
        }
        if (tid == 0) {
            S.bs = bs;
            S.block_nelim = bn;
            atomicMin(&S.fail_min, bn);
            atomic_max_nonneg(&F.max_l_bits, bml);
        }
    }
    RV_PHASE(0, 3);
    if (bn == 0) return;
    if (r < m) {
        gdouble *row = A + r;
#pragma unroll
        for (int j = 0; j < MAX_IB; ++j) y[j] = (j < bs) ? row[(k + bp[j]) * lda] : 0.0;
    }
    }
    int first = INT_MAX;
    if (r >= m) {
#pragma unroll
        for (int j = 0; j < MAX_IB; ++j) y[j] = 0.0;
    }
    panel_trsm_d(y, blk, bn, bkind, dinv, bdiag, boff);
    if (r < m) {
        gdouble *row = A + r;
#pragma unroll
        for (int j = 0; j < MAX_IB; ++j)
            if (j < bs) row[(k + j) * lda] = y[j];
        const double bound = 1.0 / threshold;
#pragma unroll
        for (int c = MAX_IB - 1; c >= 0; --c)
            if (c < bn && fabs(y[c]) >= bound) first = c;
        if (first != INT_MAX && bkind[first] == 3) first -= 1;
    }
    RV_PHASE(0, 4);
    // first failing column: wavefront minimum, one atomic per wavefront (no closing barrier)
    first = wave_min_int(first);
    if ((tid & 63) == 0 && first != INT_MAX) atomicMin(&S.fail_min, first);
    RV_PHASE(0, 5);
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
    gdouble *A = gbl(F.F);
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
    gint *mark = gbl(S.mark);
    gdouble *buf = gbl(S.swapbuf);
    for (int u = tid; u < nU; u += blockDim.x) mark[U[u]] = u + 1;
    for (int e = tid; e < nU * m; e += blockDim.x) {
        const int ui = e / m, q = e % m;
        const int u = U[ui];
        buf[e] = (u >= q) ? A[u + q * lda] : A[q + u * lda];
    }
    if (tid < nU) ptmp[tid] = gbl(F.perm)[U[tid]];
    __syncthreads();
    for (int pi = 0; pi < nU; ++pi) {
        const int p = U[pi];
        const int sp = src[pi];
        if (sp == p) continue;
        const int sidx = mark[sp] - 1;
        const gdouble *line = buf + (size_t)sidx * m;
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
        gbl(F.perm)[U[tid]] = ptmp[mark[sp] - 1];
    }
    __syncthreads();
    for (int u = tid; u < nU; u += blockDim.x) mark[U[u]] = 0;
    __syncthreads();
}

// Diagonal tile of the next inner block after this block's update, formed in LDS with the
// trailing update's arithmetic: C(i, c) -= sum_j W(i, j) L(c, j), summed over j ascending from
// zero and subtracted once.  Lr holds the next block's L rows in this block's columns, kd / dd /
// od this block's pivots; Wt receives W = L D (compute_ld_into, factor.rs:1536-1563).  Thread
// (ii = tid & 31, jg = tid >> 5) returns its updated entries (ii, jg + 8u) in nv.  Starts with the
// LDS operands in place and ends with a barrier.
__device__ __forceinline__ void next_tile_update(const double (*Lr)[MAX_IB + 1], double (*Wt)[MAX_IB + 1],
                                                 double (*Ct)[MAX_IB + 1], const int8_t *kd, const double *dd,
                                                 const double *od, int nbs, int nelim,
                                                 double (&nv)[MAX_IB / 8]) {
    const int tid = threadIdx.x, ii = tid & 31, jg = tid >> 5;
#pragma unroll
    for (int u = 0; u < MAX_IB / 8; ++u) {
        const int j = jg + 8 * u;
        double w = 0.0;
        if (ii < nbs && j < nelim) {
            const int kj = kd[j];
            if (kj == 1) w = Lr[ii][j] * dd[j];
            else if (kj == 2) w = Lr[ii][j] * dd[j] + Lr[ii][j + 1] * od[j];
            else if (kj == 3) w = Lr[ii][j - 1] * od[j - 1] + Lr[ii][j] * dd[j];
        }
        Wt[ii][j] = w;
    }
    lds_barrier();
    // the thread's four sums run side by side (independent chains, one W read per step); a full
    // block without guards, so the LDS reads of later steps are issued ahead
    double sacc[MAX_IB / 8];
#pragma unroll
    for (int u = 0; u < MAX_IB / 8; ++u) sacc[u] = 0.0;
    if (nelim == MAX_IB) {
        // operands of step j + 2 loaded during step j (three register sets in turn), a scheduling
        // barrier per step so that the loads are not regrouped behind their uses
        double wq[3], lq[3][MAX_IB / 8];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            wq[j] = Wt[ii][j];
#pragma unroll
            for (int u = 0; u < MAX_IB / 8; ++u) lq[j][u] = Lr[jg + 8 * u][j];
        }
#pragma unroll
        for (int j = 0; j < MAX_IB; ++j) {
            if (j + 2 < MAX_IB) {
                wq[(j + 2) % 3] = Wt[ii][j + 2];
#pragma unroll
                for (int u = 0; u < MAX_IB / 8; ++u) lq[(j + 2) % 3][u] = Lr[jg + 8 * u][j + 2];
            }
#pragma unroll
            for (int u = 0; u < MAX_IB / 8; ++u) sacc[u] += wq[j % 3] * lq[j % 3][u];
            __builtin_amdgcn_sched_barrier(0);
        }
    } else {
#pragma unroll
        for (int j = 0; j < MAX_IB; ++j) {
            if (j < nelim) {
                const double w = Wt[ii][j];
#pragma unroll
                for (int u = 0; u < MAX_IB / 8; ++u) sacc[u] += w * Lr[jg + 8 * u][j];
            }
        }
    }
#pragma unroll
    for (int u = 0; u < MAX_IB / 8; ++u) {
        const int c = jg + 8 * u;
        nv[u] = 0.0;
        if (ii < nbs && c <= ii) {
            nv[u] = Ct[ii][c] - sacc[u];
            Ct[ii][c] = nv[u];
        }
    }
    lds_barrier();
}

// Look-ahead factorization of the next inner block [be, be + nbs), by one extra workgroup per
// front of the update kernel.  It forms the block's diagonal tile after this block's update on its
// own -- C - sum_j w_j l_j over the block's eliminated columns in ascending order (the arithmetic
// of the trailing update, which writes the same tile to the front), with every operand loaded in
// one round trip -- and then takes the steps of k_block_apply's staging part: backup of the block,
// complete pivoting, D off-diagonals of 2x2 pivots zeroed in L11, the factored block and the
// column order written to the front; pivots and D go to the nx_* fields, which k_block_finish
// hands over.  Wr (64 x 33) and Lc (32 x 33) are its LDS.
constexpr int LOOKAHEAD_Q = 255;  // k_block_update work item of the look-ahead workgroup

__device__ void lookahead_factor(FrontDev &F, AptpState &S, double (*Wr)[MAX_IB + 1], double (*Lc)[MAX_IB + 1],
                                 int k, int be, int nelim, int nbs, double small, unsigned long long t_start) {
    // the chain below is the critical path of the update kernel: its wavefronts issue first
    __builtin_amdgcn_s_setprio(3);
#ifdef RV_PHASE_TIMING
    const bool probe = threadIdx.x == 0 && blockIdx.x < 64;
    if (probe) { atomicAdd(&g_phase[1][6], clock64() - t_start); atomicAdd(&g_phase[1][5], 1ull); }
    const unsigned long long t0 = clock64();
#endif
    __shared__ int bp[MAX_IB], bp2[MAX_IB], operm[MAX_IB];
    __shared__ int8_t bkind[MAX_IB];
    __shared__ double bdiag[MAX_IB], boff[MAX_IB];
    __shared__ double wbv[8], wbs[24];
    __shared__ int wbp[8];
    __shared__ int8_t kd[MAX_IB];
    __shared__ double dd[MAX_IB], od[MAX_IB];
    double (*Lr)[MAX_IB + 1] = Wr;            // L rows of the next block in this block's columns
    double (*Wt)[MAX_IB + 1] = Wr + MAX_IB;   // their W = L D
    double (*Ct)[MAX_IB + 1] = Lc;            // the diagonal tile, updated in place
    const int tid = threadIdx.x;
    const size_t lda = F.m;
    gdouble *A = gbl(F.F);
    gdouble *const bk = gbl(S.backup);
    // one round trip: D of this block, the L rows, the tile, the column order
    const int ii = tid & 31, jg = tid >> 5;
    double lv[MAX_IB / 8], cv[MAX_IB / 8];
#pragma unroll
    for (int u = 0; u < MAX_IB / 8; ++u) {
        const int j = jg + 8 * u;
        lv[u] = (ii < nbs && j < nelim) ? A[(be + ii) + (size_t)(k + j) * lda] : 0.0;
        cv[u] = (ii < nbs && j <= ii) ? A[(be + ii) + (size_t)(be + j) * lda] : 0.0;
    }
    int8_t kv = 0;
    double dv = 0.0, ov = 0.0;
    int opv = 0;
    if (tid < MAX_IB) {
        kv = S.bkind[tid];
        dv = S.bdiag[tid];
        ov = S.boff[tid];
        opv = (tid < nbs) ? gbl(F.perm)[be + tid] : 0;
    }
#pragma unroll
    for (int u = 0; u < MAX_IB / 8; ++u) {
        const int j = jg + 8 * u;
        Lr[ii][j] = lv[u];
        Ct[ii][j] = cv[u];
    }
    if (tid < MAX_IB) { kd[tid] = kv; dd[tid] = dv; od[tid] = ov; operm[tid] = opv; }
    lds_barrier();
    double nv[MAX_IB / 8];
    next_tile_update(Lr, Wt, Ct, kd, dd, od, nbs, nelim, nv);
#pragma unroll
    for (int u = 0; u < MAX_IB / 8; ++u) {
        const int c = jg + 8 * u;
        if (ii < nbs && c <= ii) bk[ii + (size_t)c * lda] = nv[u];
    }
    double bml;
#ifdef RV_PHASE_TIMING
    if (probe) atomicAdd(&g_phase[1][2], clock64() - t0);
    const int bn = block_factor_wg(Ct, Wr, bp, bp2, bkind, bdiag, boff, nbs, small, &bml, wbv, wbp, wbs,
                                   Wr + MAX_IB, probe);
    if (probe) atomicAdd(&g_phase[1][3], clock64() - t0);
#else
    const int bn = block_factor_wg(Ct, Wr, bp, bp2, bkind, bdiag, boff, nbs, small, &bml, wbv, wbp, wbs,
                                   Wr + MAX_IB);
#endif
    double (*blk)[MAX_IB + 1] = Ct;
    if (tid < bn && bkind[tid] == 2) blk[tid + 1][tid] = 0.0;
    lds_barrier();
    // the tile goes back to the front also when nothing was eliminated: the update kernel leaves
    // it to this workgroup, so the front holds it only from here
    {
        const int i2 = tid & 31;
        for (int j = tid >> 5; j < nbs; j += 8)
            if (i2 < nbs && j <= i2) A[(be + i2) + (size_t)(be + j) * lda] = blk[i2][j];
    }
    if (bn > 0 && tid < nbs) gbl(F.perm)[be + tid] = operm[bp[tid]];
    if (tid < MAX_IB) {
        S.nx_bperm[tid] = bp[tid];
        S.nx_bkind[tid] = (tid < bn) ? bkind[tid] : 0;
        S.nx_bdiag[tid] = bdiag[tid];
        S.nx_boff[tid] = boff[tid];
    }
    if (tid == 0) {
        S.nx_bs = nbs;
        S.nx_bn = bn;
        S.nx_valid = 1;
        atomic_max_nonneg(&F.max_l_bits, bml);
    }
#ifdef RV_PHASE_TIMING
    if (probe) atomicAdd(&g_phase[1][4], clock64() - t0);
#endif
    __builtin_amdgcn_s_setprio(0);
}

// ===========================================================================
// APTP inner block, part 2 (tiles of 64 rows x 32 columns relative to kpos; column tile 0 =
// the current block): partial restore of failed columns, cross terms, Schur update of the
// trailing columns of the outer block, max|L| statistics (factor.rs:1945-2008:
// restore_diagonal_with_perm, update_trailing, update_cross_terms), and the block row
// permutation of the already-factored columns (factor.rs:1914-1924).
// ===========================================================================
__global__ __launch_bounds__(256) void k_block_update(FrontDev *__restrict__ fr,
                                                      AptpState *__restrict__ st,
                                                      const int *__restrict__ flist,
                                                      const int4 *__restrict__ items, int ob,
                                                      int it, int nb, int ib, double small) {
    RV_PHASE_START();
    RV_PHASE_COUNT(1);
    const int4 itm = items[blockIdx.x];
    const int li = itm.x, rt = itm.y, q = itm.z;
    FrontDev &F = fr[itm.w];
    AptpState &S = st[li];
    if (!aptp_active(F.k, nb, ib, ob, it)) return;
    const int bn = S.block_nelim;
    const int k = S.kpos, bs = S.bs, be = k + bs, m = F.m;
    // trailing columns updated by this inner block: every fully-summed column, as
    // update_trailing does (factor.rs:1418-1526 with nfs_boundary, 2149), or only those of the
    // current outer block when the rest is deferred to the rank-nb MFMA update
    const int obe = (F.flags & FRONT_ALLFS) ? F.k : S.ob_start + S.ob_cols;
    const size_t lda = m;
    gdouble *A = gbl(F.F);
    gdouble *const bk = gbl(S.backup);
    const int nelim = (bn > 0) ? effective_nelim(S) : 0;
    __shared__ double Wr[64][MAX_IB + 1];
    __shared__ double Lc[32][MAX_IB + 1];
    __shared__ int bp[MAX_IB];
    __shared__ int kind[MAX_IB];
    __shared__ double dg[MAX_IB], of[MAX_IB];
    const int tid = threadIdx.x;
    // look-ahead: after a block without failures, the next block [be, be + nbs) is factored by an
    // extra workgroup (only when the next inner-block launch will process it: the look-ahead
    // writes it); its diagonal tile is then that workgroup's alone
    const int nbs = min(S.end_pos - be, ib);
    const bool ahead = bn > 0 && nelim == bs && nbs > 0 && aptp_active(F.k, nb, ib, ob, it + 1);
    if (q == LOOKAHEAD_Q) {
        if (ahead) {
#ifdef RV_PHASE_TIMING
            const unsigned long long t_start = rv_t0_;
#else
            const unsigned long long t_start = 0;
#endif
            lookahead_factor(F, S, Wr, Lc, k, be, nelim, nbs, small, t_start);
        }
        return;
    }
    RV_PHASE(1, 0);
    if (tid < MAX_IB) {
        bp[tid] = S.bperm[tid];
        kind[tid] = S.bkind[tid];
        dg[tid] = S.bdiag[tid];
        of[tid] = S.boff[tid];
    }
    lds_barrier();
    int c_lo, c_hi;
    bool active = bn > 0;
    if (q == 0) { c_lo = k; c_hi = be; }
    else {
        c_lo = be + 32 * (q - 1);
        c_hi = min(c_lo + 32, obe);
        if (c_lo >= obe || nelim == 0) active = false;
    }
    const int r_lo = k + 64 * rt;
    const int r_hi = min(r_lo + 64, m);
    if (r_lo >= m || (q >= 1 && r_hi <= c_lo)) active = false;
    if (active) {
        const int ncol = c_hi - c_lo;
        for (int e = tid; e < 64 * nelim; e += 256) {
            const int rr = e & 63, j = e >> 6;
            const int r = r_lo + rr;
            Wr[rr][j] = (r < r_hi) ? ld_entry(A + r + k * lda, lda, j, kind[j], dg, of) : 0.0;
        }
        if (q >= 1) {
            for (int e = tid; e < 32 * nelim; e += 256) {
                const int cc = e & 31, j = e >> 5;
                if (cc < ncol) Lc[cc][j] = A[(c_lo + cc) + (k + j) * lda];
            }
        } else {
            // L rows of the failed block rows: Lc[jf][c] = A(k+nelim+jf, k+c)
            const int nf = bs - nelim;
            for (int e = tid; e < 32 * nelim; e += 256) {
                const int jf = e & 31, c = e >> 5;
                if (jf < nf) Lc[jf][c] = A[(k + nelim + jf) + (k + c) * lda];
            }
        }
        lds_barrier();
        const int rr = tid & 63;
        const int r = r_lo + rr;
        const bool valid = r < r_hi;
        gdouble *row = A + r;
        if (q >= 1) {
            if (valid) {
                // the row's W values stay in registers across the tile's columns; each sum runs
                // over c ascending from zero and is subtracted once (gemm_sub_wlt order)
                double w[MAX_IB];
#pragma unroll
                for (int j = 0; j < MAX_IB; ++j) w[j] = Wr[rr][j];
                // the row's 8 entries are loaded before any is stored (a store may alias a later
                // load, so interleaved read-modify-writes would take one round trip each)
                double cv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int cc = (tid >> 6) + 4 * u;
                    const int c = c_lo + cc;
                    const bool tile_ahead = ahead && r < be + nbs && c < be + nbs;
                    cv[u] = (cc < ncol && r >= c && !tile_ahead) ? row[c * lda] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int cc = (tid >> 6) + 4 * u;
                    const int c = c_lo + cc;
                    const bool tile_ahead = ahead && r < be + nbs && c < be + nbs;
                    if (cc < ncol && r >= c && !tile_ahead) {
                        double s = 0.0;
#pragma unroll
                        for (int j = 0; j < MAX_IB; ++j)
                            if (j < nelim) s += w[j] * Lc[cc][j];
                        row[c * lda] = cv[u] - s;
                    }
                }
            }
        } else {
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
                                row[c * lda] = (br >= bc) ? bk[br + bc * lda] : bk[bc + br * lda];
                            }
                        } else {
                            row[c * lda] = bk[(r - k) + bp[j] * lda];
                        }
                    }
                    // cross terms from the passed columns
                    if (nelim > 0) {
                        const int jf = j - nelim;
                        const bool below = r >= be;
                        if (below || r - k - nelim >= jf) {
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
    }
    if (q == 0 && bn > 0) {
        // block rows permuted in the columns factored before this block: two columns per
        // wavefront instruction, lane&31 = row of the block
        const int ntr = (m + 63) / 64;
        const int per = (k + ntr - 1) / ntr;
        const int cb = rt * per, ce = min(k, cb + per);
        const int i = tid & 31, half = (tid >> 5) & 1, wv = tid >> 6;
        if (i < bs) {
            const int src = k + bp[i];
            for (int c0 = cb + 2 * wv + half; c0 < ce; c0 += 32) {
                double v[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int c = c0 + 8 * u;
                    v[u] = (c < ce) ? A[src + (size_t)c * lda] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int c = c0 + 8 * u;
                    if (c < ce) A[(k + i) + (size_t)c * lda] = v[u];
                }
            }
        }
    }
    RV_PHASE(1, 1);
}

// End of an inner block (one workgroup per front): D and the pivot counts of the accepted
// columns, then the failed columns are delayed to the end of the outer block
// (factor.rs:1962-2077).
__global__ __launch_bounds__(256) void k_block_finish(FrontDev *__restrict__ fr,
                                                      AptpState *__restrict__ st,
                                                      const int *__restrict__ flist, int ob, int it,
                                                      int nb, int ib) {
    const int li = blockIdx.x;
    FrontDev &F = fr[flist[li]];
    AptpState &S = st[li];
    if (!aptp_active(F.k, nb, ib, ob, it)) return;
    const int tid = threadIdx.x;
    const int bn = S.block_nelim;
    const int k = S.kpos, bs = S.bs;
    const int nelim = (bn > 0) ? effective_nelim(S) : 0;
    const int nxv = S.nx_valid;
    __shared__ int sa[MAX_IB], sb[MAX_IB], U[2 * MAX_IB], src[2 * MAX_IB], ptmp[2 * MAX_IB];
    __shared__ int nsw, cnt, s_end;
    int is1 = 0, is2 = 0;
    if (tid < nelim) {
        const int kd = S.bkind[tid];
        const int c = k + tid;
        if (kd == 1) {
            gbl(F.dkind)[c] = 1; gbl(F.ddiag)[c] = S.bdiag[tid]; gbl(F.doff)[c] = 0.0;
            is1 = 1;
        } else if (kd == 2 && tid + 1 < nelim) {
            gbl(F.dkind)[c] = 2; gbl(F.dkind)[c + 1] = 3;
            gbl(F.ddiag)[c] = S.bdiag[tid]; gbl(F.ddiag)[c + 1] = S.bdiag[tid + 1];
            gbl(F.doff)[c] = S.boff[tid]; gbl(F.doff)[c + 1] = 0.0;
            is2 = 1;
        }
    }
    const int n1 = __syncthreads_count(is1);
    const int n2 = __syncthreads_count(is2);
    if (tid == 0) {
        F.n1 += n1;
        F.n2 += n2;
        int end_pos = S.end_pos;
        int n = 0;
        if (bn == 0) {
            for (int off = 0; off < bs; ++off) {
                end_pos -= 1;
                if (k + off < end_pos) { sa[n] = k + off; sb[n] = end_pos; ++n; }
            }
        } else {
            for (int i = bs - nelim - 1; i >= 0; --i) {
                const int dp = k + nelim + i;
                end_pos -= 1;
                if (dp < end_pos) { sa[n] = dp; sb[n] = end_pos; ++n; }
            }
        }
        nsw = n;
        s_end = end_pos;
    }
    __syncthreads();
    apply_swaps(F, S, sa, sb, nsw, U, src, &cnt, ptmp);
    if (tid == 0) {
        S.kpos = k + nelim;
        S.end_pos = s_end;
        S.fail_min = INT_MAX;
        if (S.nib < MAX_OB_IB) S.ib_ne[S.nib] = (int8_t)nelim;
        S.nib += 1;
    }
    __syncthreads();
    if (nxv) {
        // the next block was factored ahead by the update kernel (this block had no failures,
        // so nothing moved): its pivots and D become the current block's
        if (tid < MAX_IB) {
            S.bperm[tid] = S.nx_bperm[tid];
            S.bkind[tid] = S.nx_bkind[tid];
            S.bdiag[tid] = S.nx_bdiag[tid];
            S.boff[tid] = S.nx_boff[tid];
        }
        if (tid == 0) {
            S.bs = S.nx_bs;
            S.block_nelim = S.nx_bn;
            S.fail_min = S.nx_bn;
            S.pref = 1;
            S.nx_valid = 0;
        }
    } else {
        if (tid == 0) S.pref = 0;
        if (S.kpos < S.end_pos) copy_next_block(F, S, ib);
    }
}

// ===========================================================================
// APTP outer block in one persistent launch (k_aptp_outer), for levels whose workgroups all fit
// on the device at once.  Per front one leader workgroup runs the pivot chain and NP panel
// workgroups own chunks of CH rows from the outer block's first column down:
//   leader, inner block it: factor the diagonal block (or take the one factored ahead), publish
//     its pivots (record it); then, when the block eliminated every column, form the next
//     block's L rows itself (the panel substitution for those 32 rows), its diagonal tile after
//     this block's update, and factor that tile ahead; once every panel has checked block it,
//     keep the block factored ahead when no column failed (write it to the front, backup, column
//     order) or, after failures, wait for the panels' restores, delay the failed columns to the
//     end of the outer block (k_block_finish's swaps) and factor the next block afresh.
//   panel, inner block it: backup, column permutation, substitution and threshold check of its
//     rows (k_block_apply), then, once every panel has checked, partial restore and cross terms,
//     the update of the outer block's trailing columns in its rows (the tile factored ahead left
//     to the leader) and its share of the row permutation of earlier columns (k_block_update).
// The arithmetic of every entry is the three-kernel path's, so both produce the same bits.
// Front entries, backups and records that one workgroup writes and another reads within the
// launch go through device-scope loads and stores, which bypass the per-XCD L2 (another XCD's
// copy would be stale); a producer's stores complete (s_waitcnt) before it bumps a counter, and a
// consumer's thread 0 spins on the counter with device-scope loads (bounded: a wait that runs out
// sets the front's error word, every wait of the front then gives up, and the host discards the
// launch's results).
// ===========================================================================
constexpr int CHAIN_MAX_IT = 8;                 // inner blocks per outer block
constexpr long long CHAIN_SPIN_LIMIT = 1ll << 22;
constexpr int FRONT_CHAIN_ERR = 8;              // FrontDev::flags: a wait of k_aptp_outer ran out
constexpr int FRONT_LA_ERR = 16;   // FrontDev::flags: an outer block run with its trailing update
                                   // overlapped (no failures planned) had failed columns

struct ChainRec {
    int k, bs, bn, end_pos;
    int bperm[MAX_IB];
    int bkind[MAX_IB];
    double bdiag[MAX_IB], boff[MAX_IB];
};
// Every counter on a cache line of its own: the waits poll them from many workgroups at once,
// and a line shared by several counters would queue every poll and update behind the others.
// aux is a value settled before each increment of v (the first failing column for the applied
// counters): one 8-byte load returns both, and a count seen complete comes with its final aux.
struct alignas(128) ChainCounter {
    int v;
    int aux;
    int pad[30];
};
struct ChainSync {
    ChainCounter rec;                       // inner-block records published by the leader
    ChainCounter err;                       // a wait of this front ran out
    ChainCounter applied[CHAIN_MAX_IT];     // panels done with the check of block it; aux: its
                                            // first failing column (INT_MAX before)
    ChainCounter updated[CHAIN_MAX_IT];     // panels done with the updates of block it
    ChainRec recs[CHAIN_MAX_IT];
};

__device__ __forceinline__ void chain_reset(ChainSync &C) {
    const int i = threadIdx.x;
    if (i == 0) { C.rec.v = 0; C.err.v = 0; }
    if (i < CHAIN_MAX_IT) {
        C.applied[i].v = 0;
        C.applied[i].aux = INT_MAX;
        C.updated[i].v = 0;
    }
}

__device__ __forceinline__ double co_ld(const gdouble *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void co_st(gdouble *p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int co_ldi(const int *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void co_sti(int *p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Bulk device-coherent loads: a plain buffer load carrying the same sc1 cache-policy bit as a
// device-scope atomic load, which the compiler batches like any other load (atomic loads are
// issued one round trip at a time).  The base is wavefront-uniform, the byte offset per lane; the
// resource's range turns an offset past it into zero.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t co_rsrc(const gdouble *base) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>((const double *)base), (short)0, 0x7FFFFFF0,
                                             0x00020000);
}
__device__ __forceinline__ double co_ldb(__amdgpu_buffer_rsrc_t r, size_t off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)(off * sizeof(double)), 0, 16));
}
__device__ __forceinline__ double co_ldd(const double *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void co_std(double *p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// every thread's stores complete, then the workgroup barrier
__device__ __forceinline__ void stores_done_barrier() {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
}

// All threads: wait until *p >= target (thread 0 spins).  False when the front's wait ran out
// (here or in another workgroup of the front); the caller then leaves the kernel.
__device__ __forceinline__ bool chain_wait(const int *p, int target, int *err) {
    __shared__ int s_ok;
    if (threadIdx.x == 0) {
        int ok = 1;
        long long it = 0;
        while (co_ldi(p) < target) {
            if ((++it & 63) == 0 && (it > CHAIN_SPIN_LIMIT || co_ldi(err))) {
                atomicOr(err, 1);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(4);
        }
        s_ok = ok;
    }
    __syncthreads();
    const bool ok = s_ok != 0;
    __syncthreads();  // s_ok is read by all before the next wait writes it
    return ok;
}

// chain_wait on a counter with its aux word: *aux (LDS) receives the aux read in the same 8-byte
// load that saw the count reach target.  With `also`, the same loop reads a second counter (one
// buffer load issued beside each poll, so no extra round trip) and *also_ok (LDS) tells whether it
// had reached also_target: such a count is final, a lower one says nothing.
__device__ __forceinline__ bool chain_wait_aux(ChainCounter &c, int target, int *err, int *aux,
                                               const ChainCounter *also = nullptr, int also_target = 0,
                                               int *also_ok = nullptr) {
    __shared__ int s_ok;
    if (threadIdx.x == 0) {
        int ok = 1, ao = 0;
        long long it = 0;
        unsigned long long w;
        const __amdgpu_buffer_rsrc_t ra =
            __builtin_amdgcn_make_buffer_rsrc(const_cast<ChainCounter *>(also ? also : &c), (short)0, 64, 0x00020000);
        for (;;) {
            const int a = also ? (int)__builtin_amdgcn_raw_buffer_load_b32(ra, 0, 0, 16) : 0;
            w = __hip_atomic_load((unsigned long long *)&c.v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ao = a >= also_target;
            if ((int)(unsigned)(w & 0xffffffffu) >= target) break;
            if ((++it & 63) == 0 && (it > CHAIN_SPIN_LIMIT || co_ldi(err))) {
                atomicOr(err, 1);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(4);
            __asm__ volatile("" ::: "memory");
        }
        *aux = (int)(unsigned)(w >> 32);
        if (also_ok) *also_ok = ao;
        s_ok = ok;
    }
    __syncthreads();
    const bool ok = s_ok != 0;
    __syncthreads();  // s_ok is read by all before the next wait writes it
    return ok;
}

// Symmetric line permutation (apply_swaps) with device-scope accesses to the front.
__device__ __forceinline__ void apply_swaps_co(FrontDev &F, AptpState &S, const int *sa, const int *sb, int nsw, int *U, int *src,
                               int *cnt, int *ptmp) {
    const int tid = threadIdx.x;
    const int m = F.m;
    const size_t lda = m;
    gdouble *A = gbl(F.F);
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
    gint *mark = gbl(S.mark);
    gdouble *buf = gbl(S.swapbuf);
    for (int u = tid; u < nU; u += blockDim.x) mark[U[u]] = u + 1;
    for (int e = tid; e < nU * m; e += blockDim.x) {
        const int ui = e / m, q = e % m;
        const int u = U[ui];
        buf[e] = (u >= q) ? co_ld(A + u + q * lda) : co_ld(A + q + u * lda);
    }
    if (tid < nU) ptmp[tid] = gbl(F.perm)[U[tid]];
    __syncthreads();
    for (int pi = 0; pi < nU; ++pi) {
        const int p = U[pi];
        const int sp = src[pi];
        if (sp == p) continue;
        const int sidx = mark[sp] - 1;
        const gdouble *line = buf + (size_t)sidx * m;
        for (int q = tid; q < m; q += blockDim.x) {
            const int mq = mark[q];
            const int qs = mq ? src[mq - 1] : q;
            const double v = line[qs];
            if (p >= q) co_st(A + p + q * lda, v); else co_st(A + q + p * lda, v);
        }
    }
    if (tid < nU) {
        const int sp = src[tid];
        gbl(F.perm)[U[tid]] = ptmp[mark[sp] - 1];
    }
    __syncthreads();
    for (int u = tid; u < nU; u += blockDim.x) mark[U[u]] = 0;
    stores_done_barrier();
}

struct ChainLeaderLds {
    double blk[MAX_IB][MAX_IB + 1];      // the current block, factored (L11)
    double Ct[MAX_IB][MAX_IB + 1];       // the next block's tile, factored ahead
    double Wr[2 * MAX_IB][MAX_IB + 1];   // next block's L rows and W, then the factor's buffers
    int bp[MAX_IB], bp2[MAX_IB], nbp[MAX_IB], operm[MAX_IB], noperm[MAX_IB];
    int8_t bkind[MAX_IB], nbkind[MAX_IB];
    double bdiag[MAX_IB], boff[MAX_IB], dinv[MAX_IB], nbdiag[MAX_IB], nboff[MAX_IB];
    double wbv[8], wbs[24];
    int wbp[8];
    int sa[MAX_IB], sb[MAX_IB], U[2 * MAX_IB], src[2 * MAX_IB], ptmp[2 * MAX_IB];
    int nsw, cnt, s_end, s_fail, s_upd;
};
constexpr int CHAIN_LC_COLS = 128;  // trailing columns per round of the panel update
struct ChainPanelLds {
    double blk[MAX_IB][MAX_IB + 1];      // L11 of the current block
    double WT[MAX_IB][64 + 1];           // W = L D of a 64-row sub-tile, transposed
    union {                              // one at a time, in this order within a sub-tile:
        double LT[MAX_IB][64 + 1];       //   its L in the passed columns, transposed
        double Lc[MAX_IB][MAX_IB + 1];   //   L rows of the failed block rows
        double LcT[MAX_IB][CHAIN_LC_COLS + 2];  // L rows of a round's columns, transposed
        double Yb[128][MAX_IB + 1];      // (apply) panel rows in column order, by thread
    };
    int bp[MAX_IB];
    int8_t bkind[MAX_IB];
    double bdiag[MAX_IB], boff[MAX_IB], dinv[MAX_IB];
    int rec[4];
    int s_fail;
};
union ChainLds {
    ChainLeaderLds ld;
    ChainPanelLds pn;
};

__device__ __forceinline__ void pivots_dinv(const int8_t *bkind, const double *bdiag, const double *boff,
                                            double *dinv, int bn) {
    const int tid = threadIdx.x;
    if (tid < bn) {
        const int kd = bkind[tid];
        if (kd == 1) dinv[tid] = 1.0 / bdiag[tid];
        else if (kd == 2) dinv[tid] = 1.0 / (bdiag[tid] * bdiag[tid + 1] - boff[tid] * boff[tid]);
    }
}

__device__ __forceinline__ void chain_leader(FrontDev &F, AptpState &S, ChainSync &C, ChainLeaderLds &L, int nit, int NP,
                             int ib, double small) {
    const int tid = threadIdx.x;
    const int m = F.m;
    const size_t lda = m;
    gdouble *A = gbl(F.F);
    gdouble *const bk = gbl(S.backup);
    gint *const perm = gbl(F.perm);
    int k = S.kpos, end_pos = S.end_pos, nib = S.nib;
    int n1 = 0, n2 = 0;
    bool have = false;  // the current block was factored ahead and is in the front
    bool upd_prev = false;  // the panels' updates of the previous block are known complete
    int bn = 0;
    for (int it = 0; it < nit; ++it) {
        const int bs = min(end_pos - k, ib);
        if (!have) {
            // a fresh block: the panels' updates of the previous block are in place
            if (it > 0 && !upd_prev && !chain_wait(&C.updated[it - 1].v, NP, &C.err.v)) return;
            upd_prev = true;
            {
                // every load issued before any is used (one round trip)
                const int i = tid & (MAX_IB - 1), j0 = tid / MAX_IB;
                const __amdgpu_buffer_rsrc_t rb = co_rsrc(A + k + (size_t)k * lda);
                double v[MAX_IB / 8];
#pragma unroll
                for (int u = 0; u < MAX_IB / 8; ++u) {
                    const int j = j0 + 8 * u;
                    const double x = co_ldb(rb, (size_t)min(i, bs > 0 ? bs - 1 : 0) + (size_t)min(j, bs > 0 ? bs - 1 : 0) * lda);
                    v[u] = (i < bs && j <= i) ? x : 0.0;
                }
#pragma unroll
                for (int u = 0; u < MAX_IB / 8; ++u) {
                    const int j = j0 + 8 * u;
                    if (i < bs && j <= i) {
                        L.blk[i][j] = v[u];
                        co_st(bk + i + (size_t)j * lda, v[u]);
                    }
                }
            }
            if (tid < MAX_IB) L.operm[tid] = (tid < bs) ? perm[k + tid] : 0;
            lds_barrier();
            double bml;
            bn = block_factor_wg(L.blk, L.Wr, L.bp, L.bp2, L.bkind, L.bdiag, L.boff, bs, small, &bml, L.wbv, L.wbp,
                                 L.wbs, L.Wr + MAX_IB);
            if (tid < bn && L.bkind[tid] == 2) L.blk[tid + 1][tid] = 0.0;
            lds_barrier();
            if (bn > 0) {
                const int i = tid & 31;
                for (int j = tid >> 5; j < bs; j += 8)
                    if (i < bs && j <= i) co_st(A + (k + i) + (size_t)(k + j) * lda, L.blk[i][j]);
                if (tid < bs) perm[k + tid] = L.operm[L.bp[tid]];
            }
            if (tid == 0) atomic_max_nonneg(&F.max_l_bits, bml);
        }
        pivots_dinv(L.bkind, L.bdiag, L.boff, L.dinv, bn);
        // the next block ahead: its L rows (the panel substitution of rows be..be+nbs-1), its tile
        // after this block's update, and its factorization.  Its operands are loaded before the
        // record goes out: the panels overwrite those rows' block entries once they have it.
        const int be = k + bs, nbs = min(end_pos - be, ib);
        const bool ahead = bs > 0 && bn == bs && nbs > 0 && it + 1 < nit;
        double y[MAX_IB], cv[MAX_IB / 8];
        if (ahead) {
            if (it > 0 && !upd_prev && !chain_wait(&C.updated[it - 1].v, NP, &C.err.v)) return;
            const bool live = tid < nbs;
            {
                const __amdgpu_buffer_rsrc_t rb = co_rsrc(A + be + (size_t)k * lda);
                const size_t ro = (size_t)(live ? tid : 0);
#pragma unroll
                for (int j = 0; j < MAX_IB; ++j) {
                    const double x = co_ldb(rb, ro + (size_t)L.bp[j < bs ? j : 0] * lda);
                    y[j] = (live && j < bs) ? x : 0.0;
                }
            }
            const int ii = tid & 31, jg = tid >> 5;
            {
                const __amdgpu_buffer_rsrc_t rb = co_rsrc(A + be + (size_t)be * lda);
#pragma unroll
                for (int u = 0; u < MAX_IB / 8; ++u) {
                    const int j = jg + 8 * u;
                    const double x = co_ldb(rb, (size_t)min(ii, nbs - 1) + (size_t)min(j, nbs - 1) * lda);
                    cv[u] = (ii < nbs && j <= ii) ? x : 0.0;
                }
            }
            if (tid < MAX_IB) L.noperm[tid] = (tid < nbs) ? perm[be + tid] : 0;
        }
        // record it (the barrier also completes the loads above)
        ChainRec &R = C.recs[it];
        if (tid < MAX_IB) {
            co_sti(&R.bperm[tid], L.bp[tid]);
            co_sti(&R.bkind[tid], tid < bn ? L.bkind[tid] : 0);
            co_std(&R.bdiag[tid], L.bdiag[tid]);
            co_std(&R.boff[tid], L.boff[tid]);
        }
        if (tid == 0) {
            co_sti(&R.k, k);
            co_sti(&R.bs, bs);
            co_sti(&R.bn, bn);
            co_sti(&R.end_pos, end_pos);
        }
        stores_done_barrier();
        if (tid == 0) co_sti(&C.rec.v, it + 1);
        int bn2 = 0;
        double bml2 = 0.0, nv[MAX_IB / 8];
        if (ahead) {
            double (*Lr)[MAX_IB + 1] = L.Wr;
            double (*Wt)[MAX_IB + 1] = L.Wr + MAX_IB;
            if (tid < 64) {
                // (wavefront 0 whole: rows beyond the block hold zeros)
                panel_trsm_d(y, L.blk, bn, L.bkind, L.dinv, L.bdiag, L.boff);
                if (tid < MAX_IB) {
#pragma unroll
                    for (int j = 0; j < MAX_IB; ++j) Lr[tid][j] = y[j];
                }
            }
            {
                const int ii = tid & 31, jg = tid >> 5;
#pragma unroll
                for (int u = 0; u < MAX_IB / 8; ++u) L.Ct[ii][jg + 8 * u] = cv[u];
            }
            lds_barrier();
            next_tile_update(Lr, Wt, L.Ct, L.bkind, L.bdiag, L.boff, nbs, bs, nv);
            bn2 = block_factor_wg(L.Ct, L.Wr, L.nbp, L.bp2, L.nbkind, L.nbdiag, L.nboff, nbs, small, &bml2, L.wbv,
                                  L.wbp, L.wbs, L.Wr + MAX_IB);
            if (tid < bn2 && L.nbkind[tid] == 2) L.Ct[tid + 1][tid] = 0.0;
            lds_barrier();
        }
        // the check of block it; the panels' update count of the block comes along (complete, it
        // saves the wait before the next block's operands are loaded)
        if (!chain_wait_aux(C.applied[it], NP, &C.err.v, &L.s_fail, &C.updated[it], NP, &L.s_upd)) return;
        const bool upd_done = L.s_upd != 0;
        upd_prev = upd_done;
        int nelim = 0;
        if (bn > 0) {
            nelim = min(L.s_fail, bn);
            if (nelim > 0 && L.bkind[nelim - 1] == 2) nelim -= 1;
        }
        // D and the pivot counts of the accepted columns (k_block_finish)
        int is1 = 0, is2 = 0;
        if (tid < nelim) {
            const int kd = L.bkind[tid];
            const int c = k + tid;
            if (kd == 1) {
                gbl(F.dkind)[c] = 1; gbl(F.ddiag)[c] = L.bdiag[tid]; gbl(F.doff)[c] = 0.0;
                is1 = 1;
            } else if (kd == 2 && tid + 1 < nelim) {
                gbl(F.dkind)[c] = 2; gbl(F.dkind)[c + 1] = 3;
                gbl(F.ddiag)[c] = L.bdiag[tid]; gbl(F.ddiag)[c + 1] = L.bdiag[tid + 1];
                gbl(F.doff)[c] = L.boff[tid]; gbl(F.doff)[c + 1] = 0.0;
                is2 = 1;
            }
        }
        n1 += __syncthreads_count(is1);
        n2 += __syncthreads_count(is2);
        if (nib < MAX_OB_IB && tid == 0) S.ib_ne[nib] = (int8_t)nelim;
        ++nib;
        if (ahead && nelim == bs) {
            // keep the block factored ahead: the panels left its tile to this workgroup
            {
                const int ii = tid & 31, jg = tid >> 5;
#pragma unroll
                for (int u = 0; u < MAX_IB / 8; ++u) {
                    const int j = jg + 8 * u;
                    if (ii < nbs && j <= ii) {
                        co_st(A + (be + ii) + (size_t)(be + j) * lda, L.Ct[ii][j]);
                        co_st(bk + ii + (size_t)j * lda, nv[u]);
                    }
                }
            }
            if (bn2 > 0 && tid < nbs) perm[be + tid] = L.noperm[L.nbp[tid]];
            if (tid == 0) atomic_max_nonneg(&F.max_l_bits, bml2);
            for (int e = tid; e < MAX_IB * MAX_IB; e += 256) {
                const int i = e & (MAX_IB - 1), j = e / MAX_IB;
                L.blk[i][j] = L.Ct[i][j];
            }
            if (tid < MAX_IB) {
                L.bp[tid] = L.nbp[tid];
                L.bkind[tid] = (tid < bn2) ? L.nbkind[tid] : 0;
                L.bdiag[tid] = L.nbdiag[tid];
                L.boff[tid] = L.nboff[tid];
            }
            lds_barrier();
            bn = bn2;
            k = be;
            have = true;
        } else {
            have = false;
            if (bn == 0 || nelim < bs) {
                // failed columns (the whole block when nothing was eliminated) move to the end of
                // the outer block, after every panel's restore and cross terms
                if (!upd_done && !chain_wait(&C.updated[it].v, NP, &C.err.v)) return;
                upd_prev = true;
                if (tid == 0) {
                    int n = 0;
                    if (bn == 0) {
                        for (int off = 0; off < bs; ++off) {
                            end_pos -= 1;
                            if (k + off < end_pos) { L.sa[n] = k + off; L.sb[n] = end_pos; ++n; }
                        }
                    } else {
                        for (int i = bs - nelim - 1; i >= 0; --i) {
                            const int dp = k + nelim + i;
                            end_pos -= 1;
                            if (dp < end_pos) { L.sa[n] = dp; L.sb[n] = end_pos; ++n; }
                        }
                    }
                    L.nsw = n;
                    L.s_end = end_pos;
                }
                __syncthreads();
                end_pos = L.s_end;
                apply_swaps_co(F, S, L.sa, L.sb, L.nsw, L.U, L.src, &L.cnt, L.ptmp);
            }
            k += nelim;
        }
    }
    if (tid == 0) {
        S.kpos = k;
        S.end_pos = end_pos;
        S.nib = nib;
        S.fail_min = INT_MAX;
        S.pref = 0;
        S.nx_valid = 0;
        F.n1 += n1;
        F.n2 += n2;
    }
}

// Panel rows of chunk c: [ob_start + c * CH, min(ob_start + (c + 1) * CH, m)).
template <int CH>
__device__ __forceinline__ void chain_panel(FrontDev &F, AptpState &S, ChainSync &C, ChainPanelLds &L, int pidx, int NP, int nit,
                            int ib, double threshold) {
    static_assert(CH == 64 || CH == 128, "panel rows per workgroup: 64 or 128 (the apply's LDS rows)");

    const int tid = threadIdx.x;
    const int m = F.m;
    const size_t lda = m;
    gdouble *A = gbl(F.F);
    gdouble *const bk = gbl(S.backup);
    const int ob_start = S.ob_start, obe = S.ob_start + S.ob_cols;
    const int nchunk = (m - ob_start + CH - 1) / CH;
    const double bound = 1.0 / threshold;
    for (int it = 0; it < nit; ++it) {
        if (!chain_wait(&C.rec.v, it + 1, &C.err.v)) return;
        const ChainRec &R = C.recs[it];
        if (tid < MAX_IB) {
            L.bp[tid] = co_ldi(&R.bperm[tid]);
            L.bkind[tid] = (int8_t)co_ldi(&R.bkind[tid]);
            L.bdiag[tid] = co_ldd(&R.bdiag[tid]);
            L.boff[tid] = co_ldd(&R.boff[tid]);
        }
        if (tid == 0) {
            L.rec[0] = co_ldi(&R.k);
            L.rec[1] = co_ldi(&R.bs);
            L.rec[2] = co_ldi(&R.bn);
            L.rec[3] = co_ldi(&R.end_pos);
        }
        __syncthreads();
        const int k = __builtin_amdgcn_readfirstlane(L.rec[0]), bs = __builtin_amdgcn_readfirstlane(L.rec[1]);
        const int bn = __builtin_amdgcn_readfirstlane(L.rec[2]), end_pos = __builtin_amdgcn_readfirstlane(L.rec[3]);
        const int be = k + bs;
        if (bn > 0) {
            {
                const int i = tid & (MAX_IB - 1), j0 = tid / MAX_IB;
                const __amdgpu_buffer_rsrc_t rb = co_rsrc(A + k + (size_t)k * lda);
                double v[MAX_IB / 8];
#pragma unroll
                for (int u = 0; u < MAX_IB / 8; ++u) {
                    const int j = j0 + 8 * u;
                    const double x = co_ldb(rb, (size_t)min(i, bs - 1) + (size_t)min(j, bs - 1) * lda);
                    v[u] = (i < bs && j < i) ? x : 0.0;
                }
#pragma unroll
                for (int u = 0; u < MAX_IB / 8; ++u) L.blk[i][j0 + 8 * u] = v[u];
            }
            pivots_dinv(L.bkind, L.bdiag, L.boff, L.dinv, bn);
            lds_barrier();
            for (int c = pidx; c < nchunk; c += NP) {
                const int r0 = ob_start + c * CH, r1 = min(r0 + CH, m);
                if (r1 <= be || tid >= CH) continue;
                // (wavefront-uniform: CH is a multiple of 64)
                const int r = r0 + tid;
                const bool live = r >= be && r < r1;
                double y[MAX_IB];
                {
                    // the row's block entries in column order (one base, constant offsets; rows
                    // clamped into the front, unused values discarded), backed up as they are
                    // (restore_diagonal_with_perm reads the backup in column order), then
                    // gathered into pivot order through the wavefront's own LDS rows
                    const __amdgpu_buffer_rsrc_t rb = co_rsrc(A + (size_t)k * lda);
                    const size_t ro = (size_t)min(r, m - 1);
                    double yn[MAX_IB];
#pragma unroll
                    for (int j = 0; j < MAX_IB; ++j) {
                        const double v = co_ldb(rb, ro + (size_t)j * lda);
                        yn[j] = (live && j < bs) ? v : 0.0;
                    }
                    if (live) {
                        gdouble *brow = bk + (r - k);
                        if (bs == MAX_IB) {
#pragma unroll
                            for (int j = 0; j < MAX_IB; ++j) co_st(brow + (size_t)j * lda, yn[j]);
                        } else {
#pragma unroll
                            for (int j = 0; j < MAX_IB; ++j)
                                if (j < bs) co_st(brow + (size_t)j * lda, yn[j]);
                        }
                    }
                    double *yb = L.Yb[tid];
#pragma unroll
                    for (int j = 0; j < MAX_IB; ++j) yb[j] = yn[j];
                    // (the lane reads back only what it wrote: LDS operations of a wavefront
                    // are performed in order)
#pragma unroll
                    for (int j = 0; j < MAX_IB; ++j) y[j] = yb[L.bp[j]];
                }
                panel_trsm_d(y, L.blk, bn, L.bkind, L.dinv, L.bdiag, L.boff);
                int first = INT_MAX;
                if (live) {
                    gdouble *row = A + r;
                    if (bs == MAX_IB) {
#pragma unroll
                        for (int j = 0; j < MAX_IB; ++j) co_st(row + (size_t)(k + j) * lda, y[j]);
                    } else {
#pragma unroll
                        for (int j = 0; j < MAX_IB; ++j)
                            if (j < bs) co_st(row + (size_t)(k + j) * lda, y[j]);
                    }
#pragma unroll
                    for (int q = MAX_IB - 1; q >= 0; --q)
                        if (q < bn && fabs(y[q]) >= bound) first = q;
                    if (first != INT_MAX && L.bkind[first] == 3) first -= 1;
                }
                first = wave_min_int(first);
                if ((tid & 63) == 0 && first != INT_MAX) atomicMin(&C.applied[it].aux, first);
            }
        }
        stores_done_barrier();
        if (tid == 0) __hip_atomic_fetch_add(&C.applied[it].v, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!chain_wait_aux(C.applied[it], NP, &C.err.v, &L.s_fail)) return;
        if (bn > 0) {
            int nelim = min(L.s_fail, bn);
            if (nelim > 0 && L.bkind[nelim - 1] == 2) nelim -= 1;
            nelim = __builtin_amdgcn_readfirstlane(nelim);
            const int nbs = min(end_pos - be, ib);
            const bool ahead = nelim == bs && nbs > 0 && it + 1 < nit;
            const int nf = bs - nelim;
            double ml = 0.0;
            // 64-row sub-tiles of the panel's rows at or below the block
            for (int c = pidx; c < nchunk; c += NP) {
                const int r0c = ob_start + c * CH, r1 = min(r0c + CH, m);
                for (int R0 = r0c; R0 < r1; R0 += 64) {
                    if (R0 + 64 <= k) continue;  // (uniform)
                    __syncthreads();             // the previous sub-tile's LDS reads are done
                    // L rows of the sub-tile in the passed columns -> LT[j][row]; W = L D
                    // (ld_entry, compute_ld_into factor.rs:1536-1563) -> WT[j][row]
                    {
                        const int rr = tid & 63, jg = tid >> 6, r = R0 + rr;
                        const bool valid = r >= k && r < r1;
                        double lv[MAX_IB / 4];
                        const __amdgpu_buffer_rsrc_t rb = co_rsrc(A + (size_t)k * lda);
                        const size_t ro = (size_t)min(r, m - 1);
#pragma unroll
                        for (int u = 0; u < MAX_IB / 4; ++u) {
                            const int j = jg + 4 * u;
                            const double v = co_ldb(rb, ro + (size_t)(j < nelim ? j : 0) * lda);
                            lv[u] = (valid && j < nelim) ? v : 0.0;
                        }
#pragma unroll
                        for (int u = 0; u < MAX_IB / 4; ++u) {
                            L.LT[jg + 4 * u][rr] = lv[u];
                            // max |L| over the passed columns of the rows below the block
                            if (valid && r >= be && fabs(lv[u]) > ml) ml = fabs(lv[u]);
                        }
                    }
                    __syncthreads();
                    {
                        const int rr = tid & 63, jg = tid >> 6;
#pragma unroll
                        for (int u = 0; u < MAX_IB / 4; ++u) {
                            const int j = jg + 4 * u;
                            // the three forms side by side, one selected
                            const int kd = L.bkind[j];
                            const double dj = L.bdiag[j];
                            const double lj = L.LT[j][rr];
                            const double v1 = lj * dj;
                            const double v2 = lj * dj + L.LT[j + 1 < MAX_IB ? j + 1 : j][rr] * L.boff[j];
                            const double v3 = L.LT[j > 0 ? j - 1 : 0][rr] * L.boff[j > 0 ? j - 1 : 0] + lj * dj;
                            const double v = (kd == 1) ? v1 : (kd == 2) ? v2 : (kd == 3) ? v3 : 0.0;
                            L.WT[j][rr] = (j < nelim) ? v : 0.0;
                        }
                    }
                    __syncthreads();
                    // partial restore and cross terms of the failed columns (factor.rs:1962-2008)
                    if (nf > 0) {
                        {
                            const int jf = tid & 31, c0 = tid >> 5;
                            double v[MAX_IB / 8];
#pragma unroll
                            for (int u = 0; u < MAX_IB / 8; ++u) {
                                const int cc = c0 + 8 * u;
                                v[u] = (jf < nf && cc < nelim) ? co_ld(A + (k + nelim + jf) + (size_t)(k + cc) * lda) : 0.0;
                            }
#pragma unroll
                            for (int u = 0; u < MAX_IB / 8; ++u) L.Lc[jf][c0 + 8 * u] = v[u];
                        }
                        __syncthreads();
                        const int rr = tid & 63, g = tid >> 6, r = R0 + rr;
                        if (r >= k && r < r1) {
                            gdouble *row = A + r;
                            for (int j = nelim + g; j < bs; j += 4) {
                                const size_t cj = (size_t)(k + j) * lda;
                                double v = 0.0;
                                bool have_v = false;
                                if (nelim < bn) {
                                    if (r < be) {
                                        const int i = r - k;
                                        if (i >= j) {
                                            const int br = L.bp[i], bc = L.bp[j];
                                            v = (br >= bc) ? co_ld(bk + br + (size_t)bc * lda) : co_ld(bk + bc + (size_t)br * lda);
                                            have_v = true;
                                        }
                                    } else {
                                        v = co_ld(bk + (r - k) + (size_t)L.bp[j] * lda);
                                        have_v = true;
                                    }
                                }
                                if (nelim > 0) {
                                    const int jf = j - nelim;
                                    if (r >= be || r - k - nelim >= jf) {
                                        double sx = 0.0;
                                        for (int c2 = 0; c2 < nelim; ++c2) sx += L.WT[c2][rr] * L.Lc[jf][c2];
                                        if (!have_v) v = co_ld(row + cj);
                                        v = v - sx;
                                        have_v = true;
                                    }
                                }
                                if (have_v) co_st(row + cj, v);
                            }
                        }
                    }
                    // trailing columns of the outer block, CHAIN_LC_COLS per round: thread (tx, ty)
                    // forms rows R0 + tx + 16 i (i < 4) x columns c0 + 4 ty + q + 64 p (q < 4,
                    // p < 2), each sum over j ascending from zero, subtracted once; the round's
                    // L rows staged transposed in LDS
                    if (nelim == 0 || R0 + 63 < be) continue;  // (uniform)
                    const int tx = tid & 15, ty = tid >> 4;
                    for (int c0 = be; c0 < obe; c0 += CHAIN_LC_COLS) {
                        const int ncol = min(CHAIN_LC_COLS, obe - c0);
                        if (R0 + 63 < c0) break;  // (uniform) no row of the sub-tile reaches them
                        __syncthreads();
                        {
                            constexpr int NL = CHAIN_LC_COLS * MAX_IB / 256;
                            double v[NL];
                            const __amdgpu_buffer_rsrc_t rb = co_rsrc(A + c0 + (size_t)k * lda);
#pragma unroll
                            for (int u = 0; u < NL; ++u) {
                                const int e = tid + 256 * u, cc = e % CHAIN_LC_COLS, j = e / CHAIN_LC_COLS;
                                const double x = co_ldb(rb, (size_t)(cc < ncol ? cc : 0) + (size_t)(j < nelim ? j : 0) * lda);
                                v[u] = (cc < ncol && j < nelim) ? x : 0.0;
                            }
#pragma unroll
                            for (int u = 0; u < NL; ++u) {
                                const int e = tid + 256 * u;
                                L.LcT[e / CHAIN_LC_COLS][e % CHAIN_LC_COLS] = v[u];
                            }
                        }
                        __syncthreads();
                        double acc[4][8];
#pragma unroll
                        for (int i = 0; i < 4; ++i)
#pragma unroll
                            for (int q = 0; q < 8; ++q) acc[i][q] = 0.0;
                        if (nelim == MAX_IB) {
#pragma unroll 4
                            for (int j = 0; j < MAX_IB; ++j) {
                                double wv[4], lc[8];
#pragma unroll
                                for (int i = 0; i < 4; ++i) wv[i] = L.WT[j][tx + 16 * i];
#pragma unroll
                                for (int q = 0; q < 8; ++q) lc[q] = L.LcT[j][4 * ty + (q & 3) + 64 * (q >> 2)];
#pragma unroll
                                for (int i = 0; i < 4; ++i)
#pragma unroll
                                    for (int q = 0; q < 8; ++q) acc[i][q] += wv[i] * lc[q];
                            }
                        } else {
                            for (int j = 0; j < nelim; ++j) {
                                double wv[4], lc[8];
#pragma unroll
                                for (int i = 0; i < 4; ++i) wv[i] = L.WT[j][tx + 16 * i];
#pragma unroll
                                for (int q = 0; q < 8; ++q) lc[q] = L.LcT[j][4 * ty + (q & 3) + 64 * (q >> 2)];
#pragma unroll
                                for (int i = 0; i < 4; ++i)
#pragma unroll
                                    for (int q = 0; q < 8; ++q) acc[i][q] += wv[i] * lc[q];
                            }
                        }
                        const __amdgpu_buffer_rsrc_t rbc = co_rsrc(A + R0 + (size_t)c0 * lda);
#pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            const int r = R0 + tx + 16 * i;
                            gdouble *row = A + r;
                            double cv[8];
                            bool ok[8];
#pragma unroll
                            for (int q = 0; q < 8; ++q) {
                                const int cc = 4 * ty + (q & 3) + 64 * (q >> 2), col = c0 + cc;
                                const bool tile_ahead = ahead && r < be + nbs && col < be + nbs;
                                ok[q] = r >= be && r < r1 && cc < ncol && r >= col && !tile_ahead;
                                const double x = co_ldb(rbc, (size_t)(min(r, m - 1) - R0) + (size_t)(cc < ncol ? cc : 0) * lda);
                                cv[q] = ok[q] ? x : 0.0;
                            }
#pragma unroll
                            for (int q = 0; q < 8; ++q) {
                                const int col = c0 + 4 * ty + (q & 3) + 64 * (q >> 2);
                                if (ok[q]) co_st(row + (size_t)col * lda, cv[q] - acc[i][q]);
                            }
                        }
                    }
                }
            }
            ml = wave_fmax(ml);
            if ((tid & 63) == 0) atomic_max_nonneg(&F.max_l_bits, ml);
            // block rows permuted in the columns factored before this block, this panel's share
            {
                const int per = (k + NP - 1) / NP;
                const int cb = pidx * per, ce = min(k, cb + per);
                const int i = tid & 31, half = (tid >> 5) & 1, wv = tid >> 6;
                if (i < bs) {
                    const int sr = L.bp[i];
                    // 128 columns per step, all loads issued before the stores
                    for (int cs = cb; cs < ce; cs += 128) {
                        const __amdgpu_buffer_rsrc_t rb = co_rsrc(A + k + (size_t)cs * lda);
                        double v[16];
#pragma unroll
                        for (int u = 0; u < 16; ++u) {
                            const int cc = 2 * wv + half + 8 * u;  // column cs + cc
                            const double x = co_ldb(rb, (size_t)sr + (size_t)(cs + cc < ce ? cc : 0) * lda);
                            v[u] = x;
                        }
#pragma unroll
                        for (int u = 0; u < 16; ++u) {
                            const int c = cs + 2 * wv + half + 8 * u;
                            if (c < ce) co_st(A + (k + i) + (size_t)c * lda, v[u]);
                        }
                    }
                }
            }
        }
        stores_done_barrier();
        if (tid == 0) __hip_atomic_fetch_add(&C.updated[it].v, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// One outer block of every listed front: item (APTP slot, role, front, NP); role 0 = leader,
// role 1 + p = panel p of NP.
template <int CH>
__global__ __launch_bounds__(256, 2) void k_aptp_outer(FrontDev *__restrict__ fr, AptpState *__restrict__ st,
                                                    ChainSync *__restrict__ cs, const int4 *__restrict__ items,
                                                    int nb, int ib, double small, double threshold) {
    __shared__ ChainLds lds;
    const int4 itm = items[blockIdx.x];
    FrontDev &F = fr[itm.z];
    AptpState &S = st[itm.x];
    ChainSync &C = cs[itm.x];
    const int nit = (S.ob_cols + ib - 1) / ib;
    if (itm.y == 0) {
        chain_leader(F, S, C, lds.ld, nit, itm.w, ib, small);
        if (threadIdx.x == 0 && co_ldi(&C.err.v)) atomicOr(&F.flags, FRONT_CHAIN_ERR);
    } else {
        chain_panel<CH>(F, S, C, lds.pn, itm.y - 1, itm.w, nit, ib, threshold);
    }
}

// Start of an outer block (two_level_factor, factor.rs:2126-2131); ob == 0 also initialises.
__device__ void outer_begin(FrontDev &F, AptpState &S, int ob, int nb, int ib) {
    const int p = F.k;
    const bool two = p > nb;
    const int nouter = two ? (p + nb - 1) / nb : 1;
    if (ob >= nouter) return;
    if (threadIdx.x == 0) {
        if (ob == 0) { S.gne = 0; S.remaining = p; }
        S.ob_start = S.gne;
        S.ob_cols = two ? min(S.remaining, nb) : p;
        S.kpos = S.ob_start;
        S.end_pos = S.ob_start + S.ob_cols;
        S.nib = 0;
        S.pref = 0;
        S.nx_valid = 0;
    }
    __syncthreads();
    copy_next_block(F, S, ib);
}

// End of an outer block: failed columns move to the end of the remaining fully-summed
// region (factor.rs:2195-2206), counters advance; the last block records ne.  Returns false
// when the front has no outer block ob or only one outer block (then it has no block ob + 1
// either); the return is the same for every thread of the workgroup.
__device__ bool outer_end(FrontDev &F, AptpState &S, int ob, int nb, int la) {
    const bool deferred = !(F.flags & FRONT_ALLFS);
    const int p = F.k;
    const bool two = p > nb;
    const int nouter = two ? (p + nb - 1) / nb : 1;
    if (ob >= nouter) return false;
    __shared__ int sa[MAX_IB * 8], sb[MAX_IB * 8];
    __shared__ int U[2 * MAX_IB], src[2 * MAX_IB], ptmp[2 * MAX_IB];
    __shared__ int nsw, cnt;
    const int tid = threadIdx.x;
    if (!two) {
        if (tid == 0) F.ne = S.kpos;
        return false;
    }
    const int obne = S.kpos - S.ob_start;
    const int nfail = S.ob_cols - obne;
    if (nfail > 0 && tid == 0) {
        F.obfail |= 1ull << min(ob, 63);
        // the trailing update of this block was planned to run beside the next block, which
        // is only correct without the delay swaps below: the factorization is repeated
        if (la) atomicOr(&F.flags, FRONT_LA_ERR);
    }
    if (deferred && tid == 0) {
        // algorithmic flops of this block's outer update (k_schur mode 1): 2 obne per entry of
        // the lower trapezoid rows [r0, m) x columns [r0, p)
        const double r0 = S.ob_start + S.ob_cols, nr = F.m - r0, nc = p - r0;
        if (obne > 0 && nc > 0) F.uflops += 2.0 * obne * (nc * (nc + 1.0) / 2.0 + (nr - nc) * nc);
    }
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
    return true;
}

// The first outer block of every front of a level: the APTP state is installed from its staged
// initial value (scratch pointers, fail_min = INT_MAX, the rest zero) and the front's swap marks
// are cleared, then outer block 0 starts.
__global__ __launch_bounds__(256) void k_outer_first(FrontDev *__restrict__ fr,
                                                     AptpState *__restrict__ st,
                                                     const AptpState *__restrict__ init,
                                                     const int *__restrict__ flist, int nb, int ib,
                                                     ChainSync *__restrict__ cs) {
    if (cs) chain_reset(cs[blockIdx.x]);
    FrontDev &F = fr[flist[blockIdx.x]];
    AptpState &S = st[blockIdx.x];
    const AptpState &I = init[blockIdx.x];
    static_assert(sizeof(AptpState) % sizeof(int) == 0, "AptpState copied as ints");
    const int *src = (const int *)&I;
    int *dst = (int *)&S;
    for (int i = threadIdx.x; i < (int)(sizeof(AptpState) / sizeof(int)); i += blockDim.x) dst[i] = src[i];
    gint *mark = gbl(I.mark);
    for (int i = threadIdx.x; i < F.m; i += blockDim.x) mark[i] = 0;
    __syncthreads();
    outer_begin(F, S, 0, nb, ib);
}

__global__ __launch_bounds__(256) void k_outer_end(FrontDev *__restrict__ fr,
                                                   AptpState *__restrict__ st,
                                                   const int *__restrict__ flist, int ob, int nb) {
    outer_end(fr[flist[blockIdx.x]], st[blockIdx.x], ob, nb, 0);
}

// The end of outer block ob and the start of ob + 1 in one launch (one workgroup per front).
// la: block ob's trailing update runs beside this launch and block ob + 1 (planned without failed
// columns in block ob).
__global__ __launch_bounds__(256) void k_outer_next(FrontDev *__restrict__ fr,
                                                    AptpState *__restrict__ st,
                                                    const int *__restrict__ flist, int ob, int nb,
                                                    int ib, ChainSync *__restrict__ cs, int la) {
    if (cs) chain_reset(cs[blockIdx.x]);
    FrontDev &F = fr[flist[blockIdx.x]];
    AptpState &S = st[blockIdx.x];
    if (!outer_end(F, S, ob, nb, la)) return;
    __syncthreads();
    outer_begin(F, S, ob + 1, nb, ib);
}

// ===========================================================================
// Threshold partial pivoting (tpp_factor / tpp_factor_rect, factor.rs:2268-2947).
// One wavefront per front; every lane follows the same control flow.
// mode 0: whole front from column 0; mode 1: fallback from F.ne.
// ===========================================================================
// Raw buffer loads (gfx9 resource: stride 0, range-checked against num_records): an offset at or
// beyond the range loads zero without touching memory.
constexpr unsigned BUF_RANGE = 0x7FFFFFF0u;
constexpr unsigned BUF_OOB = 0x7FFFFFF8u;  // outside every resource made by buf_rsrc
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *base) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)BUF_RANGE, 0x00020000);
}
__device__ __forceinline__ float buf_load_f32(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}
__device__ __forceinline__ double buf_load_f64(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}

constexpr int TPP_B = 4;  // rows per lane whose loads are issued together (a front below 256 rows: one batch)
// TJ (template parameter of the update routines): columns of a trailing update per memory round
// trip; more columns, fewer round trips, more registers (fewer wavefronts per SIMD)

// Reductions over the NT threads of a workgroup (NT a multiple of 64; every thread calls them in
// the same order): the wavefront result, then the wavefronts' results through LDS in wavefront
// order.  Called in uniform control flow only.
template <int NT>
__device__ __forceinline__ double blk_fmax(double v) {
    v = wave_fmax(v);
    if constexpr (NT == 64) {
        return v;
    } else {
        __shared__ double s_[NT / 64];
        __syncthreads();
        if ((threadIdx.x & 63) == 0) s_[threadIdx.x >> 6] = v;
        __syncthreads();
        double r = s_[0];
#pragma unroll
        for (int w = 1; w < NT / 64; ++w) r = fmax(r, s_[w]);
        return r;
    }
}
template <int NT>
__device__ __forceinline__ bool blk_any(bool b) {
    if constexpr (NT == 64) {
        return __any(b);
    } else {
        return __syncthreads_or(b) != 0;
    }
}
// larger value wins; equal values -> smaller index
template <int NT>
__device__ __forceinline__ void blk_argmax(double &v, int &idx) {
    wave_argmax(v, idx);
    if constexpr (NT > 64) {
        __shared__ double sv_[NT / 64];
        __shared__ int si_[NT / 64];
        __syncthreads();
        if ((threadIdx.x & 63) == 0) { sv_[threadIdx.x >> 6] = v; si_[threadIdx.x >> 6] = idx; }
        __syncthreads();
        v = sv_[0];
        idx = si_[0];
#pragma unroll
        for (int w = 1; w < NT / 64; ++w)
            if (sv_[w] > v || (sv_[w] == v && si_[w] < idx)) { v = sv_[w]; idx = si_[w]; }
    }
}

// A: the front (its first n columns are read and written).
template <class P, int NT = 64>
struct TppCtx {
    P A;
    size_t lda;
    int m, n;  // rows, fully-summed columns
    int lane;
};

template <class P, int NT>
__device__ __forceinline__ double tpp_sym(const TppCtx<P, NT> &c, int r, int q) {
    return (r >= q) ? c.A[r + q * c.lda] : c.A[q + r * c.lda];
}

// Column idx from row `from` (its row part [from, idx), then its column part [idx, m)) holds no
// entry of magnitude >= small.  Both parts go through one batch of loads per pass.
template <class P, int NT>
__device__ bool tpp_col_small(const TppCtx<P, NT> &c, int idx, int from, double small) {
    bool big = false;
    const int nrow = idx - from, tot = nrow + (c.m - idx);
    for (int e0 = c.lane; e0 < tot; e0 += NT * TPP_B) {
        double v[TPP_B];
#pragma unroll
        for (int t = 0; t < TPP_B; ++t) {
            const int e = e0 + NT * t;
            const size_t off = (e < nrow) ? idx + (size_t)(from + e) * c.lda : (size_t)(idx + e - nrow) + idx * c.lda;
            v[t] = (e < tot) ? fabs(c.A[off]) : 0.0;
        }
#pragma unroll
        for (int t = 0; t < TPP_B; ++t)
            if (e0 + NT * t < tot && v[t] >= small) big = true;
    }
    return !blk_any<NT>(big);
}

template <class P, int NT>
__device__ int tpp_row_abs_max(const TppCtx<P, NT> &c, int p, int from, int to) {
    if (from >= to) return from;
    const double f0 = fabs(tpp_sym(c, p, from));
    double bv = -1.0;
    int bi = INT_MAX;
    for (int q0 = from + 1 + c.lane; q0 < to; q0 += NT * TPP_B) {
        double v[TPP_B];
#pragma unroll
        for (int t = 0; t < TPP_B; ++t) v[t] = (q0 + NT * t < to) ? fabs(tpp_sym(c, p, q0 + NT * t)) : 0.0;
#pragma unroll
        for (int t = 0; t < TPP_B; ++t)  // ascending q per lane, as the one-by-one scan
            if (q0 + NT * t < to && v[t] > bv) { bv = v[t]; bi = q0 + NT * t; }
    }
    blk_argmax<NT>(bv, bi);
    if (f0 != f0) return from;
    return (bi != INT_MAX && bv > f0) ? bi : from;
}

// max |entry| of row/column col from column nelim (row part [nelim, col), column part (col, m)),
// index `exclude` left out; both parts go through one batch of loads per pass.
template <class P, int NT>
__device__ double tpp_rc_max_excl(const TppCtx<P, NT> &c, int col, int nelim, int exclude) {
    double best = 0.0;
    const int nrow = col - nelim, tot = nrow + (c.m - col - 1);
    for (int e0 = c.lane; e0 < tot; e0 += NT * TPP_B) {
        double v[TPP_B];
        int q[TPP_B];
#pragma unroll
        for (int t = 0; t < TPP_B; ++t) {
            const int e = e0 + NT * t;
            q[t] = (e < nrow) ? nelim + e : col + 1 + (e - nrow);  // the entry's other index
            const size_t off = (e < nrow) ? col + (size_t)q[t] * c.lda : (size_t)q[t] + col * c.lda;
            v[t] = (e < tot) ? fabs(c.A[off]) : 0.0;
        }
#pragma unroll
        for (int t = 0; t < TPP_B; ++t)
            if (e0 + NT * t < tot && q[t] != exclude) best = fmax(best, v[t]);
    }
    return blk_fmax<NT>(best);
}

template <class P, int NT>
__device__ bool tpp_test_2x2(const TppCtx<P, NT> &c, int t, int p, double maxt, double maxp, double u,
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

template <class P, int NT>
__device__ void tpp_swap(const TppCtx<P, NT> &c, gint *perm, int i, int j) {
    if (i == j) return;
    if (i > j) { const int t = i; i = j; j = t; }
    P A = c.A;
    const size_t lda = c.lda;
    // the pairs exchanged for line q (row parts left of i, between i and j, column parts below j)
    // are disjoint: every load of a pass is issued before its stores
    for (int q0 = c.lane; q0 < c.m; q0 += NT * TPP_B) {
        size_t a1[TPP_B], a2[TPP_B];
        double x1[TPP_B], x2[TPP_B];
#pragma unroll
        for (int t = 0; t < TPP_B; ++t) {
            const int q = q0 + NT * t;
            const bool ok = q < c.m && q != i && q != j;
            a1[t] = !ok ? i + (size_t)i * lda : (q < i) ? i + (size_t)q * lda : (size_t)q + i * lda;
            a2[t] = !ok ? i + (size_t)i * lda : (q < j) ? j + (size_t)q * lda : (size_t)q + j * lda;
            x1[t] = A[a1[t]];
            x2[t] = A[a2[t]];
        }
#pragma unroll
        for (int t = 0; t < TPP_B; ++t) {
            const int q = q0 + NT * t;
            if (q < c.m && q != i && q != j) { A[a1[t]] = x2[t]; A[a2[t]] = x1[t]; }
        }
    }
    if (c.lane == 0) {
        const double t = A[i + i * lda]; A[i + i * lda] = A[j + j * lda]; A[j + j * lda] = t;
        const int pt = perm[i]; perm[i] = perm[j]; perm[j] = pt;
    }
    __syncthreads();
}

template <int TJ, class P, int NT>
__device__ double tpp_apply_1x1(const TppCtx<P, NT> &c, int ne) {
    P A = c.A;
    const size_t lda = c.lda;
    const double d = A[ne + ne * lda];
    const double inv = 1.0 / d;
    double ml = 0.0;
    // TPP_B rows per lane at a time, every load of a batch before its stores
    for (int i0 = ne + 1 + c.lane; i0 < c.m; i0 += NT * TPP_B) {
        double v[TPP_B];
#pragma unroll
        for (int t = 0; t < TPP_B; ++t) v[t] = (i0 + NT * t < c.m) ? A[(i0 + NT * t) + ne * lda] : 0.0;
#pragma unroll
        for (int t = 0; t < TPP_B; ++t)
            if (i0 + NT * t < c.m) {
                const double l = v[t] * inv;
                A[(i0 + NT * t) + ne * lda] = l;
                ml = fmax(ml, fabs(l));
            }
    }
    ml = blk_fmax<NT>(ml);
    __syncthreads();
    for (int i0 = ne + 1 + c.lane; i0 < c.m; i0 += NT * TPP_B) {
        double li[TPP_B];
#pragma unroll
        for (int t = 0; t < TPP_B; ++t) li[t] = (i0 + NT * t < c.m) ? A[(i0 + NT * t) + ne * lda] : 0.0;
        // TJ columns per memory round trip: their loads (range-checked buffer loads, zero
        // outside the column's live rows: no branch, no wait per load) before any store
        for (int j0 = ne + 1; j0 < c.n; j0 += TJ) {
            const __amdgpu_buffer_rsrc_t rc = buf_rsrc((const double *)(A + (size_t)j0 * lda));
            double ldl[TJ], cv[TJ][TPP_B];
#pragma unroll
            for (int q = 0; q < TJ; ++q) {
                const int j = j0 + q;
                ldl[q] = (j < c.n) ? A[j + ne * lda] * d : 0.0;
#pragma unroll
                for (int t = 0; t < TPP_B; ++t) {
                    const int i = i0 + NT * t;
                    const bool ok = j < c.n && i >= j && i < c.m;
                    cv[q][t] = buf_load_f64(rc, ok ? (unsigned)(i + (size_t)q * lda) * 8u : BUF_OOB);
                }
            }
#pragma unroll
            for (int q = 0; q < TJ; ++q) {
                const int j = j0 + q;
#pragma unroll
                for (int t = 0; t < TPP_B; ++t) {
                    const int i = i0 + NT * t;
                    if (j < c.n && i >= j && i < c.m) A[i + j * lda] = cv[q][t] - li[t] * ldl[q];
                }
            }
        }
    }
    __syncthreads();
    return ml;
}

template <int TJ, class P, int NT>
__device__ double tpp_apply_2x2(const TppCtx<P, NT> &c, int ne) {
    P A = c.A;
    const size_t lda = c.lda;
    const double a11 = A[ne + ne * lda], a21 = A[(ne + 1) + ne * lda], a22 = A[(ne + 1) + (ne + 1) * lda];
    const double det = a11 * a22 - a21 * a21;
    const double inv_det = 1.0 / det;
    double ml = 0.0;
    for (int i0 = ne + 2 + c.lane; i0 < c.m; i0 += NT * TPP_B) {
        double v1[TPP_B], v2[TPP_B];
#pragma unroll
        for (int t = 0; t < TPP_B; ++t) {
            const int i = i0 + NT * t;
            v1[t] = (i < c.m) ? A[i + ne * lda] : 0.0;
            v2[t] = (i < c.m) ? A[i + (ne + 1) * lda] : 0.0;
        }
#pragma unroll
        for (int t = 0; t < TPP_B; ++t) {
            const int i = i0 + NT * t;
            if (i < c.m) {
                const double ai1 = v1[t], ai2 = v2[t];
                const double l1 = (ai1 * a22 - ai2 * a21) * inv_det;
                const double l2 = (ai2 * a11 - ai1 * a21) * inv_det;
                A[i + ne * lda] = l1;
                A[i + (ne + 1) * lda] = l2;
                ml = fmax(fmax(ml, fabs(l1)), fabs(l2));
            }
        }
    }
    ml = blk_fmax<NT>(ml);
    __syncthreads();
    for (int i0 = ne + 2 + c.lane; i0 < c.m; i0 += NT * TPP_B) {
        double l1[TPP_B], l2[TPP_B];
#pragma unroll
        for (int t = 0; t < TPP_B; ++t) {
            const int i = i0 + NT * t;
            l1[t] = (i < c.m) ? A[i + ne * lda] : 0.0;
            l2[t] = (i < c.m) ? A[i + (ne + 1) * lda] : 0.0;
        }
        for (int j0 = ne + 2; j0 < c.n; j0 += TJ) {
            const __amdgpu_buffer_rsrc_t rc = buf_rsrc((const double *)(A + (size_t)j0 * lda));
            double wj1[TJ], wj2[TJ], cv[TJ][TPP_B];
#pragma unroll
            for (int q = 0; q < TJ; ++q) {
                const int j = j0 + q;
                const double lj1 = (j < c.n) ? A[j + ne * lda] : 0.0;
                const double lj2 = (j < c.n) ? A[j + (ne + 1) * lda] : 0.0;
                wj1[q] = lj1 * a11 + lj2 * a21;
                wj2[q] = lj1 * a21 + lj2 * a22;
#pragma unroll
                for (int t = 0; t < TPP_B; ++t) {
                    const int i = i0 + NT * t;
                    const bool ok = j < c.n && i >= j && i < c.m;
                    cv[q][t] = buf_load_f64(rc, ok ? (unsigned)(i + (size_t)q * lda) * 8u : BUF_OOB);
                }
            }
#pragma unroll
            for (int q = 0; q < TJ; ++q) {
                const int j = j0 + q;
#pragma unroll
                for (int t = 0; t < TPP_B; ++t) {
                    const int i = i0 + NT * t;
                    if (j < c.n && i >= j && i < c.m) A[i + j * lda] = cv[q][t] - (l1[t] * wj1[q] + l2[t] * wj2[q]);
                }
            }
        }
    }
    __syncthreads();
    if (c.lane == 0) A[(ne + 1) + ne * lda] = 0.0;
    __syncthreads();
    return ml;
}

// Threshold partial pivoting of one front from column nelim (the wavefront of the calling
// workgroup).
template <int TJ, class P, int NT>
__device__ void tpp_run(const TppCtx<P, NT> &c, FrontDev &F, int nelim, double u, double small) {
    const int n = c.n;
    gint *perm = gbl(F.perm);
    int n1 = 0, n2 = 0;
    double maxl = 0.0;
    auto set1 = [&](int col, double v) {
        if (c.lane == 0) { gbl(F.dkind)[col] = 1; gbl(F.ddiag)[col] = v; gbl(F.doff)[col] = 0.0; }
    };
    while (nelim < n) {
        if (tpp_col_small(c, nelim, nelim, small)) {
            set1(nelim, 0.0);
            ++n1;
            for (int r = nelim + 1 + c.lane; r < c.m; r += NT) c.A[r + nelim * c.lda] = 0.0;
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
                for (int r = nelim + 1 + c.lane; r < c.m; r += NT) c.A[r + nelim * c.lda] = 0.0;
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
                const double ml = tpp_apply_2x2<TJ>(c, nelim);
                if (c.lane == 0) {
                    gbl(F.dkind)[nelim] = 2; gbl(F.dkind)[nelim + 1] = 3;
                    gbl(F.ddiag)[nelim] = d11; gbl(F.ddiag)[nelim + 1] = d22;
                    gbl(F.doff)[nelim] = d21; gbl(F.doff)[nelim + 1] = 0.0;
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
                const double ml = tpp_apply_1x1<TJ>(c, nelim);
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
                const double ml = tpp_apply_1x1<TJ>(c, nelim);
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

// One wavefront per front, factored in place.  (Staging a front's first k columns in LDS was
// measured slower on the leaf levels: the LDS per workgroup cut the wavefronts per CU more than the
// shorter latency gained.)  TJ = 2 keeps three wavefronts per SIMD (levels with more fronts than
// SIMDs); TJ = 8 takes a quarter of the update round trips (levels of few, larger fronts).
template <int TJ, int NT>
__global__ __launch_bounds__(NT) void k_tpp(FrontDev *__restrict__ fr, const int *__restrict__ flist,
                                            int mode, double u, double small) {
    FrontDev &F = fr[flist[blockIdx.x]];
    const int n = F.k, m = F.m;
    const int nelim = (mode == 0) ? 0 : F.ne;
    if (nelim >= n) return;
    tpp_run<TJ>(TppCtx<gdouble *, NT>{gbl(F.F), (size_t)m, m, n, (int)threadIdx.x}, F, nelim, u, small);
}

// ===========================================================================
// Schur updates on FP64 MFMA: C -= W L_c^T on the lower part, with W = L_r D formed once per
// update by k_form_ld.
//   mode 0 (contribution, compute_contribution_gemm factor.rs:1590-1652):
//           rows/cols [k, m), K = ne columns from 0.
//   mode 1 (outer-block update of the deferred trailing columns):
//           rows [obe, m), cols [obe, k), K = columns eliminated in the outer block.
// ===========================================================================
typedef double d4 __attribute__((ext_vector_type(4)));

struct UpdGeom {
    int r0, nr, nc, kc0, K;
};

__device__ __forceinline__ UpdGeom upd_geom(const FrontDev &F, const AptpState *st, int aux, int mode) {
    UpdGeom g;
    if (mode == 0) {
        g.r0 = F.k; g.nr = F.m - F.k; g.nc = g.nr; g.kc0 = 0; g.K = F.ne;
    } else {
        const AptpState &S = st[aux];
        g.r0 = S.ob_start + S.ob_cols; g.nr = F.m - g.r0; g.nc = F.k - g.r0; g.kc0 = S.ob_start;
        g.K = S.kpos - S.ob_start;
    }
    return g;
}

constexpr int FC = 32;  // columns per k_form_ld work item

// W[r, c] = (L D)[r0 + r, kc0 + c] (compute_ld_into, factor.rs:1536-1563), column-major with
// leading dimension nr.  Work item (front, aux, 256 rows, FC columns).  For an FP32 front
// (FRONT_W32) the workspace holds W and a copy of L[r0 + r, kc0 + c] as FP32, each nr x K with
// leading dimension nr (together the bytes of the FP64 W): the rounding k_schur<1> applies to its
// operands, done once per front instead of once per tile, and half the operand bytes per tile.
// snap (mode 1, may be null): the first item of each front copies the front's APTP state there,
// for an outer update that runs after the state has moved on to the next outer block.
__global__ __launch_bounds__(256) void k_form_ld(const FrontDev *__restrict__ fr,
                                                 const AptpState *__restrict__ st,
                                                 const int4 *__restrict__ items, int mode,
                                                 double *const *__restrict__ wp,
                                                 AptpState *__restrict__ snap) {
    const int4 it = items[blockIdx.x];
    const FrontDev &F = fr[it.x];
    if (snap && it.z == 0 && it.w == 0) {
        const int *src = (const int *)&st[it.y];
        int *dst = (int *)&snap[it.y];
        for (int i = threadIdx.x; i < (int)(sizeof(AptpState) / sizeof(int)); i += blockDim.x) dst[i] = src[i];
    }
    const UpdGeom g = upd_geom(F, st, it.y, mode);
    if (g.K <= 0 || g.nr <= 0) return;
    const int c0 = it.w * FC;
    if (c0 >= g.K) return;
    const int cn = min(FC, g.K - c0);
    __shared__ int sk[FC + 2];
    __shared__ double sd[FC + 2], so[FC + 2];
    const int tid = threadIdx.x;
    if (tid < FC + 2) {
        const int c = c0 - 1 + tid;
        const bool in = c >= 0 && c < g.K;
        sk[tid] = in ? gbl(F.dkind)[g.kc0 + c] : 0;
        sd[tid] = in ? gbl(F.ddiag)[g.kc0 + c] : 0.0;
        so[tid] = in ? gbl(F.doff)[g.kc0 + c] : 0.0;
    }
    __syncthreads();
    const int r = it.z * 256 + tid;
    if (r >= g.nr) return;
    const size_t lda = F.m;
    const gdouble *row = gbl(F.F) + (g.r0 + r);
    double lv[FC + 2];
#pragma unroll
    for (int u = 0; u < FC + 2; ++u) {
        const int c = c0 - 1 + u;
        lv[u] = (c >= 0 && c < g.K) ? row[(size_t)(g.kc0 + c) * lda] : 0.0;
    }
    const bool w32 = (F.flags & FRONT_W32) != 0;
    gdouble *W = gbl(wp[it.y]) + r;
    __attribute__((address_space(1))) float *W32 = (__attribute__((address_space(1))) float *)gbl(wp[it.y]) + r;
    __attribute__((address_space(1))) float *L32 = W32 + (size_t)g.nr * g.K;
#pragma unroll
    for (int u = 1; u <= FC; ++u) {
        if (u - 1 < cn) {
            const int kind = sk[u];
            double v = 0.0;
            if (kind == 1) v = lv[u] * sd[u];
            else if (kind == 2) v = lv[u] * sd[u] + lv[u + 1] * so[u];
            else if (kind == 3) v = lv[u - 1] * so[u - 1] + lv[u] * sd[u];
            const size_t o = (size_t)(c0 + u - 1) * g.nr;
            if (w32) {
                W32[o] = (float)v;
                L32[o] = (float)lv[u];
            } else {
                W[o] = v;
            }
        }
    }
}

typedef float f4 __attribute__((ext_vector_type(4)));
typedef short s4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

// Workgroup -> tile-list index.  Consecutive workgroups are dispatched round-robin over the 8
// XCDs (workgroup bid runs on XCD bid & 7).
//  * contiguous (contribution updates, every listed tile active): each XCD takes one contiguous
//    run of the list, so tiles of a row block share W rows in its L2;
//  * spread (outer-block updates): the list is planned for the largest trailing block, and only a
//    prefix of each front's tiles is active once earlier outer blocks have been eliminated; runs
//    of TILE_RUN consecutive tiles go to successive XCDs, so every prefix is spread over all of
//    them.  The grid is rounded up to a multiple of 8 * TILE_RUN; indices >= ntiles are idle.
constexpr unsigned TILE_RUN = 2;
__host__ __device__ constexpr unsigned spread_grid(size_t ntiles) {
    return (unsigned)((ntiles + 8 * TILE_RUN - 1) / (8 * TILE_RUN) * (8 * TILE_RUN));
}
__device__ __forceinline__ unsigned tile_index(unsigned bid, unsigned nbk, bool spread) {
    const unsigned xcd = bid & 7, s = bid >> 3;
    if (spread) return ((s / TILE_RUN) * 8 + xcd) * TILE_RUN + s % TILE_RUN;
    const unsigned qn = nbk >> 3, rem = nbk & 7;
    return xcd * qn + min(xcd, rem) + s;
}

// FP64 -> BF16 bits, round to nearest even (through FP32, itself rounded to nearest)
__device__ __forceinline__ unsigned short to_bf16(double x) {
    unsigned u = __float_as_uint((float)x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

// C -= W L_c^T: 128 x 128 output tile per workgroup (4 wavefronts x 64 x 64 = 4 x 4 MFMA
// 16x16 blocks), K in chunks of SKC through LDS with the next chunk prefetched into registers
// (a chunk's MFMA work outlasts the global-load latency).  The operands enter the MFMA as
// (L_c, W), so a lane's four results lie in one row of C.
//   PREC 0: FP64 operands, v_mfma_f64_16x16x4f64;
//   PREC 1: operands rounded to FP32, v_mfma_f32_16x16x4f32, FP32 accumulation;
//   PREC 2: operands rounded to BF16, v_mfma_f32_16x16x32_bf16, FP32 accumulation.
// The product is subtracted from the FP64 C in every case.
constexpr int SKC = 32;

// Two workgroups per CU for FP64 and BF16, three for FP32 (registers fit 256 resp. 168 per lane);
// FP64 and FP32 take K in chunks of 16 so that their double-buffered operand tiles fit that many
// times in the LDS.  TILE = 64 (FP64 / FP32 only): 64 x 64 output tiles, each wavefront a 32 x 32
// quarter, for the thin updates on the critical path (four times the workgroups of 128 x 128
// tiles over the same columns); every entry is summed in the same order either way.
template <int PREC, int TILE = 128>
__global__ __launch_bounds__(256, PREC == 1 ? 3 : 2) void k_schur(const FrontDev *__restrict__ fr,
                                               const AptpState *__restrict__ st,
                                               const TileItem *__restrict__ tiles, int mode,
                                               double *const *__restrict__ wp, unsigned ntiles) {
    static_assert(TILE == 128 || (TILE == 64 && PREC != 2), "tile: 128, or 64 for FP64 / FP32");
    constexpr int KCH = (PREC == 2) ? SKC : 16;  // K per chunk (the LDS of two FP64 / three FP32 workgroups per CU)
    constexpr int NB = TILE / 32;                 // 16 x 16 MFMA blocks per wavefront and dimension
    constexpr int WS = TILE / 2;                  // a wavefront's sub-tile
    constexpr int KS = 256 / TILE;                // k slices of a chunk's operand loads
    constexpr int P64 = TILE + 8, P32 = TILE + 16;  // LDS row pitches (FP64, FP32)
    const unsigned ti_ = tile_index(blockIdx.x, gridDim.x, mode == 1);
    if (ti_ >= ntiles) return;
    const TileItem t = tiles[ti_];
    const FrontDev &F = fr[t.front];
    const UpdGeom g = upd_geom(F, st, t.aux, mode);
    const int K = g.K, nr = g.nr, nc = g.nc, r0 = g.r0;
    if (K <= 0 || nr <= 0 || nc <= 0) return;
    const int bi = t.ti * TILE, bj = t.tj * TILE;
    if (bi >= nr || bj >= nc || bi + TILE - 1 < bj) return;
    const size_t lda = F.m;
    gdouble *A = gbl(F.F);
    // LDS operand tiles: [k][row] for FP64/FP32, [row][k] for BF16 (four consecutive k per lane)
    // two chunk buffers: a chunk is staged while the previous one feeds the MFMAs (one barrier
    // per chunk)
    constexpr int LDS_BYTES = (PREC == 0) ? 2 * KCH * P64 * 8 : (PREC == 1) ? 2 * KCH * P32 * 4
                                                                             : 2 * 128 * (KCH + 8) * 2;
    __shared__ __align__(16) char lds2[2 * LDS_BYTES];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wr = wv >> 1, wc = wv & 1;
    // a wavefront whose sub-tile lies strictly above the diagonal only helps staging
    const bool busy = bi + wr * WS + WS - 1 >= bj + wc * WS;
    const int lr = tid % TILE, ko = tid / TILE;
    const bool arow = bi + lr < nr, brow = bj + lr < nc;
    // operand loads through buffer resources based at the chunk's first column: an entry outside
    // the tile or beyond K gets an out-of-range offset and loads as zero, with no branch and no
    // select, so the waits count the loads of the chunk in flight exactly
    // FP32: W and L come as FP32 pairs from the workspace (k_form_ld, FRONT_W32); otherwise W
    // (FP64) from the workspace and L from the front
    typedef typename std::conditional<PREC == 1, float, double>::type op_t;
    const double *Wb = wp[t.aux] + bi;
    const double *Lb = F.F + (r0 + bj) + (size_t)g.kc0 * lda;
    const float *W32 = (const float *)wp[t.aux] + bi;
    const float *L32 = (const float *)wp[t.aux] + (size_t)nr * K + bj;
    auto load_chunk = [&](op_t *ra, op_t *rb, int kc) {
        if constexpr (PREC == 1) {
            const __amdgpu_buffer_rsrc_t rw = buf_rsrc(W32 + (size_t)kc * nr);
            const __amdgpu_buffer_rsrc_t rl = buf_rsrc(L32 + (size_t)kc * nr);
#pragma unroll
            for (int u = 0; u < KCH / KS; ++u) {
                const int cc = ko + KS * u, c = kc + cc;
                ra[u] = buf_load_f32(rw, (arow && c < K) ? (unsigned)(lr + cc * nr) * 4u : BUF_OOB);
                rb[u] = buf_load_f32(rl, (brow && c < K) ? (unsigned)(lr + cc * nr) * 4u : BUF_OOB);
            }
        } else {
            const __amdgpu_buffer_rsrc_t rw = buf_rsrc(Wb + (size_t)kc * nr);
            const __amdgpu_buffer_rsrc_t rl = buf_rsrc(Lb + (size_t)kc * lda);
#pragma unroll
            for (int u = 0; u < KCH / KS; ++u) {
                const int cc = ko + KS * u, c = kc + cc;
                ra[u] = buf_load_f64(rw, (arow && c < K) ? (unsigned)(lr + cc * nr) * 8u : BUF_OOB);
                rb[u] = buf_load_f64(rl, (brow && c < K) ? (unsigned)(lr + cc * (int)lda) * 8u : BUF_OOB);
            }
        }
    };
    constexpr int NU = KCH / KS;  // k values per thread and operand in one chunk
    typedef typename std::conditional<PREC == 0, d4, f4>::type acc_t;
    acc_t acc[NB][NB];
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) acc[i][j] = acc_t{0, 0, 0, 0};
    // a chunk's operands in registers into its LDS buffer
    auto stage = [&](const op_t *ra, const op_t *rb, char *lds) {
        if constexpr (PREC == 0) {
            double (*As)[P64] = (double (*)[P64])lds;
            double (*Bs)[P64] = (double (*)[P64])(lds + KCH * P64 * 8);
#pragma unroll
            for (int u = 0; u < NU; ++u) {
                As[ko + KS * u][lr] = ra[u];
                Bs[ko + KS * u][lr] = rb[u];
            }
        } else if constexpr (PREC == 1) {
            float (*As)[P32] = (float (*)[P32])lds;
            float (*Bs)[P32] = (float (*)[P32])(lds + KCH * P32 * 4);
#pragma unroll
            for (int u = 0; u < NU; ++u) {
                As[ko + KS * u][lr] = (float)ra[u];
                Bs[ko + KS * u][lr] = (float)rb[u];
            }
        } else {
            unsigned short (*As)[KCH + 8] = (unsigned short (*)[KCH + 8])lds;
            unsigned short (*Bs)[KCH + 8] = (unsigned short (*)[KCH + 8])(lds + 128 * (KCH + 8) * 2);
#pragma unroll
            for (int u = 0; u < NU; ++u) {
                As[lr][ko + 2 * u] = to_bf16(ra[u]);
                Bs[lr][ko + 2 * u] = to_bf16(rb[u]);
            }
        }
    };
    // the MFMAs of one staged chunk
    auto compute = [&](char *lds) {
        if (busy) {
            if constexpr (PREC == 0) {
                double (*As)[P64] = (double (*)[P64])lds;
                double (*Bs)[P64] = (double (*)[P64])(lds + KCH * P64 * 8);
#pragma unroll
                for (int kk = 0; kk < KCH; kk += 4) {
                    const int kr = kk + (lane >> 4);
                    double a[NB], b[NB];
#pragma unroll
                    for (int i = 0; i < NB; ++i) a[i] = As[kr][wr * WS + i * 16 + (lane & 15)];
#pragma unroll
                    for (int j = 0; j < NB; ++j) b[j] = Bs[kr][wc * WS + j * 16 + (lane & 15)];
#pragma unroll
                    for (int i = 0; i < NB; ++i)
#pragma unroll
                        for (int j = 0; j < NB; ++j)
                            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[j], a[i], acc[i][j], 0, 0, 0);
                }
            } else if constexpr (PREC == 1) {
                float (*As)[P32] = (float (*)[P32])lds;
                float (*Bs)[P32] = (float (*)[P32])(lds + KCH * P32 * 4);
#pragma unroll
                for (int kk = 0; kk < KCH; kk += 4) {
                    const int kr = kk + (lane >> 4);
                    float a[NB], b[NB];
#pragma unroll
                    for (int i = 0; i < NB; ++i) a[i] = As[kr][wr * WS + i * 16 + (lane & 15)];
#pragma unroll
                    for (int j = 0; j < NB; ++j) b[j] = Bs[kr][wc * WS + j * 16 + (lane & 15)];
#pragma unroll
                    for (int i = 0; i < NB; ++i)
#pragma unroll
                        for (int j = 0; j < NB; ++j)
                            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(b[j], a[i], acc[i][j], 0, 0, 0);
                }
            } else {
                unsigned short (*As)[KCH + 8] = (unsigned short (*)[KCH + 8])lds;
                unsigned short (*Bs)[KCH + 8] = (unsigned short (*)[KCH + 8])(lds + 128 * (KCH + 8) * 2);
                // v_mfma_f32_16x16x32_bf16: lane l holds eight consecutive k, 8 (l / 16) .. + 7,
                // of row / column l % 16
#pragma unroll
                for (int kk = 0; kk < KCH; kk += 32) {
                    const int k8 = kk + 8 * (lane >> 4);
                    bf8 a[4], b[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) a[i] = *(const bf8 *)&As[wr * 64 + i * 16 + (lane & 15)][k8];
#pragma unroll
                    for (int j = 0; j < 4; ++j) b[j] = *(const bf8 *)&Bs[wc * 64 + j * 16 + (lane & 15)][k8];
#pragma unroll
                    for (int i = 0; i < 4; ++i)
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[j], a[i], acc[i][j], 0, 0, 0);
                }
            }
        }
    };
    // the next chunk's operands are loaded while the current one feeds the MFMAs
    op_t ra[NU], rb[NU];
    load_chunk(ra, rb, 0);
    for (int kc = 0, buf = 0; kc < K; kc += KCH, buf ^= 1) {
        char *lds = lds2 + buf * LDS_BYTES;
        stage(ra, rb, lds);
        __syncthreads();
        if (kc + KCH < K) load_chunk(ra, rb, kc + KCH);
        compute(lds);
    }
    if (!busy) return;
    // result layouts, (L_c, W) product: v_mfma_f64_16x16x4f64 lane l element q holds D row
    // (l/16) + 4q, column l%16; the FP32-accumulating 16x16 MFMAs hold D row 4(l/16) + q.
    // D row = C column, D column = C row.  C's entries are loaded before any is stored, half of
    // the tile at a time (a store may alias a later load, so interleaved read-modify-writes would
    // take one memory round trip each); entries outside the lower tile load as zero and are not
    // stored.
    const int rl0 = bi + wr * WS + (lane & 15);
    const __amdgpu_buffer_rsrc_t rc = buf_rsrc(F.F + (r0 + bi) + (size_t)(r0 + bj) * lda);  // tile origin
    constexpr int NH = (NB + 1) / 2;  // row blocks per half
#pragma unroll
    for (int h = 0; h < NB / NH; ++h) {
        double cv[NH][NB][4];
#pragma unroll
        for (int i = 0; i < NH; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int rl = rl0 + (NH * h + i) * 16;
                    const int dr = (PREC == 0) ? (lane >> 4) + 4 * q : 4 * (lane >> 4) + q;
                    const int cl = bj + wc * WS + j * 16 + dr;
                    const bool ok = rl < nr && cl < nc && rl >= cl;
                    const unsigned off = (unsigned)((rl - bi) + (size_t)(cl - bj) * lda) * 8u;
                    cv[i][j][q] = buf_load_f64(rc, ok ? off : BUF_OOB);
                }
#pragma unroll
        for (int i = 0; i < NH; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int rl = rl0 + (NH * h + i) * 16;
                    const int dr = (PREC == 0) ? (lane >> 4) + 4 * q : 4 * (lane >> 4) + q;
                    const int cl = bj + wc * WS + j * 16 + dr;
                    if (rl < nr && cl < nc && rl >= cl)
                        A[(r0 + rl) + (size_t)(r0 + cl) * lda] = cv[i][j][q] - (double)acc[NH * h + i][j][q];
                }
    }
}

// C -= W L_c^T in the oracle's summation order (gemm_sub_wlt): every entry accumulates
// W[i,c] * L[j,c] over c ascending from zero, multiply and add rounded separately (this file
// is compiled without FMA contraction), and is subtracted once.  128 x 128 tile per workgroup,
// each thread 8 rows x 8 columns; K staged through LDS in chunks of SKC.
__global__ __launch_bounds__(256) void k_schur_exact(const FrontDev *__restrict__ fr,
                                                     const AptpState *__restrict__ st,
                                                     const TileItem *__restrict__ tiles, int mode,
                                                     double *const *__restrict__ wp, unsigned ntiles) {
    const unsigned ti_ = tile_index(blockIdx.x, gridDim.x, mode == 1);
    if (ti_ >= ntiles) return;
    const TileItem t = tiles[ti_];
    const FrontDev &F = fr[t.front];
    const UpdGeom g = upd_geom(F, st, t.aux, mode);
    const int K = g.K, nr = g.nr, nc = g.nc, r0 = g.r0;
    if (K <= 0 || nr <= 0 || nc <= 0) return;
    const int bi = t.ti * 128, bj = t.tj * 128;
    if (bi >= nr || bj >= nc || bi + 127 < bj) return;
    const size_t lda = F.m;
    gdouble *A = gbl(F.F);
    __shared__ double As[SKC][132];
    __shared__ double Bs[SKC][132];
    const int tid = threadIdx.x;
    const int tr = tid & 15, tc = tid >> 4;  // rows tr + 16a, columns tc + 16b
    const int lr = tid & 127, ko = tid >> 7;
    const bool arow = bi + lr < nr, brow = bj + lr < nc;
    const gdouble *Wp = gbl(wp[t.aux]) + bi + lr;
    const gdouble *Lp = A + (r0 + bj + lr) + (size_t)g.kc0 * lda;
    auto load_chunk = [&](double *ra, double *rb, int kc) {
#pragma unroll
        for (int u = 0; u < SKC / 2; ++u) {
            const int c = kc + ko + 2 * u;
            ra[u] = (arow && c < K) ? Wp[(size_t)c * nr] : 0.0;
            rb[u] = (brow && c < K) ? Lp[(size_t)c * lda] : 0.0;
        }
    };
    constexpr int NU = SKC / 2;
    double ra[NU], rb[NU];
    load_chunk(ra, rb, 0);
    double acc[8][8];
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 8; ++b) acc[a][b] = 0.0;
    for (int kc = 0; kc < K; kc += SKC) {
#pragma unroll
        for (int u = 0; u < NU; ++u) {
            As[ko + 2 * u][lr] = ra[u];
            Bs[ko + 2 * u][lr] = rb[u];
        }
        __syncthreads();
        if (kc + SKC < K) load_chunk(ra, rb, kc + SKC);
        const int kn = min(SKC, K - kc);  // products beyond K are not added
        for (int kk = 0; kk < kn; ++kk) {
            double av[8], bv[8];
#pragma unroll
            for (int a = 0; a < 8; ++a) av[a] = As[kk][tr + 16 * a];
#pragma unroll
            for (int b = 0; b < 8; ++b) bv[b] = Bs[kk][tc + 16 * b];
#pragma unroll
            for (int a = 0; a < 8; ++a)
#pragma unroll
                for (int b = 0; b < 8; ++b) acc[a][b] += av[a] * bv[b];
        }
        __syncthreads();
    }
    // C's entries are loaded before any is stored, half of the tile at a time (a store may alias
    // a later load, so interleaved read-modify-writes would take one memory round trip each)
    const __amdgpu_buffer_rsrc_t rc = buf_rsrc(F.F + (r0 + bi) + (size_t)(r0 + bj) * lda);  // tile origin
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        double cv[4][8];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const int rl = bi + tr + 16 * (4 * h + a), cl = bj + tc + 16 * b;
                const bool ok = rl < nr && cl < nc && rl >= cl;
                const unsigned off = (unsigned)((rl - bi) + (size_t)(cl - bj) * lda) * 8u;
                cv[a][b] = buf_load_f64(rc, ok ? off : BUF_OOB);
            }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const int rl = bi + tr + 16 * (4 * h + a), cl = bj + tc + 16 * b;
                if (rl < nr && cl < nc && rl >= cl)
                    A[(r0 + rl) + (size_t)(r0 + cl) * lda] = cv[a][b] - acc[4 * h + a][b];
            }
    }
}

// Deferred outer-block update in the reference's order.  update_trailing (factor.rs:1418-1526)
// subtracts each inner block's rank-nelim product from every fully-summed column beyond it; the
// columns beyond the current outer block are not read until their own outer block, so their
// updates are applied here, once per outer block, and still one inner block at a time: each
// entry of rows [obe, m) x columns [obe, p) becomes C - s_0 - s_1 - ..., s_b summed over block
// b's columns in ascending order from zero (gemm_sub_wlt), each subtraction rounded.  64 x 64
// tile per workgroup, each thread 4 x 4 entries kept in registers across the blocks.
__global__ __launch_bounds__(256, 3) void k_outer_exact(const FrontDev *__restrict__ fr,
                                                     const AptpState *__restrict__ st,
                                                     const TileItem *__restrict__ tiles,
                                                     double *const *__restrict__ wpp, unsigned ntiles) {
    const unsigned ti_ = tile_index(blockIdx.x, gridDim.x, true);
    if (ti_ >= ntiles) return;
    const TileItem t = tiles[ti_];
    const FrontDev &F = fr[t.front];
    const AptpState &S = st[t.aux];
    const int r0 = S.ob_start + S.ob_cols, nr = F.m - r0, nc = F.k - r0;
    const int nib = min(S.nib, MAX_OB_IB);
    if (nr <= 0 || nc <= 0 || S.kpos == S.ob_start) return;
    const int bi = t.ti * 64, bj = t.tj * 64;
    if (bi >= nr || bj >= nc || bi + 63 < bj) return;
    const size_t lda = F.m;
    const gdouble *A = gbl(F.F);
    // W = L D of the outer block's eliminated columns, rows [r0, m), leading dimension nr
    const gdouble *W = gbl(wpp[t.aux]);
    __shared__ double Ws[MAX_IB][65];
    __shared__ double Ls[MAX_IB][65];
    const int tid = threadIdx.x;
    const int tr = tid & 15, tc = tid >> 4;  // rows tr + 16a, columns tc + 16b
    const int lr = tid & 63, lk = tid >> 6;  // staging: row lr, columns lk + 4u
    // operands of one inner block prefetched into registers while the previous one computes
    double pw[MAX_IB / 4], pl[MAX_IB / 4];
    // range-checked buffer loads based at the block's first column (zero outside the tile): no
    // branch or select per load, so the prefetch stays in flight while the current block computes
    const int obs = S.ob_start;
    const bool wrow = bi + lr < nr, lrow = bj + lr < nc;
    auto prefetch = [&](int blk, int col0) {
        const int ne = S.ib_ne[blk];
        const __amdgpu_buffer_rsrc_t rw = buf_rsrc(wpp[t.aux] + bi + (size_t)(col0 - obs) * nr);
        const __amdgpu_buffer_rsrc_t rl = buf_rsrc(F.F + (r0 + bj) + (size_t)col0 * lda);
#pragma unroll
        for (int u = 0; u < MAX_IB / 4; ++u) {
            const int cc = lk + 4 * u;
            pw[u] = buf_load_f64(rw, (wrow && cc < ne) ? (unsigned)(lr + cc * nr) * 8u : BUF_OOB);
            pl[u] = buf_load_f64(rl, (lrow && cc < ne) ? (unsigned)(lr + cc * (int)lda) * 8u : BUF_OOB);
        }
    };
    int blk = 0, col0 = S.ob_start;
    while (blk < nib && S.ib_ne[blk] == 0) ++blk;
    if (blk < nib) prefetch(blk, col0);
    double c[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int rl = bi + tr + 16 * a, cl = bj + tc + 16 * b;
            c[a][b] = (rl < nr && cl < nc && rl >= cl) ? A[(r0 + rl) + (size_t)(r0 + cl) * lda] : 0.0;
        }
    while (blk < nib) {
        const int ne = S.ib_ne[blk];
        lds_barrier();  // the previous block's operands are consumed
#pragma unroll
        for (int u = 0; u < MAX_IB / 4; ++u) {
            const int cc = lk + 4 * u;
            if (cc < ne) {
                Ws[cc][lr] = pw[u];
                Ls[cc][lr] = pl[u];
            }
        }
        lds_barrier();
        int nblk = blk + 1;
        const int ncol0 = col0 + ne;
        while (nblk < nib && S.ib_ne[nblk] == 0) ++nblk;
        if (nblk < nib) prefetch(nblk, ncol0);
        double s[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) s[a][b] = 0.0;
        // two operand sets in turn: the next step's operands are read from LDS before the
        // current step's products, so their latency is covered (a row past ne, or row 0 after
        // the last, is read and unused)
        double w0[4], l0[4], w1[4], l1[4];
        auto rd = [&](double *w, double *l, int row) {
#pragma unroll
            for (int a = 0; a < 4; ++a) w[a] = Ws[row][tr + 16 * a];
#pragma unroll
            for (int b = 0; b < 4; ++b) l[b] = Ls[row][tc + 16 * b];
        };
        auto acc = [&](const double *w, const double *l) {
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) s[a][b] += w[a] * l[b];
        };
        rd(w0, l0, 0);
        int cc = 0;
#pragma unroll 1
        for (; cc + 1 < ne; cc += 2) {
            // scheduling barriers keep each read ahead of the products it overlaps
            rd(w1, l1, cc + 1);
            __builtin_amdgcn_sched_barrier(0);
            acc(w0, l0);
            __builtin_amdgcn_sched_barrier(0);
            rd(w0, l0, (cc + 2) & (MAX_IB - 1));
            __builtin_amdgcn_sched_barrier(0);
            acc(w1, l1);
            __builtin_amdgcn_sched_barrier(0);
        }
        if (cc < ne) acc(w0, l0);
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) c[a][b] -= s[a][b];
        blk = nblk;
        col0 = ncol0;
    }
    gdouble *Aw = gbl(F.F);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int rl = bi + tr + 16 * a, cl = bj + tc + 16 * b;
            if (rl < nr && cl < nc && rl >= cl) Aw[(r0 + rl) + (size_t)(r0 + cl) * lda] = c[a][b];
        }
}

constexpr int KF_CHUNK = 8192;  // doubles per k_keep_factor item

// Eliminated columns of a front held in the pool into its kept storage (the same column-major
// layout, lda = m): the first ne * m entries of the front.  One item = KF_CHUNK entries, 32 per
// thread, every load issued before the stores.
__global__ __launch_bounds__(256) void k_keep_factor(const FrontDev *__restrict__ fr,
                                                     const int2 *__restrict__ items) {
    const int2 it = items[blockIdx.x];
    const FrontDev &F = fr[it.x];
    const size_t n = (size_t)F.ne * F.m;
    const size_t base = (size_t)it.y * KF_CHUNK + threadIdx.x;
    if ((size_t)it.y * KF_CHUNK >= n) return;
    const gdouble *src = gbl((const double *)F.F);
    gdouble *dst = gbl(F.Lf);
    constexpr int U = KF_CHUNK / 256;
    double v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const size_t i = base + (size_t)u * 256;
        v[u] = i < n ? src[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const size_t i = base + (size_t)u * 256;
        if (i < n) dst[i] = v[u];
    }
}

constexpr int RES_INTS = 10;  // per-front results of k_collect

// Per-front results (one workgroup of 256 threads per front): counts, inertia, max|L|, update
// flops, the solve's index map and the delayed columns.  Each loop issues the loads of COLB
// entries per thread before using any of them.
constexpr int COLB = 4;

__global__ __launch_bounds__(256) void k_collect(FrontDev *__restrict__ fr,
                                                 const int *__restrict__ flist,
                                                 const long long *__restrict__ doffs,
                                                 int *__restrict__ res, int *__restrict__ delayed) {
    const int li = blockIdx.x;
    const int f = flist[li];
    FrontDev &F = fr[f];
    const int ne = F.ne, k = F.k, m = F.m;
    const int tid = threadIdx.x;
    const gi8 *dkind = gbl(F.dkind);
    const gdouble *ddiag = gbl(F.ddiag), *doff = gbl(F.doff);
    // zero 1x1 pivots and the inertia of D (MixedDiagonal::compute_inertia, diagonal.rs:379-416)
    int z = 0, pos = 0, neg = 0, zer = 0;
    for (int c0 = tid; c0 < ne; c0 += 256 * COLB) {
        int kd[COLB];
        double a[COLB], b[COLB], cn[COLB];
#pragma unroll
        for (int u = 0; u < COLB; ++u) {
            const int c = c0 + 256 * u;
            kd[u] = (c < ne) ? dkind[c] : 0;
            a[u] = (c < ne) ? ddiag[c] : 0.0;
            b[u] = (c < ne) ? doff[c] : 0.0;
            cn[u] = (c + 1 < ne) ? ddiag[c + 1] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < COLB; ++u) {
            if (kd[u] == 1) {
                const double d = a[u];
                if (d > 0.0) ++pos;
                else if (d < 0.0) ++neg;
                else { ++zer; ++z; }
            } else if (kd[u] == 2) {
                const double det = a[u] * cn[u] - b[u] * b[u], tr = a[u] + cn[u];
                if (det > 0.0) {
                    if (tr > 0.0) pos += 2; else neg += 2;
                } else if (det < 0.0) {
                    ++pos; ++neg;
                } else if (tr > 0.0) {
                    ++pos; ++zer;
                } else if (tr < 0.0) {
                    ++neg; ++zer;
                } else {
                    zer += 2;
                }
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        z += __shfl_xor(z, o);
        pos += __shfl_xor(pos, o);
        neg += __shfl_xor(neg, o);
        zer += __shfl_xor(zer, o);
    }
    __shared__ int part[4][4];
    if ((tid & 63) == 0) {
        part[tid >> 6][0] = z;
        part[tid >> 6][1] = pos;
        part[tid >> 6][2] = neg;
        part[tid >> 6][3] = zer;
    }
    __syncthreads();
    if (tid == 0) {
        int *r = res + li * RES_INTS;
        r[0] = ne;
        r[1] = F.n1;
        r[2] = F.n2;
        r[7] = ((F.flags & FRONT_CHAIN_ERR) ? 1 : 0) | ((F.flags & FRONT_LA_ERR) ? 2 : 0);
        r[8] = (int)(unsigned)(F.obfail & 0xffffffffu);
        r[9] = (int)(unsigned)(F.obfail >> 32);
#pragma unroll
        for (int q = 0; q < 4; ++q) r[3 + q] = part[0][q] + part[1][q] + part[2][q] + part[3][q];
        unsigned long long *ml = (unsigned long long *)(res + RES_INTS * gridDim.x);
        ml[li] = F.max_l_bits;
        ((double *)(ml + gridDim.x))[li] = F.uflops;
    }
    int *dl = delayed + doffs[li];
    const gint *rows = gbl(F.rows), *perm = gbl(F.perm);
    gint *sidx = gbl(F.sidx);
    for (int i0 = tid; i0 < m; i0 += 256 * COLB) {
        int src[COLB], g[COLB];
#pragma unroll
        for (int u = 0; u < COLB; ++u) {
            const int i = i0 + 256 * u;
            src[u] = (i < k) ? perm[i] : i;
        }
#pragma unroll
        for (int u = 0; u < COLB; ++u) g[u] = (i0 + 256 * u < m) ? rows[src[u]] : 0;
#pragma unroll
        for (int u = 0; u < COLB; ++u) {
            const int i = i0 + 256 * u;
            if (i < m) {
                sidx[i] = g[u];
                if (i >= ne && i < k) dl[i - ne] = g[u];
            }
        }
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
    fprec_.assign(ns, (signed char)o_.update_precision);
    const char *ch = std::getenv("RV_APTP_CHAIN");
    chain_on_ = !(ch && std::strcmp(ch, "0") == 0);
    const char *la = std::getenv("RV_OUTER_LOOKAHEAD");
    la_on_ = !(la && std::strcmp(la, "0") == 0);
    const char *t64 = std::getenv("RV_OUTER_TILE64");
    outer_t64_ = !(t64 && std::strcmp(t64, "0") == 0);
    const char *kf = std::getenv("RV_KEEP_FRONTS");
    keep_fronts_ = kf && std::strcmp(kf, "1") == 0;
    obfail_.assign(ns, 0);
    pool_block_.assign(ns, -1);
}

size_t NumericGPU::chain_capacity(int ch) {
    int &cap = chain_cap_[ch == 64 ? 0 : 1];
    if (cap < 0) {
        int dev = 0, cus = 0, per = 0;
        RV_HIP(hipGetDevice(&dev));
        RV_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        if (ch == 64)
            RV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_aptp_outer<64>, 256, 0));
        else
            RV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_aptp_outer<128>, 256, 0));
        cap = per * cus;
    }
    return (size_t)cap;
}

void NumericGPU::set_front_precision(const signed char *prec) {
    std::vector<signed char> next(fprec_.size(), (signed char)o_.update_precision);
    if (prec) next.assign(prec, prec + fprec_.size());
    if (next != fprec_) {
        fprec_ = std::move(next);
        ++prec_gen_;  // every level plan is rebuilt at the next factorization
    }
}

void NumericGPU::drop_graph() {
    if (gexec_) (void)hipGraphExecDestroy(gexec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    gexec_ = nullptr;
    graph_ = nullptr;
}

NumericGPU::~NumericGPU() {
    drop_graph();
    if (ev_fork_) (void)hipEventDestroy(ev_fork_);
    if (ev_join_) (void)hipEventDestroy(ev_join_);
    if (stream2_) (void)hipStreamDestroy(stream2_);
    if (d_fr_) (void)hipFree(d_fr_);
    for (auto &b : pin_blocks_) (void)hipHostFree(b.first);
    for (LevelPlan &L : plans_)
        for (auto &e : L.sev) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
    if (d_sol_meta_) (void)hipFree(d_sol_meta_);
}

char *NumericGPU::pin_take(size_t bytes) {
    bytes = (bytes + 255) / 256 * 256;
    if (bytes == 0) bytes = 256;
    if (pin_cur_ < pin_blocks_.size() && pin_used_ + bytes <= pin_blocks_[pin_cur_].second) {
        char *p = pin_blocks_[pin_cur_].first + pin_used_;
        pin_used_ += bytes;
        return p;
    }
    if (pin_cur_ + 1 < pin_blocks_.size() && bytes <= pin_blocks_[pin_cur_ + 1].second) {
        ++pin_cur_;
        pin_used_ = bytes;
        return pin_blocks_[pin_cur_].first;
    }
    // a new block; blocks in use by transfers in flight are never freed here
    size_t total = 0;
    for (auto &b : pin_blocks_) total += b.second;
    const size_t sz = std::max({bytes, (size_t)4 << 20, total});
    char *p = nullptr;
    RV_HIP(hipHostMalloc((void **)&p, sz, hipHostMallocDefault));
    pin_blocks_.emplace_back(p, sz);
    pin_cur_ = pin_blocks_.size() - 1;
    pin_used_ = bytes;
    return p;
}

// Start of a factorization (no transfers in flight): merge the blocks into one.
void NumericGPU::pin_reset() {
    if (pin_blocks_.size() > 1) {
        size_t total = 0;
        for (auto &b : pin_blocks_) {
            total += b.second;
            (void)hipHostFree(b.first);
        }
        pin_blocks_.clear();
        char *p = nullptr;
        RV_HIP(hipHostMalloc((void **)&p, total, hipHostMallocDefault));
        pin_blocks_.emplace_back(p, total);
    }
    pin_cur_ = 0;
    pin_used_ = 0;
}

// Levels are planned and enqueued back to back with one synchronization at the end; each plan
// assumes the children's results the previous factorization produced (none delayed on the
// first).  Levels above the first one whose results differ from that assumption are planned
// again from the actual results and run one at a time.
void NumericGPU::factor(const double *d_val, const double *d_scaling) {
    const Prepared &P = *prep_;
    const int ns = (int)P.sns.size();
    const int nl = (int)P.levels.size();
    RV_HIP(hipStreamSynchronize(stream_));  // nothing of an earlier call may still be in flight
    if ((int)plans_.size() != nl) plans_.assign(nl, LevelPlan());
    for (const LevelPlan &L : plans_) belief_valid_ = belief_valid_ && L.valid;
    if (!belief_valid_) {
        for (LevelPlan &L : plans_) L.valid = false;
        ne_.assign(ns, 0);
        delayed_.assign(ns, {});
        m_.assign(ns, 0);
        k_.assign(ns, 0);
        rows_.assign(ns, {});
    }
    // the solve metadata depends only on each front's order, eliminated count and delayed
    // columns: keep it when a refactorization reproduces them
    const bool had_solve = solve_ready_ && belief_valid_;
    const bool had_solve_plans = belief_valid_;
    const std::vector<int> prev_ne = ne_, prev_m = m_;
    const std::vector<std::vector<int>> prev_delayed = delayed_;
    solve_ready_ = false;
    inv_valid_ = false;
    belief_valid_ = false;  // until this factorization completes
    n1_.assign(ns, 0); n2_.assign(ns, 0); zp_.assign(ns, 0);
    inert_.assign(ns, {0, 0, 0});
    maxl_.assign(ns, 0.0);
    times_ = PhaseTimes();
    pin_reset();
    // pass 1: every level enqueued on its plan
    std::vector<std::vector<int>> bel_ne(nl);
    std::vector<std::vector<std::vector<int>>> bel_del(nl);
    bool invalidate_rest = false;
    bool stable = graph_mode && had_solve_plans;
    for (int lv = 0; stable && lv < nl; ++lv)
        stable = plans_[lv].valid && plans_[lv].uploaded && plan_matches(lv);
    if (stable) {
        // every level runs on its existing plan: one launch of the captured level sequence
        // (captured on the first such call, and again whenever plans or inputs changed)
        if (!gexec_ || g_gen_ != plan_gen_ || g_val_ != d_val || g_scal_ != d_scaling) {
            drop_graph();
            const bool t = timing;
            timing = false;
            RV_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
            capturing_ = true;  // (one stream in the graph: no outer look-ahead)
            for (int lv = 0; lv < nl; ++lv) enqueue_level(lv, d_val, d_scaling);
            capturing_ = false;
            RV_HIP(hipStreamEndCapture(stream_, &graph_));
            timing = t;
            RV_HIP(hipGraphInstantiate(&gexec_, graph_, nullptr, nullptr, 0));
            g_gen_ = plan_gen_;
            g_val_ = d_val;
            g_scal_ = d_scaling;
        }
        for (int lv = 0; lv < nl; ++lv)
            for (int s : P.levels[lv]) {
                bel_ne[lv].push_back(ne_[s]);
                bel_del[lv].push_back(delayed_[s]);
            }
        RV_HIP(hipGraphLaunch(gexec_, stream_));
        ++graph_launches_;
    }
    for (int lv = 0; !stable && lv < nl; ++lv) {
        LevelPlan &L = plans_[lv];
        const bool rebuild = invalidate_rest || !L.valid || !plan_matches(lv);
        if (rebuild) {
            plan_level(lv, !invalidate_rest);
            invalidate_rest = true;  // the arena behind this level is handed out again
            // assume the rebuilt level delays nothing
            for (int s : P.levels[lv]) {
                ne_[s] = k_[s];
                delayed_[s].clear();
            }
        }
        for (int s : P.levels[lv]) {
            bel_ne[lv].push_back(ne_[s]);
            bel_del[lv].push_back(delayed_[s]);
        }
        enqueue_level(lv, d_val, d_scaling);
    }
    RV_HIP(hipStreamSynchronize(stream_));
    int first_bad = -1;
    for (int lv = 0; lv < nl && first_bad < 0; ++lv) {
        read_level(lv);
        const std::vector<int> &fr = P.levels[lv];
        for (size_t i = 0; i < fr.size(); ++i)
            if (ne_[fr[i]] != bel_ne[lv][i] || delayed_[fr[i]] != bel_del[lv][i]) first_bad = lv;
    }
    // pass 2 (only when a level's delays differed from the plans above it).  Pass 1's later
    // levels may have taken the pool blocks of contributions that the levels planned again read:
    // the levels up to first_bad run again first (same plans, same results).
    if (first_bad >= 0) {
        if (!keep_fronts_)
            for (int lv = 0; lv <= first_bad; ++lv) enqueue_level(lv, d_val, d_scaling);
        for (int lv = first_bad + 1; lv < nl; ++lv) {
            plan_level(lv, lv == first_bad + 1);
            enqueue_level(lv, d_val, d_scaling);
            RV_HIP(hipStreamSynchronize(stream_));
            read_level(lv);
        }
    }
    if (chain_err_) {
        // a wait of the persistent outer-block launch ran out (its workgroups were not all resident
        // at once): the results are void; factor again with the per-inner-block launches
        chain_err_ = false;
        la_err_ = false;
        chain_on_ = false;
        for (LevelPlan &L : plans_) L.valid = false;
        drop_graph();
        factor(d_val, d_scaling);
        return;
    }
    if (la_err_) {
        // an outer block planned without failed columns had some, and its trailing update ran
        // beside the delay swaps: the results are void; factor again in stream order (this
        // factorization's failures plan the next one)
        la_err_ = false;
        la_suspend_ = true;
        factor(d_val, d_scaling);
        la_suspend_ = false;
        return;
    }
    belief_valid_ = true;
    obfail_valid_ = true;
    // the solve plan holds the fronts' device addresses: any level planned since it was built
    // (plan_level hands out arena memory anew) makes it stale
    solve_ready_ = had_solve && prev_ne == ne_ && prev_m == m_ && prev_delayed == delayed_ &&
                   sol_plan_gen_ == plan_gen_;
    // the solve plan (and with it the workspace size rv_solve_scratch reports) follows every
    // factorization
    if (!solve_ready_) prepare_solve();
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

bool NumericGPU::plan_matches(int lv) const {
    const LevelPlan &L = plans_[lv];
    const Prepared &P = *prep_;
    if (L.prec_gen != prec_gen_) return false;
    size_t q = 0;
    for (int s : P.levels[lv])
        for (int c : P.children[s]) {
            if (q >= L.child_ne.size() || L.child_ne[q] != ne_[c] || L.child_del[q] != delayed_[c]) return false;
            ++q;
        }
    return q == L.child_ne.size();
}

// Pool blocks of level lv in placement order (largest first): every non-root front of at least
// 1 MiB is a block of its own, the smaller ones share one; each with the highest level of its
// fronts' parents (local indices into the level's front list, m per supernode).  None with
// keep_fronts_.
void NumericGPU::pool_groups(int lv, const std::vector<int> &m,
                             std::vector<std::pair<std::vector<int>, int>> *groups) const {
    const Prepared &P = *prep_;
    groups->clear();
    if (keep_fronts_) return;
    const std::vector<int> &fronts = P.levels[lv];
    std::vector<int> small;
    std::vector<std::pair<size_t, int>> big;
    int small_death = -1;
    size_t small_bytes = 0;
    for (int li = 0; li < (int)fronts.size(); ++li) {
        const int s = fronts[li];
        const int par = P.sns[s].parent;
        if (par < 0) continue;
        const size_t bytes = (size_t)m[s] * m[s] * sizeof(double);
        if (bytes >= ((size_t)1 << 20)) {
            big.emplace_back(bytes, li);
        } else {
            small.push_back(li);
            small_bytes += bytes;
            small_death = std::max(small_death, P.level_of[par]);
        }
    }
    std::stable_sort(big.begin(), big.end(), [](const std::pair<size_t, int> &a, const std::pair<size_t, int> &b) {
        return a.first > b.first;
    });
    bool small_done = small.empty();
    for (const auto &b : big) {
        if (!small_done && small_bytes >= b.first) {
            groups->emplace_back(small, small_death);
            small_done = true;
        }
        groups->emplace_back(std::vector<int>{b.second}, P.level_of[P.sns[fronts[b.second]].parent]);
    }
    if (!small_done) groups->emplace_back(small, small_death);
}

// Highest pool offset of a placement of every level's blocks at the fronts' sizes without
// delays: the size of the first pool chunk (plus a margin for fronts grown by delays).
size_t NumericGPU::pool_estimate() const {
    const Prepared &P = *prep_;
    const int ns = (int)P.sns.size();
    std::vector<int> m0(ns);
    for (int s = 0; s < ns; ++s) m0[s] = P.sns[s].n_owned() + (int)P.sns[s].pattern.size();
    FrontPool vp;
    vp.set_virtual();
    std::vector<std::pair<std::vector<int>, int>> groups;
    for (int lv = 0; lv < (int)P.levels.size(); ++lv) {
        pool_groups(lv, m0, &groups);
        for (const auto &g : groups) {
            size_t bytes = 0;
            for (int li : g.first) {
                const size_t mm = m0[P.levels[lv][li]];
                bytes += (mm * mm + 31) / 32 * 32 * sizeof(double);
            }
            vp.place(lv, g.second, bytes);
        }
    }
    return vp.top();
}

// reset_arena: the first level planned in a pass hands out the arena from its old mark (from
// the start for level 0); later levels of the pass allocate after it.
void NumericGPU::plan_level(int lv, bool reset_arena) {
    const Prepared &P = *prep_;
    LevelPlan &L = plans_[lv];
    if (reset_arena) {
        if (lv == 0) arena_.reset_keep();
        else arena_.reset_to(L.mark);
    }
    for (auto &e : L.sev) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    L = LevelPlan();
    L.mark = arena_.mark();
    ++plan_gen_;
    // working memory of a level's kernels (W panels, block backups, swap lines, APTP state) is
    // dead once the level's launches have run: every level takes it from the start of one
    // shared scratch arena, as the levels run in stream order
    scratch_.reset_keep();
    const std::vector<int> &fronts = P.levels[lv];
    const int nf = (int)fronts.size();
    const int nb = o_.outer_block_size, ib = o_.inner_block_size;
    L.prec_gen = prec_gen_;
    // updates beyond the current outer block: FP64 fronts defer them to k_outer_exact
    // (reference order, one pass per outer block; it holds at most MAX_OB_IB inner blocks, else
    // the inner blocks update every fully-summed column themselves); reduced-precision and
    // FP64-MFMA fronts (codes 1, 2, 4) to one rank-nb MFMA GEMM
    const bool exact_outer = cdiv(nb, ib) <= MAX_OB_IB;
    auto fp = [&](int s) { return (int)fprec_[s]; };
    auto deferred_f = [&](int s) { return fp(s) == 1 || fp(s) == 2 || fp(s) == 4; };
    for (int s : fronts)
        for (int c : P.children[s]) {
            L.child_ne.push_back(ne_[c]);
            L.child_del.push_back(delayed_[c]);
        }
    // ---- front structure (numeric.rs:1238-1262) ----
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
        totM += m_[s];
        totK += std::max(k, 1);
    }
    // ---- working fronts: roots (every front with keep_fronts_) in arena_, each with its factor
    // in place; the others in pool_ blocks that later levels take again once the parent level has
    // assembled them, their eliminated columns copied at the level's end to kept storage in arena_
    // (offsets in doubles rounded to 32: 256-byte aligned fronts) ----
    auto r32 = [](size_t x) { return (x + 31) / 32 * 32; };
    std::vector<double *> fptr(nf, nullptr), lptr(nf, nullptr);
    {
        pool_.drop_from(lv);
        if (!keep_fronts_ && pool_.bytes_reserved() == 0) {
            const size_t est = pool_estimate();
            pool_.reserve(est + est / 8);
        }
        std::vector<std::pair<std::vector<int>, int>> groups;
        pool_groups(lv, m_, &groups);
        for (int li = 0; li < nf; ++li) pool_block_[fronts[li]] = -1;
        for (const auto &g : groups) {
            size_t bytes = 0;
            for (int li : g.first) bytes += r32((size_t)m_[fronts[li]] * m_[fronts[li]]) * sizeof(double);
            const int id = pool_.place(lv, g.second, bytes);
            double *base = (double *)pool_.ptr(id);
            size_t o = 0;
            for (int li : g.first) {
                const int s = fronts[li];
                fptr[li] = base + o;
                o += r32((size_t)m_[s] * m_[s]);
                pool_block_[s] = id;
            }
        }
        size_t totF = 0, totL = 0;
        for (int li = 0; li < nf; ++li) {
            const int s = fronts[li];
            if (fptr[li]) totL += r32((size_t)m_[s] * k_[s]);
            else totF += r32((size_t)m_[s] * m_[s]);
        }
        double *Fbase = (double *)arena_.alloc(std::max<size_t>(totF + totL, 1) * sizeof(double));
        size_t fo = 0, lo = totF;
        for (int li = 0; li < nf; ++li) {
            const int s = fronts[li];
            if (fptr[li]) {
                lptr[li] = Fbase + lo;
                lo += r32((size_t)m_[s] * k_[s]);
            } else {
                fptr[li] = lptr[li] = Fbase + fo;
                fo += r32((size_t)m_[s] * m_[s]);
            }
        }
        L.Fbase = Fbase;
        L.totF = totF + totL;
    }
    // ---- device memory of the level ----
    int *ibase = (int *)arena_.alloc(std::max<size_t>(3 * totM, 1) * sizeof(int));
    int8_t *kbase = (int8_t *)arena_.alloc(std::max<size_t>(totK, 1));
    double *dbase = (double *)arena_.alloc(std::max<size_t>(2 * totK, 1) * sizeof(double));
    L.ibase = ibase;
    L.totM = totM;
    std::vector<FrontDev> lfr(nf);
    {
        size_t mo = 0, ko = 0;
        for (int li = 0; li < nf; ++li) {
            const int s = fronts[li];
            FrontDev d;
            std::memset(&d, 0, sizeof(d));
            d.F = fptr[li];
            d.Lf = lptr[li];
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
            d.flags = (tpp_whole ? FRONT_TPP : 0) | (!deferred_f(s) && !exact_outer ? FRONT_ALLFS : 0) |
                      (fp(s) == 1 ? FRONT_W32 : 0);
            mo += d.m;
            ko += std::max(d.k, 1);
            lfr[li] = d;
            h_fr_[s] = d;
        }
    }
    // ---- work lists ----
    std::vector<ScatterChunk> sch;
    std::vector<int> tpp_list, aptp_list;
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
    struct EaPending { int child, parent, round, size, cne; size_t mapoff; int skip; };
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
                e.child = c; e.parent = s; e.round = (int)ci; e.size = sz; e.cne = ne_[c];
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
    // ---- staging blob ----
    std::vector<char> &stg = L.stage;
    std::vector<int> ids(fronts.begin(), fronts.end());
    L.nf = nf;
    L.o_ids = stage_push(stg, ids.data(), ids.size());
    L.o_lfr = stage_push(stg, lfr.data(), lfr.size());
    {
        std::vector<int2> zitems;
        for (int li = 0; li < nf; ++li)
            for (int c = 0; c < lfr[li].m; c += ZCOLS) zitems.push_back(make_int2(li, c));
        L.o_zl = stage_push(stg, zitems.data(), zitems.size());
        L.nzl = zitems.size();
        // eliminated columns of the pooled fronts into their kept storage (supernode, chunk)
        std::vector<int2> kitems;
        for (int li = 0; li < nf; ++li)
            if (lfr[li].Lf != lfr[li].F)
                for (size_t c = 0; c * KF_CHUNK < (size_t)lfr[li].m * lfr[li].k; ++c)
                    kitems.push_back(make_int2(fronts[li], (int)c));
        L.o_kf = stage_push(stg, kitems.data(), kitems.size());
        L.nkf = kitems.size();
    }
    std::vector<int> rowsflat;
    rowsflat.reserve(totM);
    for (int s : fronts) rowsflat.insert(rowsflat.end(), rows_[s].begin(), rows_[s].end());
    L.o_rows = stage_push(stg, rowsflat.data(), rowsflat.size());
    L.o_sch = stage_push(stg, sch.data(), sch.size());
    L.nsch = sch.size();
    const size_t o_maps = stage_push(stg, maps.data(), maps.size());
    L.o_tpp = stage_push(stg, tpp_list.data(), tpp_list.size());
    L.ntpp = tpp_list.size();
    L.o_aptp = stage_push(stg, aptp_list.data(), aptp_list.size());
    // extend-add items per round; map fields hold blob offsets until the device copy exists
    std::vector<std::vector<EaItem>> ea_rounds(max_children);
    for (const EaPending &e : eap) {
        const int nt = cdiv(e.size, 64);
        for (int ti = 0; ti < nt; ++ti)
            for (int tj = 0; tj <= ti; ++tj) {
                EaItem it;
                it.child = e.child; it.parent = e.parent; it.ti = ti; it.tj = tj;
                it.map = (const unsigned *)(uintptr_t)(o_maps + e.mapoff * sizeof(unsigned));
                it.skip_zero = e.skip; it.cne = e.cne;
                ea_rounds[e.round].push_back(it);
            }
    }
    L.o_ea.resize(max_children);
    L.n_ea.resize(max_children);
    for (int r = 0; r < max_children; ++r) {
        L.o_ea[r] = stage_push(stg, ea_rounds[r].data(), ea_rounds[r].size());
        L.n_ea[r] = ea_rounds[r].size();
    }
    // APTP item lists
    const int na = (int)aptp_list.size();
    L.na = na;
    // work items carry the front's index next to its APTP slot: the kernels read the front
    // descriptor without a dependent load of the slot list
    std::vector<int4> k2items;
    std::vector<int4> k3items;
    // the look-ahead workgroups first, so they are dispatched with the first wave of workgroups
    for (int li = 0; li < na; ++li)
        if (k_[aptp_list[li]] > ib) k3items.push_back(make_int4(li, 0, LOOKAHEAD_Q, aptp_list[li]));
    int max_outer = 0;
    for (int li = 0; li < na; ++li) {
        const int s = aptp_list[li];
        const int m = m_[s], p = k_[s];
        for (int c = 0; c < cdiv(m, 256); ++c) k2items.push_back(make_int4(li, c, s, 0));
        const int qmax = 1 + cdiv((deferred_f(s) || exact_outer) ? std::min(p, nb) : p, 32);
        for (int rt = 0; rt < cdiv(m, 64); ++rt)
            for (int q = 0; q < qmax; ++q) k3items.push_back(make_int4(li, rt, q, s));
        const int nouter = p > nb ? cdiv(p, nb) : 1;
        max_outer = std::max(max_outer, nouter);
    }
    L.max_outer = max_outer;
    L.o_k2 = stage_push(stg, k2items.data(), k2items.size());
    L.nk2 = k2items.size();
    L.o_k3 = stage_push(stg, k3items.data(), k3items.size());
    L.nk3 = k3items.size();
    L.ninner.assign(max_outer, 0);
    for (int ob = 0; ob < max_outer; ++ob)
        for (int s : aptp_list) {
            const int p = k_[s];
            const bool two = p > nb;
            const int nouter = two ? cdiv(p, nb) : 1;
            if (ob >= nouter) continue;
            const int obcols = two ? std::min(nb, p - ob * nb) : p;
            L.ninner[ob] = std::max(L.ninner[ob], cdiv(obcols, ib));
        }
    // one persistent launch per outer block when every workgroup of the level's largest one
    // (outer block 0: leader + panels of CH rows per front) is resident at once and the inner
    // blocks fit the launch's records
    L.chain_ch = 0;
    L.o_chain.clear();
    L.n_chain.clear();
    L.dC = nullptr;
    // (at most CHAIN_MAX_IT inner blocks per outer block also means no front of the level updates every
    // fully-summed column from its inner blocks: exact_outer holds)
    if (na > 0 && chain_on_ && cdiv(nb, ib) <= CHAIN_MAX_IT && ib <= MAX_IB) {
        for (int ch : {64, 128}) {
            size_t tot = 0;
            for (int s : aptp_list) tot += 1 + cdiv(m_[s], ch);
            if (tot <= chain_capacity(ch)) {
                L.chain_ch = ch;
                break;
            }
        }
    }
    if (L.chain_ch) {
        const int ch = L.chain_ch;
        L.o_chain.assign(max_outer, 0);
        L.n_chain.assign(max_outer, 0);
        for (int ob = 0; ob < max_outer; ++ob) {
            // leaders first, then the panels front by front
            std::vector<int4> items;
            std::vector<int> npan(na, 0);
            for (int li = 0; li < na; ++li) {
                const int s = aptp_list[li], p = k_[s];
                const int nouter = p > nb ? cdiv(p, nb) : 1;
                if (ob >= nouter) continue;
                npan[li] = cdiv(m_[s] - ob * nb, ch);
                items.push_back(make_int4(li, 0, s, npan[li]));
            }
            for (int li = 0; li < na; ++li)
                for (int q = 0; q < npan[li]; ++q) items.push_back(make_int4(li, 1 + q, aptp_list[li], npan[li]));
            L.o_chain[ob] = stage_push(stg, items.data(), items.size());
            L.n_chain[ob] = items.size();
        }
        L.dC = (ChainSync *)scratch_.alloc((size_t)na * sizeof(ChainSync));
    }
    // contribution GEMM tiles (mode 0), one list per update precision
    std::array<std::vector<TileItem>, 4> ctiles;
    for (int li = 0; li < nf; ++li) {
        const int s = fronts[li];
        if (P.sns[s].parent < 0) continue;
        const int nfs = m_[s] - k_[s];
        if (nfs <= 0) continue;
        const int nt = cdiv(nfs, 128);
        for (int ti = 0; ti < nt; ++ti)
            for (int tj = 0; tj <= ti; ++tj) ctiles[fp(s) == 4 ? 0 : fp(s)].push_back(TileItem{s, ti, tj, li});
    }
    for (int q = 0; q < 4; ++q) {
        L.o_ct[q] = stage_push(stg, ctiles[q].data(), ctiles[q].size());
        L.nct[q] = ctiles[q].size();
    }
    // W = L D buffers and their k_form_ld items: mode 0 per level front, mode 1 per APTP front
    std::vector<double *> wp0(nf, nullptr), wp1(aptp_list.size(), nullptr), wp1b(aptp_list.size(), nullptr);
    std::vector<int4> f0items, f1items;
    for (int li = 0; li < nf; ++li) {
        const int s = fronts[li];
        if (P.sns[s].parent < 0) continue;
        const int nfs = m_[s] - k_[s];
        if (nfs <= 0 || k_[s] <= 0) continue;
        wp0[li] = (double *)scratch_.alloc((size_t)nfs * k_[s] * sizeof(double));
        for (int rc = 0; rc < cdiv(nfs, 256); ++rc)
            for (int cc = 0; cc < cdiv(k_[s], FC); ++cc) f0items.push_back(make_int4(s, li, rc, cc));
    }
    for (size_t li = 0; li < aptp_list.size(); ++li) {
        const int s = aptp_list[li];
        if (k_[s] <= nb || !(deferred_f(s) || exact_outer)) continue;
        wp1[li] = (double *)scratch_.alloc((size_t)m_[s] * nb * sizeof(double));
        wp1b[li] = (double *)scratch_.alloc((size_t)m_[s] * nb * sizeof(double));
        for (int rc = 0; rc < cdiv(m_[s], 256); ++rc)
            for (int cc = 0; cc < cdiv(nb, FC); ++cc) f1items.push_back(make_int4(s, (int)li, rc, cc));
    }
    L.o_wp0 = stage_push(stg, wp0.data(), wp0.size());
    L.o_wp1 = stage_push(stg, wp1.data(), wp1.size());
    L.o_wp1b = stage_push(stg, wp1b.data(), wp1b.size());
    L.o_f0 = stage_push(stg, f0items.data(), f0items.size());
    L.nf0 = f0items.size();
    L.o_f1 = stage_push(stg, f1items.data(), f1items.size());
    L.nf1 = f1items.size();
    L.ota64 = outer_t64_ && nb % 128 == 0;
    // outer-update tiles (mode 1) per outer block index and kind (exact, FP32, BF16, FP64 MFMA):
    // the tiles in the next outer block's columns first (otiles), then the rest (otiles_b)
    std::vector<std::array<std::vector<TileItem>, 4>> otiles(max_outer), otiles_b(max_outer);
    for (int li = 0; li < na; ++li) {
        const int s = aptp_list[li];
        const int m = m_[s], p = k_[s];
        const bool dfr = deferred_f(s);
        if (p <= nb || !(dfr || exact_outer)) continue;
        const int nouter = cdiv(p, nb);
        const int kind = dfr ? (fp(s) == 4 ? 3 : fp(s)) : 0;
        const int T = dfr ? 128 : 64;  // k_schur tiles / k_outer_exact tiles
        // the next block's columns in 64 x 64 k_schur tiles (FP32, FP64 MFMA; see L.ota64)
        const int TA = (dfr && L.ota64 && kind != 2) ? 64 : T;
        // every outer block: columns beyond it hold later blocks and earlier blocks' delays
        for (int ob = 0; ob < nouter; ++ob) {
            const int obcols = std::min(nb, p - ob * nb);
            const int nr = m - obcols, nc = p - obcols;
            if (nc <= 0) continue;
            const int ntr = cdiv(nr, T), ntc = cdiv(nc, T);
            if (TA != T) {
                // (nb a multiple of 128: both tilings split the columns at nb)
                for (int ti = 0; ti < cdiv(nr, TA); ++ti)
                    for (int tj = 0; tj <= std::min(ti, std::min(cdiv(nc, TA), nb / TA) - 1); ++tj)
                        otiles[ob][kind].push_back(TileItem{s, ti, tj, li});
                for (int ti = 0; ti < ntr; ++ti)
                    for (int tj = nb / T; tj <= std::min(ti, ntc - 1); ++tj)
                        otiles_b[ob][kind].push_back(TileItem{s, ti, tj, li});
                continue;
            }
            for (int ti = 0; ti < ntr; ++ti)
                for (int tj = 0; tj <= std::min(ti, ntc - 1); ++tj)
                    (tj < cdiv(nb, T) ? otiles : otiles_b)[ob][kind].push_back(TileItem{s, ti, tj, li});
        }
    }
    L.o_ot.resize(max_outer);
    L.n_ot.resize(max_outer);
    L.na_ot.resize(max_outer);
    for (int ob = 0; ob < max_outer; ++ob)
        for (int q = 0; q < 4; ++q) {
            L.na_ot[ob][q] = otiles[ob][q].size();
            otiles[ob][q].insert(otiles[ob][q].end(), otiles_b[ob][q].begin(), otiles_b[ob][q].end());
            L.o_ot[ob][q] = stage_push(stg, otiles[ob][q].data(), otiles[ob][q].size());
            L.n_ot[ob][q] = otiles[ob][q].size();
        }
    L.doffs.assign(nf + 1, 0);
    for (int li = 0; li < nf; ++li) L.doffs[li + 1] = L.doffs[li] + std::max(k_[fronts[li]], 1);
    L.o_doffs = stage_push(stg, L.doffs.data(), L.doffs.size());
    // APTP state and scratch: backup (m x ib), swap lines (2 ib x m), line marks (m)
    if (na > 0) {
        std::vector<AptpState> hs(na);
        size_t markbytes = 0;
        for (int li = 0; li < na; ++li) {
            AptpState &S = hs[li];
            std::memset(&S, 0, sizeof(S));
            S.fail_min = INT_MAX;
            const int m = m_[aptp_list[li]];
            S.backup = (double *)scratch_.alloc((size_t)m * ib * sizeof(double));
            S.blkcopy = (double *)scratch_.alloc((size_t)MAX_IB * MAX_IB * sizeof(double));
            S.swapbuf = (double *)scratch_.alloc((size_t)2 * MAX_IB * m * sizeof(double));
            markbytes += ((size_t)m * sizeof(int) + 255) / 256 * 256;
        }
        L.mark_p = (char *)scratch_.alloc(markbytes);
        L.mark_bytes = markbytes;
        size_t mo = 0;
        for (int li = 0; li < na; ++li) {
            hs[li].mark = (int *)(L.mark_p + mo);
            mo += ((size_t)m_[aptp_list[li]] * sizeof(int) + 255) / 256 * 256;
        }
        L.dS = (AptpState *)scratch_.alloc((size_t)na * sizeof(AptpState));
        L.dSnap[0] = (AptpState *)scratch_.alloc((size_t)na * sizeof(AptpState));
        L.dSnap[1] = (AptpState *)scratch_.alloc((size_t)na * sizeof(AptpState));
        L.o_hs = stage_push(stg, hs.data(), hs.size());
    }
    // device copy of the blob; extend-add map fields become device addresses
    L.dstage = (char *)arena_.alloc(stg.size());
    for (int r = 0; r < max_children; ++r) {
        EaItem *items = (EaItem *)(stg.data() + L.o_ea[r]);
        for (size_t i = 0; i < L.n_ea[r]; ++i)
            items[i].map = (const unsigned *)(L.dstage + (size_t)(uintptr_t)items[i].map);
    }
    // k_collect results: ints, max|L| bits, update flops, then the delayed columns
    const size_t rbytes = (size_t)nf * RES_INTS * sizeof(int) + (size_t)nf * 16;
    L.res_del_off = (rbytes + 255) / 256 * 256;
    L.res_bytes = L.res_del_off + (size_t)L.doffs[nf] * sizeof(int);
    L.dres = (char *)arena_.alloc(L.res_bytes);
    L.valid = true;
}

void NumericGPU::ensure_stream2() {
    if (stream2_) return;
    RV_HIP(hipStreamCreate(&stream2_));
    RV_HIP(hipEventCreateWithFlags(&ev_fork_, hipEventDisableTiming));
    RV_HIP(hipEventCreateWithFlags(&ev_join_, hipEventDisableTiming));
}

void NumericGPU::enqueue_level(int lv, const double *d_val, const double *d_scaling) {
    const Prepared &P = *prep_;
    LevelPlan &L = plans_[lv];
    const int nb = o_.outer_block_size, ib = o_.inner_block_size;
    if (!L.uploaded) {
        char *hp = pin_take(L.stage.size());
        std::memcpy(hp, L.stage.data(), L.stage.size());
        RV_HIP(hipMemcpyAsync(L.dstage, hp, L.stage.size(), hipMemcpyHostToDevice, stream_));
        L.uploaded = true;
    }
    for (auto &e : L.sev) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    L.sev.clear();
    L.sev_contrib.clear();
    auto D = [&](size_t off) { return L.dstage + off; };
    // Schur GEMM launches, bracketed by HIP events on their stream when timing is on
    auto timed_schur = [&](hipStream_t sm, int prec, size_t ntiles, const TileItem *tl, AptpState *stp, int mode,
                           double *const *wpp, bool t64 = false) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (timing) {
            RV_HIP(hipEventCreate(&e0));
            RV_HIP(hipEventCreate(&e1));
            RV_HIP(hipEventRecord(e0, sm));
        }
        const dim3 grid(mode == 1 ? spread_grid(ntiles) : (unsigned)ntiles), block(256);
        const unsigned nt = (unsigned)ntiles;
        if (prec == 3)
            hipLaunchKernelGGL(k_schur_exact, grid, block, 0, sm, d_fr_, (const AptpState *)stp, tl, mode, wpp, nt);
        else if (prec == 1 && t64)
            hipLaunchKernelGGL((k_schur<1, 64>), grid, block, 0, sm, d_fr_, (const AptpState *)stp, tl, mode, wpp, nt);
        else if (prec == 1)
            hipLaunchKernelGGL(k_schur<1>, grid, block, 0, sm, d_fr_, (const AptpState *)stp, tl, mode, wpp, nt);
        else if (prec == 2)
            hipLaunchKernelGGL(k_schur<2>, grid, block, 0, sm, d_fr_, (const AptpState *)stp, tl, mode, wpp, nt);
        else if (t64)
            hipLaunchKernelGGL((k_schur<0, 64>), grid, block, 0, sm, d_fr_, (const AptpState *)stp, tl, mode, wpp, nt);
        else
            hipLaunchKernelGGL(k_schur<0>, grid, block, 0, sm, d_fr_, (const AptpState *)stp, tl, mode, wpp, nt);
        if (timing) {
            RV_HIP(hipEventRecord(e1, sm));
            L.sev.emplace_back(e0, e1);
            L.sev_contrib.push_back(mode == 0 ? 1 : 0);
        }
    };
    // deferred exact outer updates, timed with the Schur GEMMs
    auto timed_outer = [&](hipStream_t sm, size_t ntiles, const TileItem *tl, AptpState *stp, double *const *wpp) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (timing) {
            RV_HIP(hipEventCreate(&e0));
            RV_HIP(hipEventCreate(&e1));
            RV_HIP(hipEventRecord(e0, sm));
        }
        hipLaunchKernelGGL(k_outer_exact, dim3(spread_grid(ntiles)), dim3(256), 0, sm, d_fr_,
                           (const AptpState *)stp, tl, wpp, (unsigned)ntiles);
        if (timing) {
            RV_HIP(hipEventRecord(e1, sm));
            L.sev.emplace_back(e0, e1);
            L.sev_contrib.push_back(0);
        }
    };
    // outer look-ahead for this level: the trailing update of outer block ob may run on stream2_
    // beside block ob + 1 when no front of the level had failed columns in block ob in the last
    // factorization (failed columns are swapped into the trailing columns at the block's end)
    auto lookahead_ok = [&](int ob) {
        if (!la_on_ || la_suspend_ || capturing_ || !obfail_valid_ || ob + 1 >= L.max_outer) return false;
        const int *al = (const int *)(L.stage.data() + L.o_aptp);
        for (int li = 0; li < L.na; ++li)
            if (ob >= 63 || ((obfail_[al[li]] >> ob) & 1ull)) return false;
        return true;
    };
    const int nf = L.nf, na = L.na;
    // ---- lower triangles zeroed, per-front descriptors, perm, rows: one launch ----
    {
        const long long nrows = (long long)L.totM;
        const unsigned nb_setup = (unsigned)(L.nzl + cdiv(nf, 256) + nf + (nrows + ROWCOPY - 1) / ROWCOPY);
        hipLaunchKernelGGL(k_level_setup, dim3(nb_setup), dim3(256), 0, stream_, d_fr_, (const int *)D(L.o_ids),
                           (const FrontDev *)D(L.o_lfr), nf, (const int2 *)D(L.o_zl), (int)L.nzl, (int *)L.ibase,
                           (const int *)D(L.o_rows), nrows);
    }
    // ---- assembly ----
    if (L.nsch)
        hipLaunchKernelGGL(k_scatter_orig, dim3((unsigned)L.nsch), dim3(256), 0, stream_, d_fr_,
                           (const ScatterChunk *)D(L.o_sch), P.d_amap, d_val, d_scaling);
    for (size_t r = 0; r < L.n_ea.size(); ++r)
        if (L.n_ea[r])
            hipLaunchKernelGGL(k_extend_add, dim3((unsigned)L.n_ea[r]), dim3(256), 0, stream_, d_fr_,
                               (const EaItem *)D(L.o_ea[r]));
    // ---- TPP fronts ----
    if (L.ntpp) {
        // a level of more fronts than SIMDs: one wavefront per front; otherwise a workgroup of four
        const bool wide = L.ntpp > 1024;
        hipLaunchKernelGGL((wide ? k_tpp<2, 64> : k_tpp<8, 256>), dim3((unsigned)L.ntpp), dim3(wide ? 64 : 256), 0,
                           stream_, d_fr_, (const int *)D(L.o_tpp), 0, o_.threshold, o_.small);
    }
    // ---- APTP fronts ----
    if (na > 0) {
        AptpState *dS = L.dS;
        const int *dl = (const int *)D(L.o_aptp);
        bool la_pending = false;  // a trailing update runs on stream2_ (joined before the next one)
        for (int ob = 0; ob < L.max_outer; ++ob) {
            if (ob == 0)
                hipLaunchKernelGGL(k_outer_first, dim3(na), dim3(256), 0, stream_, d_fr_, dS,
                                   (const AptpState *)D(L.o_hs), dl, nb, ib, L.dC);
            if (L.chain_ch && L.n_chain[ob]) ++chain_launches_;
            if (L.chain_ch == 64 && L.n_chain[ob])
                hipLaunchKernelGGL(k_aptp_outer<64>, dim3((unsigned)L.n_chain[ob]), dim3(256), 0, stream_, d_fr_,
                                   dS, L.dC, (const int4 *)D(L.o_chain[ob]), nb, ib, o_.small, o_.threshold);
            else if (L.chain_ch == 128 && L.n_chain[ob])
                hipLaunchKernelGGL(k_aptp_outer<128>, dim3((unsigned)L.n_chain[ob]), dim3(256), 0, stream_, d_fr_,
                                   dS, L.dC, (const int4 *)D(L.o_chain[ob]), nb, ib, o_.small, o_.threshold);
            for (int it = 0; !L.chain_ch && it < L.ninner[ob]; ++it) {
                hipLaunchKernelGGL(k_block_apply, dim3((unsigned)L.nk2), dim3(256), 0, stream_, d_fr_, dS,
                                   dl, (const int4 *)D(L.o_k2), ob, it, nb, ib, o_.small, o_.threshold);
                hipLaunchKernelGGL(k_block_update, dim3((unsigned)L.nk3), dim3(256), 0, stream_, d_fr_,
                                   dS, dl, (const int4 *)D(L.o_k3), ob, it, nb, ib, o_.small);
                hipLaunchKernelGGL(k_block_finish, dim3(na), dim3(256), 0, stream_, d_fr_, dS, dl, ob,
                                   it, nb, ib);
            }
            // W of block ob in one of two buffers by parity: the trailing update of block ob - 1
            // may still read the other
            const int par = ob & 1;
            double *const *wpp = (double *const *)D(par ? L.o_wp1b : L.o_wp1);
            size_t n_rest = 0;
            for (int q = 0; q < 4; ++q) n_rest += L.n_ot[ob][q] - L.na_ot[ob][q];
            const bool la = n_rest > 0 && lookahead_ok(ob);
            if (la) ensure_stream2();
            // W of block ob; with look-ahead also the states as block ob left them (k_outer_next
            // moves them on to block ob + 1 while the rest of this update still runs)
            if ((L.n_ot[ob][0] || L.n_ot[ob][1] || L.n_ot[ob][2] || L.n_ot[ob][3]) && L.nf1)
                hipLaunchKernelGGL(k_form_ld, dim3((unsigned)L.nf1), dim3(256), 0, stream_, d_fr_, dS,
                                   (const int4 *)D(L.o_f1), 1, wpp, la ? L.dSnap[par] : (AptpState *)nullptr);
            // the columns of block ob + 1 on from here received block ob - 1's trailing update
            if (la_pending) {
                RV_HIP(hipStreamWaitEvent(stream_, ev_join_, 0));
                la_pending = false;
            }
            // next block's columns (all columns without look-ahead) here, the rest beside block ob + 1
            // (kind 3: FP64 MFMA, k_schur<0>); with L.ota64 the next block's columns come in 64 x 64
            // tiles (FP32, FP64 MFMA) and the rest in 128 x 128 tiles, one launch each
            for (int q = 1; q <= 3; ++q) {
                const bool t64 = L.ota64 && q != 2;
                const size_t na_q = L.na_ot[ob][q], n_q = L.n_ot[ob][q];
                const size_t nhere = (la || t64) ? na_q : n_q;
                if (nhere) timed_schur(stream_, q == 3 ? 0 : q, nhere, (const TileItem *)D(L.o_ot[ob][q]), dS, 1, wpp, t64);
                if (!la && t64 && n_q > na_q)
                    timed_schur(stream_, q == 3 ? 0 : q, n_q - na_q, (const TileItem *)D(L.o_ot[ob][q]) + na_q, dS, 1, wpp);
            }
            {
                const size_t nhere = la ? L.na_ot[ob][0] : L.n_ot[ob][0];
                if (nhere) timed_outer(stream_, nhere, (const TileItem *)D(L.o_ot[ob][0]), dS, wpp);
            }
            if (ob + 1 < L.max_outer)
                hipLaunchKernelGGL(k_outer_next, dim3(na), dim3(256), 0, stream_, d_fr_, dS, dl, ob, nb, ib, L.dC,
                                   la ? 1 : 0);
            else
                hipLaunchKernelGGL(k_outer_end, dim3(na), dim3(256), 0, stream_, d_fr_, dS, dl, ob, nb);
            if (la) {
                // forked once block ob + 1 can start, so that its persistent launch and this update
                // begin dispatching together (an update dispatched first would hold every slot)
                RV_HIP(hipEventRecord(ev_fork_, stream_));
                RV_HIP(hipStreamWaitEvent(stream2_, ev_fork_, 0));
                for (int q = 1; q <= 3; ++q)
                    if (L.n_ot[ob][q] > L.na_ot[ob][q])
                        timed_schur(stream2_, q == 3 ? 0 : q, L.n_ot[ob][q] - L.na_ot[ob][q],
                                    (const TileItem *)D(L.o_ot[ob][q]) + L.na_ot[ob][q], L.dSnap[par], 1, wpp);
                if (L.n_ot[ob][0] > L.na_ot[ob][0])
                    timed_outer(stream2_, L.n_ot[ob][0] - L.na_ot[ob][0],
                                (const TileItem *)D(L.o_ot[ob][0]) + L.na_ot[ob][0], L.dSnap[par], wpp);
                RV_HIP(hipEventRecord(ev_join_, stream2_));
                la_pending = true;
                ++la_launches_;
            }
        }
        if (la_pending) {
            RV_HIP(hipStreamWaitEvent(stream_, ev_join_, 0));
            la_pending = false;
        }
        if (o_.failed_pivot_tpp)
            hipLaunchKernelGGL((k_tpp<8, 256>), dim3(na), dim3(256), 0, stream_, d_fr_, dl, 1, o_.threshold, o_.small);
    }
    // ---- kept factor columns of the pooled fronts: final from here on, read beside the
    // contribution GEMMs (which only read them) on stream2_, joined after the collect (measured
    // against a copy beside the next level's pivoting, which slows that latency-bound chain
    // more) ----
    const bool kf_fork = L.nkf && !capturing_;
    if (L.nkf) {
        hipStream_t ks = stream_;
        if (kf_fork) {
            ensure_stream2();
            RV_HIP(hipEventRecord(ev_fork_, stream_));
            RV_HIP(hipStreamWaitEvent(stream2_, ev_fork_, 0));
            ks = stream2_;
        }
        hipLaunchKernelGGL(k_keep_factor, dim3((unsigned)L.nkf), dim3(256), 0, ks, (const FrontDev *)d_fr_,
                           (const int2 *)D(L.o_kf));
        if (kf_fork) RV_HIP(hipEventRecord(ev_join_, stream2_));
    }
    // ---- contribution blocks ----
    if (L.nct[0] + L.nct[1] + L.nct[2] + L.nct[3]) {
        hipLaunchKernelGGL(k_form_ld, dim3((unsigned)L.nf0), dim3(256), 0, stream_, d_fr_,
                           (const AptpState *)nullptr, (const int4 *)D(L.o_f0), 0, (double *const *)D(L.o_wp0),
                           (AptpState *)nullptr);
        for (int q = 0; q < 4; ++q)
            if (L.nct[q])
                timed_schur(stream_, q, L.nct[q], (const TileItem *)D(L.o_ct[q]), nullptr, 0, (double *const *)D(L.o_wp0));
    }
    // ---- collect ----
    hipLaunchKernelGGL(k_collect, dim3(nf), dim3(256), 0, stream_, d_fr_, (const int *)D(L.o_ids),
                       (const long long *)D(L.o_doffs), (int *)L.dres, (int *)(L.dres + L.res_del_off));
    if (kf_fork) RV_HIP(hipStreamWaitEvent(stream_, ev_join_, 0));
    L.hres = pin_take(L.res_bytes);
    RV_HIP(hipMemcpyAsync(L.hres, L.dres, L.res_bytes, hipMemcpyDeviceToHost, stream_));
}

void NumericGPU::read_level(int lv) {
    const Prepared &P = *prep_;
    LevelPlan &L = plans_[lv];
    const std::vector<int> &fronts = P.levels[lv];
    const int nf = L.nf;
    const int *hr = (const int *)L.hres;
    const unsigned long long *hml = (const unsigned long long *)(L.hres + (size_t)nf * RES_INTS * sizeof(int));
    const double *huf = (const double *)(hml + nf);
    const int *hdel = (const int *)(L.hres + L.res_del_off);
    for (int li = 0; li < nf; ++li) {
        const int s = fronts[li];
        const int *r = hr + li * RES_INTS;
        if (r[7] & 1) chain_err_ = true;
        if (r[7] & 2) la_err_ = true;
        obfail_[s] = (unsigned long long)(unsigned)r[8] | ((unsigned long long)(unsigned)r[9] << 32);
        ne_[s] = r[0];
        n1_[s] = r[1];
        n2_[s] = r[2];
        zp_[s] = r[3];
        inert_[s][0] = r[4];
        inert_[s][1] = r[5];
        inert_[s][2] = r[6];
        unsigned long long bits = hml[li];
        double v;
        std::memcpy(&v, &bits, sizeof(v));
        maxl_[s] = v;
        h_fr_[s].ne = ne_[s];
        delayed_[s].assign(hdel + L.doffs[li], hdel + L.doffs[li] + (k_[s] - ne_[s]));
    }
    if (!L.sev.empty()) {
        // algorithmic flops: contribution blocks 2 ne nfs(nfs+1)/2, outer updates from k_outer_end
        double fl = 0, flc = 0, maxnfs = 0;
        for (int li = 0; li < nf; ++li) {
            const int s = fronts[li];
            fl += huf[li];
            if (P.sns[s].parent < 0) continue;
            const double nfs = m_[s] - k_[s];
            flc += (double)ne_[s] * nfs * (nfs + 1.0);
            maxnfs = std::max(maxnfs, nfs);
        }
        fl += flc;
        const bool big = maxnfs >= 2048;
        for (size_t q = 0; q < L.sev.size(); ++q) {
            auto &e = L.sev[q];
            float ms = 0.f;
            RV_HIP(hipEventElapsedTime(&ms, e.first, e.second));
            times_.schur_ms += ms;
            times_.schur_launches += 1;
            if (big && q < L.sev_contrib.size() && L.sev_contrib[q]) {
                times_.big_ms += ms;
                times_.big_launches += 1;
            }
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        L.sev.clear();
        L.sev_contrib.clear();
        times_.schur_flops += fl;
        if (big) times_.big_flops += flc;
    }
}

#ifdef RV_PHASE_TIMING
}  // namespace rv
extern "C" void rv_debug_phases(double *out) {
    unsigned long long h[2][16], n[2];
    (void)hipMemcpyFromSymbol(h, HIP_SYMBOL(rv::g_phase), sizeof(h));
    (void)hipMemcpyFromSymbol(n, HIP_SYMBOL(rv::g_phase_launches), sizeof(n));
    for (int k = 0; k < 2; ++k)
        for (int i = 0; i < 16; ++i) out[k * 17 + i] = n[k] ? (double)h[k][i] / (double)n[k] : 0.0;
    out[16] = (double)n[0];
    out[33] = (double)n[1];
}
namespace rv {
#endif

bool NumericGPU::front_held(int s) const {
    return pool_block_[s] < 0 || !pool_.clobbered(pool_block_[s]);
}

void NumericGPU::export_raw(int s, double *Fh, int *perm, int *rows) const {
    if (!front_held(s))
        throw std::runtime_error("the working front of supernode " + std::to_string(s) +
                                 " was handed to a later level (RV_KEEP_FRONTS=1 keeps it)");
    const FrontDev &F = h_fr_[s];
    const int m = F.m;
    if (Fh) RV_HIP(hipMemcpy(Fh, F.F, (size_t)m * m * sizeof(double), hipMemcpyDeviceToHost));
    if (perm) RV_HIP(hipMemcpy(perm, F.perm, m * sizeof(int), hipMemcpyDeviceToHost));
    if (rows) RV_HIP(hipMemcpy(rows, F.rows, m * sizeof(int), hipMemcpyDeviceToHost));
}

// Contribution block of front s (ContributionBlock, numeric.rs:516-523): the trailing
// (m - ne) x (m - ne) block of the front after its Schur update, rows [delayed columns in APTP
// order, pattern rows].  Returns its order, 0 for a root or a fully eliminated front.
int NumericGPU::export_contribution(int s, double *data, int *rows, int *num_delayed) const {
    const FrontDev &F = h_fr_[s];
    const int m = F.m, k = F.k, ne = ne_[s];
    const int size = m - ne;
    if (prep_->sns[s].parent < 0 || size <= 0) return 0;
    if (!front_held(s))
        throw std::runtime_error("the contribution block of supernode " + std::to_string(s) +
                                 " was handed to a later level (RV_KEEP_FRONTS=1 keeps it)");
    if (num_delayed) *num_delayed = k - ne;
    if (data)
        RV_HIP(hipMemcpy2D(data, (size_t)size * sizeof(double), F.F + ne + (size_t)ne * m,
                           (size_t)m * sizeof(double), (size_t)size * sizeof(double), size,
                           hipMemcpyDeviceToHost));
    if (rows) {
        std::vector<int> perm(m), r(m);
        RV_HIP(hipMemcpy(perm.data(), F.perm, m * sizeof(int), hipMemcpyDeviceToHost));
        RV_HIP(hipMemcpy(r.data(), F.rows, m * sizeof(int), hipMemcpyDeviceToHost));
        int q = 0;
        for (int i = ne; i < k; ++i) rows[q++] = r[perm[i]];
        for (int i = k; i < m; ++i) rows[q++] = r[i];
    }
    return size;
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

// Features of every front for the update-precision classifier, from the supernode structure and
// the last factorization (16 per front, row-major): sizes, update and front flops, position in the
// tree, assembled contribution volume, delays in and out, 2x2 and negative-pivot fractions,
// largest |L|, and the ratio to the parent's order.
void NumericGPU::front_features(float *out) const {
    const Prepared &P = *prep_;
    const int ns = (int)P.sns.size();
    std::vector<int> lvl(ns, 0);
    const int nl = (int)P.levels.size();
    for (int l = 0; l < nl; ++l)
        for (int s : P.levels[l]) lvl[s] = l;
    auto lg = [](double v) { return (float)std::log2(1.0 + std::max(v, 0.0)); };
    for (int s = 0; s < ns; ++s) {
        const double m = m_[s], k = k_[s], ne = ne_[s], nfs = m - k;
        double din = 0.0, cvol = 0.0;
        for (int c : P.children[s]) {
            din += (double)delayed_[c].size();
            const double sc = m_[c] - ne_[c];
            cvol += sc * sc;
        }
        const double ff = ((m - 1) * m * (m + 1) - (m - k - 1) * (m - k) * (m - k + 1)) / 3.0;
        const int par = P.sns[s].parent;
        float *f = out + (size_t)s * 16;
        f[0] = lg(m);
        f[1] = lg(k);
        f[2] = lg(nfs);
        f[3] = (float)(m > 0 ? k / m : 0.0);
        f[4] = lg(ne * nfs * (nfs + 1));
        f[5] = lg(ff);
        f[6] = (float)(nl > 1 ? (double)lvl[s] / (nl - 1) : 0.0);
        f[7] = lg((double)P.children[s].size());
        f[8] = lg(cvol);
        f[9] = (float)(k > 0 ? din / k : 0.0);
        f[10] = (float)(k > 0 ? (k - ne) / k : 0.0);
        f[11] = (float)(2.0 * n2_[s] / std::max(1.0, ne));
        f[12] = (float)std::log10(1.0 + maxl_[s]);
        f[13] = (float)(inert_[s][1] / std::max(1.0, ne));
        f[14] = par < 0 ? 1.0f : 0.0f;
        f[15] = (float)(par < 0 || m_[par] == 0 ? 0.0 : m / m_[par]);
    }
}

// inertia from the D factors (diagonal.rs:391-461), summed over fronts (solver.rs:504-519)
void NumericGPU::inertia(long long out[3]) const {
    // per-front counts from k_collect; the unresolved delays at the roots are not in D
    out[0] = out[1] = out[2] = 0;
    for (const auto &c : inert_)
        for (int i = 0; i < 3; ++i) out[i] += c[i];
}

void NumericGPU::export_front(int s, double *l11, double *l21, signed char *dkind, double *ddiag,
                              double *doff, int *col_indices, int *row_indices,
                              int *local_perm) const {
    const FrontDev &F = h_fr_[s];
    const int m = F.m, k = F.k, ne = ne_[s];
    std::vector<double> A((size_t)m * std::max(ne, 1));
    if (ne) RV_HIP(hipMemcpy(A.data(), F.Lf, (size_t)m * ne * sizeof(double), hipMemcpyDeviceToHost));
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
