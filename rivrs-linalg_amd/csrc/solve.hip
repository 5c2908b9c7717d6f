// solve.hip -- supernodal forward / diagonal / backward solve on the device-resident factors
// (aptp_solve, sparse/src/symmetric/solve.rs:57-326), for nrhs right-hand sides.
//
// The factors are read from each front's kept columns (the front itself for a root, else the
// copy of its first ne columns made at its level's end, same layout): L11 = F[0:ne, 0:ne] (unit
// lower, 2x2 off-diagonals zeroed), L21 = F[ne:m, 0:ne], row/column indices from the solve index.
//
// Each sweep is ONE launch over the whole assembly tree, run as a dataflow: work items are
// ordered so that every item depends only on items before it, each workgroup takes the next
// item with a ticket (atomic counter) when it starts, and waits for what it needs on flags and
// counters in device memory.  A waiting workgroup therefore only waits for workgroups that
// already run, so every wait ends (bounded in any case: a wait that exceeds its limit sets an
// error word the host reports instead of hanging).
//
//  forward (L y = b, solve.rs:89-92), per front with ne > 0:
//    diagonal item c  -- pulls the descendants' updates into the rows of column block c (the
//                        L21*x vectors Y of the fronts below, in postorder, as the reference
//                        subtracts them), streams L[c block rows, 0 : 128c] against the blocks
//                        already solved, solves the 128 x 128 unit triangle and publishes x_c;
//    rows item g      -- Y = L21[128 rows] * x, streamed block by block as they are published.
//  backward (D z = y, L^T x = z, solve.rs:97-141), per front:
//    column chunk item (C, h) -- partial products L21[256 rows of chunk h, block C]^T x_rows
//                        once every ancestor has finished;
//    diagonal item C  -- z = D^-1 y for its columns (MixedDiagonal::solve_in_place,
//                        diagonal.rs:312-365), minus the chunk partials in order, minus
//                        L[block C', block C]^T x_C' for the later blocks C' as they are
//                        published, then the 128 x 128 transposed unit triangle.
// A right-hand-side group of up to 8 columns shares every load of L.  All sums run in a fixed
// order: results do not depend on scheduling.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/rivrs_ldlt.h"
#include "rv_numeric.h"

namespace rv {

namespace {

constexpr int SB = 128;  // diagonal block of the blocked triangular solves (two 64-column halves)
constexpr int GH = 256;  // L21 rows per backward column-chunk item
constexpr int CW = 32;   // columns per streamed chunk of the forward row products (16 per half)
constexpr long long SPIN_LIMIT = 1LL << 22;

// Item kinds.  A front with one diagonal block and at most FUSE_R rows of L21 is one fused item
// per sweep; larger fronts split into diagonal, rows and chunk items.
enum { IT_FDIAG = 0, IT_FROWS = 1, IT_FSMALL = 2, IT_BCHUNK = 3, IT_BDIAG = 4, IT_BSMALL = 5 };
constexpr int FUSE_R = 512;

// Everything the items of one front need, in one record (one dependent load after the item).
struct SolveFront {
    const double *F;     // front, lda = m
    const int *sidx;     // global index of local row i
    const int8_t *dkind;
    const double *ddiag, *doff;
    long long wbase;     // offset of the front's columns in W / Xb
    long long ybase;     // offset of the front's L21 rows in Y
    long long pbase;     // offset of the front's chunk partials in P
    int m, ne;
    int fbase;           // first block flag of the front
    int ndone;           // forward: increments of the front's done counter (rows items, 1 if fused)
    int nchk;            // backward: chunks per column block
    int ech0, ech1;      // effective children (nearest descendants with ne > 0) in ech
    int epar;            // effective parent or -1
    int epar_nblk;       // its column blocks (backward done target)
};

// The solve plan as the kernels see it.
struct PlanDev {
    const SolveFront *sf;
    const int2 *ech;             // (child front, its forward done target)
    const int *ptr;              // pull lists: Y entries per global index
    const long long *ent;
    int nfront, nflag;
};

// Flags and counters, zeroed by k_solve_start; every array holds one slice per RHS group.
struct SyncDev {
    int *tick;   // [2] ticket counters of the two sweeps
    int *err;    // [1] set when a wait exceeded its limit
    int *fflag;  // forward: block solved
    int *fdone;  // forward: rows items of the front finished
    int *bflag;  // backward: block solved
    int *bgc;    // backward: chunk items of the block finished
    int *bdone;  // backward: blocks of the front finished
};

__device__ __forceinline__ double readlane_f64(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Data that one workgroup of a sweep writes and another reads within the same launch (x blocks,
// Y, chunk partials) goes through device-coherent accesses: stores and loads with device scope
// (they bypass the per-XCD L2, whose copies another XCD's writes would leave stale), so no cache
// maintenance fence is needed.  A producer stores its data, waits until its stores have
// completed (s_waitcnt), joins a workgroup barrier, and one thread then sets the flag or bumps the
// counter with a device-scope atomic; a consumer's thread 0 spins on the flag with device-scope
// loads, and after a workgroup barrier every thread reads the data with device-scope loads.
// The factor L is never written during a solve and is read through the caches as usual.
__device__ __forceinline__ double ld_co(const double *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st_co(double *p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// every thread's device-coherent stores complete, then the workgroup barrier
__device__ __forceinline__ void stores_done_barrier() {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
}

// Thread 0 waits until *p >= target; the caller follows with a workgroup barrier.
__device__ __forceinline__ void wait_at_least(const int *p, int target, int *err) {
    long long it = 0;
    while (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        if ((++it & 63) == 0 &&
            (it > SPIN_LIMIT || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
            atomicOr(err, 1);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// after stores_done_barrier(), by one thread
__device__ __forceinline__ void publish_set(int *p) {
    __hip_atomic_store(p, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void publish_add(int *p) {
    __hip_atomic_fetch_add(p, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int take_ticket(int *tick) {
    __shared__ int s_t;
    if (threadIdx.x == 0) s_t = atomicAdd(tick, 1);
    __syncthreads();
    return s_t;
}

// Start of a solve: x_perm[i] = rhs[perm_fwd[i]] (* s[i]) (solver.rs:396-410) for every RHS
// column, and the sweeps' flags and counters cleared.
__global__ void k_solve_start(const double *__restrict__ rhs, long long ld, double *__restrict__ xp, long long n,
                              int nrhs, const int *__restrict__ pf, const double *__restrict__ s,
                              int *__restrict__ sync, long long nsync) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nsync) sync[i] = 0;
    if (i >= n) return;
    const int p = pf[i];
    const double si = s ? s[i] : 1.0;
    for (int q = 0; q < nrhs; ++q) {
        double v = rhs[p + q * ld];
        if (s) v *= si;
        xp[i + q * n] = v;
    }
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

// Pull of columns that no front eliminated (unresolved delays at the roots).
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

// Inverses of the diagonal triangles.  Every 64 x 64 unit lower triangle T of L11 the solve
// uses (columns jb .. jb + n - 1 of a front, n <= 64) is inverted once per factorization, at the
// first solve after it, and X = T^-1 (unit lower) is kept in the unused upper triangle of the
// same diagonal block: X(i, j), i > j, at front position (jb + j, jb + i).  The sweeps then form
// x = X w and x = X^T w as dot products, where the substitution needed 63 dependent steps.
// (Every other reader of a front -- factorization kernels, extend-add, exports -- uses its lower
// triangle only.)
// One wavefront per triangle: lane c holds column c of X in registers and forms it column by
// column of T (x = e_c; for j = 0, 1, ...: x_i -= T(i, j) x_j for i > j, as one FMA per entry),
// with T in LDS by columns (Tt[j][i] = T(i, j)), read by all lanes at the same address.  The
// triangle's order n is rounded up to S = 16, 32 or 64 (rows and columns from n on are zeros and
// are not stored back), so the loops are fixed and unrolled.  The updates run in groups of eight
// rows of one column, each group's T entries loaded three groups ahead; a scheduling barrier after
// every group keeps the unrolled loop's 2016 loads from being hoisted all at once.  Global loads
// and stores go through a buffer resource based at the triangle, with an out-of-range offset for
// the entries a lane skips: no divergent branch, which would keep x's registers apart per branch.
template <int S, int J, int I0>
struct InvNext {
    static constexpr bool more = I0 + 8 < S;
    static constexpr int j = more ? J : J + 1;
    static constexpr int i0 = more ? I0 + 8 : J + 2;
    static constexpr bool valid = j < S - 1;
};
template <int S, int J, int I0>
__device__ __forceinline__ void inv_load(double (&c)[8], const double (*Tt)[64]) {
#pragma unroll
    for (int u = 0; u < 8; ++u) c[u] = (I0 + u < S) ? Tt[J][I0 + u < S ? I0 + u : 0] : 0.0;
}
template <int S, int J, int I0>
__device__ __forceinline__ void inv_groups(double (&x)[S], const double (*Tt)[64], const double (&c0)[8],
                                           const double (&c1)[8], const double (&c2)[8]) {
    using N1 = InvNext<S, J, I0>;
    using N2 = InvNext<S, N1::j, N1::i0>;
    using N3 = InvNext<S, N2::j, N2::i0>;
    double c3[8] = {};
    if constexpr (N1::valid && N2::valid && N3::valid) inv_load<S, N3::j, N3::i0>(c3, Tt);
    const double xj = x[J];
#pragma unroll
    for (int u = 0; u < 8; ++u)
        if (I0 + u < S) x[I0 + u] = fma(-c0[u], xj, x[I0 + u]);
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (N1::valid) inv_groups<S, N1::j, N1::i0>(x, Tt, c1, c2, c3);
}

template <int S>
__device__ __forceinline__ void invert_unit_lower(double (*Tt)[64], __amdgpu_buffer_rsrc_t rb, size_t lda, int n) {
    const int l = threadIdx.x;
    constexpr int OUT = 0x7FFFFFF8;  // beyond the resource: the load returns 0, the store is dropped
    // lane l: row l of T, columns j < min(l, S - 1), 16 loads in flight per round trip (n <= S)
#pragma unroll
    for (int j0 = 0; j0 < S - 1; j0 += 16) {
        double v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int j = j0 + u;
            const int off = (l < n && j < l) ? (int)((l + (size_t)j * lda) * sizeof(double)) : OUT;
            v[u] = (j < S - 1) ? __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rb, off, 0, 0)) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 16; ++u)
            if (j0 + u < S - 1) Tt[j0 + u][l] = v[u];
    }
    __syncthreads();
    double x[S];
#pragma unroll
    for (int i = 0; i < S; ++i) x[i] = (i == l) ? 1.0 : 0.0;
    using N1 = InvNext<S, 0, 1>;
    using N2 = InvNext<S, N1::j, N1::i0>;
    double c0[8], c1[8], c2[8];
    inv_load<S, 0, 1>(c0, Tt);
    inv_load<S, N1::j, N1::i0>(c1, Tt);
    inv_load<S, N2::j, N2::i0>(c2, Tt);
    inv_groups<S, 0, 1>(x, Tt, c0, c1, c2);
    // X(i, l), i > l, at front position (jb + l, jb + i)
#pragma unroll
    for (int i = 1; i < S; ++i) {
        const int off = (i > l && i < n) ? (int)((l + (size_t)i * lda) * sizeof(double)) : OUT;
        typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, x[i]), rb, off, 0, 0);
    }
}

// One launch per order class S (the host groups the triangles), so each holds only its LDS.
template <int S>
__global__ __launch_bounds__(64) void k_diag_inverse(const SolveFront *__restrict__ sf,
                                                     const int4 *__restrict__ tris) {
    __shared__ __align__(16) double Tt[S - 1][64];
    const int4 t = tris[blockIdx.x];
    const SolveFront &Fr = sf[t.x];
    const int jb = t.y, n = t.z;
    const size_t lda = (size_t)Fr.m;
    const __amdgpu_buffer_rsrc_t rb = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<double *>(Fr.F) + jb + (size_t)jb * lda, (short)0, 0x7FFFFFF0, 0x00020000);
    invert_unit_lower<S>(Tt, rb, lda, n);
}

// The diagonal block [jb, jb + SB) in LDS: the inverses X_A, X_B of its two 64 x 64 unit lower
// triangles (rows and columns jb..jb+63 and jb+64..jb+127) as compact strictly-lower triangles
// (entry (r, c), c < r, at r (r - 1) / 2 + c), read from the block's upper triangle, and the tile
// C = L[jb+64 : jb+128, jb : jb+64] between them (64 x 65, padded).  Thread (lane, quarter
// q = tid >> 6) loads X entries (q + 4u, lane) and C entries (lane, q + 4u): lanes over
// consecutive front rows in both (coalesced); the loads are issued at the start of an item, off
// the chain.
constexpr int TRI = 64 * 63 / 2;
struct DiagLds {
    double *tA, *tB;
    double (*C)[65];
};

__device__ __forceinline__ void load_diag_block(const DiagLds &d, const gdouble *L, size_t lda, int jb, int n1,
                                                int n2) {
    const int l = threadIdx.x & 63, q = threadIdx.x >> 6;
    double a[16], b[16], c[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int i = q + 4 * u;  // X row, C column
        a[u] = (i < n1 && l < i) ? L[(jb + l) + (size_t)(jb + i) * lda] : 0.0;
        b[u] = (i < n2 && l < i) ? L[(jb + 64 + l) + (size_t)(jb + 64 + i) * lda] : 0.0;
        c[u] = (l < n2) ? L[(jb + 64 + l) + (size_t)(jb + i) * lda] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int i = q + 4 * u;
        if (l < i) {
            d.tA[i * (i - 1) / 2 + l] = a[u];
            d.tB[i * (i - 1) / 2 + l] = b[u];
        }
        d.C[l][i] = c[u];
    }
}

// x = X w for a 64 x 64 unit lower triangle's inverse X (compact): wavefront w takes right-hand
// sides w, w + 4, ...; lane i forms x_i = w_i + sum_{j<i} X(i, j) w_j with its row of X in
// registers and w_j by readlane, the sum in four interleaved partial sums combined in a fixed
// order.  ws[q * SB + off + i] holds w on entry and x on return; rows >= nn take no part.
// part (may be null): four column-quarter sums per right-hand side, part[(4q + k) * 64 + i], taken
// off w first as ((p0 + p1) + p2) + p3 (the update by the tile before the triangle, folded into
// the read so that it needs no barrier of its own).
__device__ __forceinline__ void tri_fwd(const double *X, double *ws, int off, int nn, int nq,
                                        const double *part = nullptr) {
    const int lane = threadIdx.x & 63;
    if ((threadIdx.x >> 6) >= nq) return;
    const int rb = lane * (lane - 1) / 2;
    double row[64];
#pragma unroll
    for (int j = 0; j < 63; ++j) row[j] = (j < lane) ? X[rb + j] : 0.0;
    row[63] = 0.0;
    for (int q = threadIdx.x >> 6; q < nq; q += 4) {
        double wi = (lane < nn) ? ws[q * SB + off + lane] : 0.0;
        if (part && lane < nn)
            wi -= ((part[(q * 4) * 64 + lane] + part[(q * 4 + 1) * 64 + lane]) + part[(q * 4 + 2) * 64 + lane]) +
                  part[(q * 4 + 3) * 64 + lane];
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
#pragma unroll
        for (int j = 0; j < 64; j += 4) {
            a0 += row[j] * readlane_f64(wi, j);
            a1 += row[j + 1] * readlane_f64(wi, j + 1);
            a2 += row[j + 2] * readlane_f64(wi, j + 2);
            a3 += row[j + 3] * readlane_f64(wi, j + 3);
        }
        if (lane < nn) ws[q * SB + off + lane] = wi + ((a0 + a1) + (a2 + a3));
    }
}

// x = X^T w: lane i forms x_i = w_i + sum_{j>i} X(j, i) w_j with column i of X in registers;
// part (may be null): part[q * 64 + i] is taken off w first.
__device__ __forceinline__ void tri_bwd(const double *X, double *ws, int off, int nn, int nq,
                                        const double *part = nullptr) {
    const int lane = threadIdx.x & 63;
    if ((threadIdx.x >> 6) >= nq) return;
    double col[64];
    col[0] = 0.0;
#pragma unroll
    for (int j = 1; j < 64; ++j) col[j] = (j > lane) ? X[j * (j - 1) / 2 + lane] : 0.0;
    for (int q = threadIdx.x >> 6; q < nq; q += 4) {
        double wi = (lane < nn) ? ws[q * SB + off + lane] : 0.0;
        if (part && lane < nn) wi -= part[q * 64 + lane];
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
#pragma unroll
        for (int j = 0; j < 64; j += 4) {
            a0 += col[j] * readlane_f64(wi, j);
            a1 += col[j + 1] * readlane_f64(wi, j + 1);
            a2 += col[j + 2] * readlane_f64(wi, j + 2);
            a3 += col[j + 3] * readlane_f64(wi, j + 3);
        }
        if (lane < nn) ws[q * SB + off + lane] = wi + ((a0 + a1) + (a2 + a3));
    }
}

// Forward substitution of one diagonal block for nq right-hand sides (all threads of the
// workgroup): ws[q * SB + i] holds the block's w on entry and its x on return.
template <int NR>
__device__ __forceinline__ void fwd_block(const DiagLds &d, double *ws, double *part, int n1, int n2, int nq) {
    const int tid = threadIdx.x, i = tid & 63, jq = tid >> 6;
    tri_fwd(d.tA, ws, 0, n1, nq);
    __syncthreads();
    if (n2 > 0) {
        // C x_A: column quarters combined in a fixed order
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            if (q < nq) {
                double s = 0.0;
#pragma unroll
                for (int u = 0; u < 16; ++u) s += d.C[i][jq + 4 * u] * ws[q * SB + jq + 4 * u];
                part[(q * 4 + jq) * 64 + i] = s;
            }
        }
        __syncthreads();
        tri_fwd(d.tB, ws, 64, n2, nq, part);  // w_B - C x_A, then the triangle
        __syncthreads();
    }
}

// Cross-lane moves for the butterflies, all on the VALU (no LDS round trip): the 32- and 16-lane
// exchanges by v_permlane32_swap / v_permlane16_swap, the 8-lane one by a DPP rotation within rows
// of 16, the 2- and 1-lane ones by DPP quad permutations.
template <bool D32>
__device__ __forceinline__ void perm_swap(double &a, double &b) {
    const unsigned long long x = (unsigned long long)__double_as_longlong(a);
    const unsigned long long y = (unsigned long long)__double_as_longlong(b);
    const unsigned xl = (unsigned)x, xh = (unsigned)(x >> 32), yl = (unsigned)y, yh = (unsigned)(y >> 32);
    auto lo = D32 ? __builtin_amdgcn_permlane32_swap(xl, yl, false, false) : __builtin_amdgcn_permlane16_swap(xl, yl, false, false);
    auto hi = D32 ? __builtin_amdgcn_permlane32_swap(xh, yh, false, false) : __builtin_amdgcn_permlane16_swap(xh, yh, false, false);
    a = __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0]));
    b = __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
}

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Column sums over the 64 lanes of a wavefront for 8 columns at once, a fixed order: a
// reduce-scatter butterfly over lane bits 5, 4, 3 (8 -> 1 values per lane), then sums over lane
// bits 2, 1, 0; at every stage the lower lane of a pair contributes first.  p[c] is the lane's
// partial of column c; on return every lane holds the sum of column ((lane >> 3) & 7) in p[0].
__device__ __forceinline__ void wave_colsum8(double (&p)[8]) {
    const int lane = threadIdx.x & 63;
    // bit 5: after the swap a holds (own, partner's) p[i] on lower lanes and (partner's, own)
    // p[i + 4] on upper lanes, paired with b: a + b is (keep + received) resp. (received + keep)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        double a = p[i], b = p[i + 4];
        perm_swap<true>(a, b);
        p[i] = a + b;
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        double a = p[i], b = p[i + 2];
        perm_swap<false>(a, b);
        p[i] = a + b;
    }
    {
        const bool hi = (lane & 8) != 0;
        const double keep = hi ? p[1] : p[0];
        const double send = hi ? p[0] : p[1];
        const double recv = dpp_f64<0x128>(send);  // row_ror:8 = lane ^ 8 within rows of 16
        p[0] = hi ? recv + keep : keep + recv;
    }
    {
        const double other = __shfl_xor(p[0], 4);
        p[0] = (lane & 4) ? other + p[0] : p[0] + other;
    }
    {
        const double other = dpp_f64<0x4E>(p[0]);  // quad_perm [2,3,0,1]
        p[0] = (lane & 2) ? other + p[0] : p[0] + other;
    }
    {
        const double other = dpp_f64<0xB1>(p[0]);  // quad_perm [1,0,3,2]
        p[0] = (lane & 1) ? other + p[0] : p[0] + other;
    }
}

// Backward substitution of one diagonal block with L11^T (all threads): ws holds the block's
// right-hand side, from which every later block and L21 have been removed, and its x on return.
// C^T x_B: wavefront jq sums its sixteen columns jq + 4u over the 64 rows with the butterfly.
template <int NR>
__device__ __forceinline__ void bwd_block(const DiagLds &d, double *ws, double *part, int n1, int n2, int nq) {
    const int tid = threadIdx.x, i = tid & 63, jq = tid >> 6;
    if (n2 > 0) {
        tri_bwd(d.tB, ws, 64, n2, nq);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            if (q < nq) {
                const double xb = (i < n2) ? ws[q * SB + 64 + i] : 0.0;
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    double p[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) p[u] = d.C[i][jq + 4 * (8 * h + u)] * xb;
                    wave_colsum8(p);
                    if ((i & 7) == 0) part[q * 64 + jq + 4 * (8 * h + ((i >> 3) & 7))] = p[0];
                }
            }
        }
        __syncthreads();
    }
    tri_bwd(d.tA, ws, 0, n1, nq, n2 > 0 ? part : nullptr);  // w_A - C^T x_B, then the triangle
    __syncthreads();
}

// Pull of the descendants' updates into column i of a front: x[t] - sum of the Y entries of
// global index t = sidx[i], in list order (postorder of the fronts that produced them).
__device__ __forceinline__ double pull_column(int t, const int *__restrict__ ptr, const long long *__restrict__ ent,
                                              const double *__restrict__ Yq, const double *__restrict__ xq) {
    const int e1 = ptr[t + 1];
    int e = ptr[t];
    double acc = 0.0;
    for (; e + 4 <= e1; e += 4) {
        const long long a0 = ent[e], a1 = ent[e + 1], a2 = ent[e + 2], a3 = ent[e + 3];
        const double y0 = ld_co(Yq + a0), y1 = ld_co(Yq + a1), y2 = ld_co(Yq + a2), y3 = ld_co(Yq + a3);
        acc = (((acc + y0) + y1) + y2) + y3;
    }
    for (; e < e1; ++e) acc += ld_co(Yq + ent[e]);
    return xq[t] - acc;
}

// Row products of the forward sweep: acc[q] (thread tid < 128, row r0 + tid) -= L[r0 + tid, 0:ncols] x,
// the columns streamed in chunks of CW (thread (row ri, half ch) holds 16 columns of a chunk,
// the next chunk's loads in flight during the current products), each 128-column block's x
// read from W once its flag is set -- or, with flags == nullptr (ncols <= SB), taken from xs as
// the caller filled it (zeros beyond ncols).  Per chunk the two halves' sums are subtracted in
// order, the same arithmetic either way.
template <int NR>
__device__ __forceinline__ void fwd_chunk(const double (&v)[16], int k, int nq, double (&acc)[NR], const double *xs,
                                          double *ps) {
    const int tid = threadIdx.x, ri = tid & 127, ch = tid >> 7;
    const int xo = (k * CW) % SB + 16 * ch;
    double *pk = ps + (k & 1) * (NR * 256);
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        if (q < nq) {
            double s = 0.0;
#pragma unroll
            for (int u = 0; u < 16; ++u) s += v[u] * xs[q * SB + xo + u];
            pk[q * 256 + ch * 128 + ri] = s;
        }
    }
    __syncthreads();
    if (tid < 128) {
#pragma unroll
        for (int q = 0; q < NR; ++q)
            if (q < nq) acc[q] -= pk[q * 256 + tid] + pk[q * 256 + 128 + tid];
    }
}

// The four chunks of one block at once (the last block, whose loads were all issued before its
// flag): the four partial sums per thread into four slices of ps, one barrier, then the chunks'
// sums subtracted in chunk order -- the arithmetic of four fwd_chunk calls.
template <int NR>
__device__ __forceinline__ void fwd_chunk4(const double (&v0)[16], const double (&v1)[16], const double (&v2)[16],
                                           const double (&v3)[16], int k, int nq, double (&acc)[NR],
                                           const double *xs, double *ps) {
    const int tid = threadIdx.x, ri = tid & 127, ch = tid >> 7;
    auto part = [&](const double (&v)[16], int kk, int slot) {
        const int xo = (kk * CW) % SB + 16 * ch;
        double *pk = ps + slot * (NR * 256);
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            if (q < nq) {
                double s = 0.0;
#pragma unroll
                for (int u = 0; u < 16; ++u) s += v[u] * xs[q * SB + xo + u];
                pk[q * 256 + ch * 128 + ri] = s;
            }
        }
    };
    part(v0, k, 0);
    part(v1, k + 1, 1);
    part(v2, k + 2, 2);
    part(v3, k + 3, 3);
    __syncthreads();
    if (tid < 128) {
#pragma unroll
        for (int slot = 0; slot < 4; ++slot) {
            const double *pk = ps + slot * (NR * 256);
#pragma unroll
            for (int q = 0; q < NR; ++q)
                if (q < nq) acc[q] -= pk[q * 256 + tid] + pk[q * 256 + 128 + tid];
        }
    }
}

template <int NR, bool FLAGS>
__device__ __forceinline__ void fwd_rows_product(const gdouble *L, size_t lda, int r0, int nr, int ncols,
                                                 const int *flags, int *err, const double *Wf, long long wtotal,
                                                 int nq, double (&acc)[NR], double *xs, double *ps) {
    const int tid = threadIdx.x, ri = tid & 127, ch = tid >> 7;
    const bool rin = ri < nr;
    const int nchunk = (ncols + CW - 1) / CW;
    const gdouble *Lr = L + (r0 + ri);
    auto load = [&](int k, double (&v)[16]) {
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int col = k * CW + 16 * ch + u;
            v[u] = (rin && col < ncols) ? Lr[(size_t)col * lda] : 0.0;
        }
    };
    auto take_block = [&](int blk) {  // wait for block blk's x and stage it in xs
        __syncthreads();  // the previous block's x is no longer read
        if (tid == 0) wait_at_least(flags + blk, 1, err);
        __syncthreads();
        for (int e = tid; e < nq * SB; e += 256) {
            const int q = e / SB, j = e % SB, col = blk * SB + j;
            xs[e] = (col < ncols) ? ld_co(Wf + q * wtotal + col) : 0.0;
        }
        __syncthreads();
    };
    // the last block's chunks are all loaded before its flag is awaited: that wait ends the chain
    // step of the block before, and only the products remain after it
    const int klast = FLAGS ? ((ncols + SB - 1) / SB - 1) * (SB / CW) : nchunk;
    double cur[16], nxt[16];
    if (nchunk > 0) load(0, cur);
    for (int k = 0; k < nchunk; ++k) {
        if (FLAGS && k == klast) {
            static_assert(SB / CW == 4, "four chunks per block");
            double c1[16], c2[16], c3[16];
            load(k + 1, c1);
            load(k + 2, c2);
            load(k + 3, c3);
            take_block(k / (SB / CW));
            fwd_chunk4<NR>(cur, c1, c2, c3, k, nq, acc, xs, ps);
            break;
        }
        if (FLAGS && (k * CW) % SB == 0) take_block(k / (SB / CW));
        if (k + 1 < nchunk) load(k + 1, nxt);
        fwd_chunk<NR>(cur, k, nq, acc, xs, ps);
#pragma unroll
        for (int u = 0; u < 16; ++u) cur[u] = nxt[u];
    }
}

// Transposed products for up to 128 columns [c0, c0 + ncols): out[q * SB + j] = sum over rows
// i in [r0, r0 + nr) of L[i, c0 + j] * x(q, i).  Wavefront w takes columns 32w .. 32w + 31 in
// four batches of 8; lanes take rows r0 + lane + 64u (u < RPL), xv[q][u] being the lane's x of
// that row.  Every load of L serves all nq right-hand sides.
template <int NR, int RPL>
__device__ __forceinline__ void tcol_product(const gdouble *L, size_t lda, int r0, int nr, int c0, int ncols,
                                             const double (&xv)[NR][RPL], int nq, double *out) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll 1
    for (int cb = 32 * wv; cb < 32 * wv + 32; cb += 8) {
        if (cb >= ncols) break;
        double lv[RPL][8];
#pragma unroll
        for (int u = 0; u < RPL; ++u) {
            const int i = lane + 64 * u;
#pragma unroll
            for (int c = 0; c < 8; ++c)
                lv[u][c] = (i < nr && cb + c < ncols) ? L[(size_t)(r0 + i) + (size_t)(c0 + cb + c) * lda] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            if (q < nq) {
                double p[8];
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    double s = 0.0;
#pragma unroll
                    for (int u = 0; u < RPL; ++u) s += lv[u][c] * xv[q][u];
                    p[c] = s;
                }
                wave_colsum8(p);
                const int col = cb + ((lane >> 3) & 7);
                if ((lane & 7) == 0 && col < ncols) out[q * SB + col] = p[0];
            }
        }
    }
}

// A transposed tile of at most 128 rows [r0, r0 + nr) x 128 columns [c0, c0 + ncols) wholly in
// registers (wavefront w: columns 32w .. 32w + 31 in four batches of 8; lanes: rows lane, lane + 64),
// so it can be loaded before the x it multiplies is available.
__device__ __forceinline__ void tcol_load2(const gdouble *L, size_t lda, int r0, int nr, int c0, int ncols,
                                           double (&lv)[4][2][8]) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int i = lane + 64 * u;
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const int col = 32 * wv + 8 * b + c;
                lv[b][u][c] = (i < nr && col < ncols) ? L[(size_t)(r0 + i) + (size_t)(c0 + col) * lda] : 0.0;
            }
        }
}

// out[q * SB + j] = sum over the tile's rows of tile(i, j) x(q, i), the arithmetic of tcol_product
template <int NR>
__device__ __forceinline__ void tcol_compute2(const double (&lv)[4][2][8], const double (&xv)[NR][2], int nq,
                                              int ncols, double *out) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const int cb = 32 * wv + 8 * b;
        if (cb >= ncols) break;
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            if (q < nq) {
                double p[8];
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    double sum = 0.0;
#pragma unroll
                    for (int u = 0; u < 2; ++u) sum += lv[b][u][c] * xv[q][u];
                    p[c] = sum;
                }
                wave_colsum8(p);
                const int col = cb + ((lane >> 3) & 7);
                if ((lane & 7) == 0 && col < ncols) out[q * SB + col] = p[0];
            }
        }
    }
}

// ===========================================================================
// Forward sweep (one launch): ticket order = item order.
// ===========================================================================
template <int NR>
__global__ __launch_bounds__(256, 2) void k_fwd_sweep(PlanDev P, const int4 *__restrict__ items, int nwork, int ngrp,
                                                   int nrhs, double *__restrict__ x, long long n,
                                                   double *__restrict__ W, long long wtotal, double *__restrict__ Y,
                                                   long long ytotal, SyncDev S) {
    const int t = take_ticket(S.tick);
    if (t >= nwork) return;
    const int4 it = items[t / ngrp];
    const int grp = t % ngrp;
    const int q0 = grp * NR, nq = min(NR, nrhs - q0);
    const int li = it.x;
    const SolveFront &Fr = P.sf[li];
    const int ne = Fr.ne, m = Fr.m;
    const size_t lda = (size_t)m;
    const gdouble *L = gbl(Fr.F);
    const int tid = threadIdx.x;
    const int *flags = S.fflag + (size_t)grp * P.nflag + Fr.fbase;
    const double *Wf = W + (size_t)q0 * wtotal + Fr.wbase;
    __shared__ double xs[NR * SB];
    __shared__ double ps[4 * NR * 256];  // chunk partials (two slices in turn, four for a block)
    __shared__ double tri[2 * TRI];
    __shared__ double cti[64][65];
    __shared__ double ws[NR * SB];
    const DiagLds dl{tri, tri + TRI, cti};
    double acc[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) acc[q] = 0.0;
    if (it.y == IT_FDIAG || it.y == IT_FSMALL) {
        const int c = it.z, j0 = c * SB;
        const int nb = min(SB, ne - j0), n1 = min(64, nb), n2 = nb - n1;
        load_diag_block(dl, L, lda, j0, n1, n2);
        // the descendants' Y vectors are complete once the effective children are
        if (tid == 0) {
            for (int e = Fr.ech0; e < Fr.ech1; ++e) {
                const int2 ch = P.ech[e];
                wait_at_least(S.fdone + (size_t)grp * P.nfront + ch.x, ch.y, S.err);
            }
        }
        __syncthreads();
        if (tid < nb) {
            const int tg = gbl(Fr.sidx)[j0 + tid];
            if ((unsigned)tg < (unsigned)n) {
#pragma unroll
                for (int q = 0; q < NR; ++q)
                    if (q < nq)
                        acc[q] = pull_column(tg, P.ptr, P.ent, Y + (size_t)(q0 + q) * ytotal, x + (size_t)(q0 + q) * n);
            } else {
                atomicOr(S.err, 2);  // an index outside the system: the solve plan does not fit the fronts
            }
        }
        if (it.y == IT_FDIAG) fwd_rows_product<NR, true>(L, lda, j0, nb, j0, flags, S.err, Wf, wtotal, nq, acc, xs, ps);
        if (tid < nb) {
#pragma unroll
            for (int q = 0; q < NR; ++q) ws[q * SB + tid] = acc[q];
        }
        __syncthreads();
        fwd_block<NR>(dl, ws, ps, n1, n2, nq);
        if (it.y == IT_FDIAG) {
            // later blocks and the rows items of this front read x_c
            if (tid < nb) {
#pragma unroll
                for (int q = 0; q < NR; ++q)
                    if (q < nq) st_co(W + (size_t)(q0 + q) * wtotal + Fr.wbase + j0 + tid, ws[q * SB + tid]);
            }
            stores_done_barrier();
            if (tid == 0) publish_set(S.fflag + (size_t)grp * P.nflag + Fr.fbase + c);
            return;
        }
        // fused front: x stays in LDS for its own rows (W is read by the next launch only)
        for (int e = tid; e < nq * SB; e += 256) {
            const int q = e / SB, j = e % SB;
            const double v = (j < nb) ? ws[e] : 0.0;
            xs[e] = v;
            if (j < nb) W[(size_t)(q0 + q) * wtotal + Fr.wbase + j] = v;
        }
        __syncthreads();
        for (int r0 = ne; r0 < m; r0 += SB) {
            const int nr = min(SB, m - r0);
            double a2[NR];
#pragma unroll
            for (int q = 0; q < NR; ++q) a2[q] = 0.0;
            fwd_rows_product<NR, false>(L, lda, r0, nr, ne, nullptr, S.err, Wf, wtotal, nq, a2, xs, ps);
            if (tid < nr) {
#pragma unroll
                for (int q = 0; q < NR; ++q)
                    if (q < nq) st_co(Y + (size_t)(q0 + q) * ytotal + Fr.ybase + (r0 - ne) + tid, -a2[q]);
            }
            __syncthreads();  // ps of this row block is read before the next one writes it
        }
        stores_done_barrier();
        if (tid == 0) publish_add(S.fdone + (size_t)grp * P.nfront + li);
        return;
    }
    // IT_FROWS: 128 rows of L21
    const int r0 = ne + it.z * SB, nr = min(SB, m - r0);
    fwd_rows_product<NR, true>(L, lda, r0, nr, ne, flags, S.err, Wf, wtotal, nq, acc, xs, ps);
    if (tid < nr) {
#pragma unroll
        for (int q = 0; q < NR; ++q)
            if (q < nq) st_co(Y + (size_t)(q0 + q) * ytotal + Fr.ybase + (r0 - ne) + tid, -acc[q]);
    }
    stores_done_barrier();
    if (tid == 0) publish_add(S.fdone + (size_t)grp * P.nfront + li);
}

// z = D^-1 y for column jj of a front (MixedDiagonal::solve_in_place, diagonal.rs:312-365):
// 1x1 -> y / d (0 for a zero pivot); a 2x2 by Cramer's rule (0 for a singular block).
__device__ __forceinline__ double dsolve_col(const SolveFront &F, int jj, const double *__restrict__ y) {
    const int kd = gbl(F.dkind)[jj];
    if (kd == 1) {
        const double d = gbl(F.ddiag)[jj];
        return (d == 0.0) ? 0.0 : y[jj] / d;
    }
    const int f = (kd == 2) ? jj : jj - 1;  // first column of the 2x2
    const double a = gbl(F.ddiag)[f], b = gbl(F.doff)[f], cc = gbl(F.ddiag)[f + 1];
    const double det = a * cc - b * b;
    if (det == 0.0) return 0.0;
    const double r1 = y[f], r2 = y[f + 1];
    return (kd == 2) ? (cc * r1 - b * r2) / det : (a * r2 - b * r1) / det;
}

// x of the L21 rows [ne + rb, ne + rb + nr) of a front, lane rows rb + lane + 64u, for the chunk
// products of the backward sweep (device-coherent: ancestors write them in the same launch)
template <int NR>
__device__ __forceinline__ void gather_rows_x(const SolveFront &Fr, int rb, int nr, const double *__restrict__ x,
                                              long long n, int q0, int nq, double (&xv)[NR][GH / 64], int *err) {
    const int lane = threadIdx.x & 63;
    const gint *ri = gbl(Fr.sidx) + Fr.ne + rb;
#pragma unroll
    for (int u = 0; u < GH / 64; ++u) {
        const int i = lane + 64 * u;
        int g = (i < nr) ? ri[i] : -1;
        if (g >= n) {
            atomicOr(err, 2);  // an index outside the system: the solve plan does not fit the fronts
            g = -1;
        }
#pragma unroll
        for (int q = 0; q < NR; ++q) xv[q][u] = (g >= 0 && q < nq) ? ld_co(x + (size_t)(q0 + q) * n + g) : 0.0;
    }
}

// ===========================================================================
// Backward sweep (one launch).
// ===========================================================================
template <int NR>
__global__ __launch_bounds__(256, 2) void k_bwd_sweep(PlanDev P, const int4 *__restrict__ items, int nwork, int ngrp,
                                                   int nrhs, double *__restrict__ x, long long n,
                                                   const double *__restrict__ W, double *__restrict__ Xb,
                                                   long long wtotal, double *__restrict__ Pp, long long ptotal,
                                                   SyncDev S) {
    const int t = take_ticket(S.tick + 1);
    if (t >= nwork) return;
    const int4 it = items[t / ngrp];
    const int grp = t % ngrp;
    const int q0 = grp * NR, nq = min(NR, nrhs - q0);
    const int li = it.x;
    const SolveFront &Fr = P.sf[li];
    const int ne = Fr.ne, m = Fr.m, r = m - ne;
    const size_t lda = (size_t)m;
    const gdouble *L = gbl(Fr.F);
    const int tid = threadIdx.x, lane = tid & 63;
    const int nblk = (ne + SB - 1) / SB;
    const int fb = grp * P.nflag + Fr.fbase;
    const int nch = Fr.nchk;
    __shared__ double cs[NR * SB];
    __shared__ double tri[2 * TRI];
    __shared__ double cti[64][65];
    __shared__ double ws[NR * SB];
    __shared__ double part[NR * 4 * 64];
    const DiagLds dl{tri, tri + TRI, cti};
    // chunk items and fused fronts read the x of L21 rows: every ancestor has finished
    if (it.y != IT_BDIAG) {
        if (tid == 0 && Fr.epar >= 0)
            wait_at_least(S.bdone + (size_t)grp * P.nfront + Fr.epar, Fr.epar_nblk, S.err);
        __syncthreads();
    }
    if (it.y == IT_BCHUNK) {
        const int C = it.z / nch, h = it.z % nch;
        const int c0 = C * SB, ncols = min(SB, ne - c0);
        const int rb = h * GH, nr = min(GH, r - rb);
        double xv[NR][GH / 64];
        gather_rows_x<NR>(Fr, rb, nr, x, n, q0, nq, xv, S.err);
        tcol_product<NR, GH / 64>(L, lda, ne + rb, nr, c0, ncols, xv, nq, cs);
        __syncthreads();
        for (int e = tid; e < nq * SB; e += 256) {
            const int q = e / SB, j = e % SB;
            if (j < ncols) st_co(Pp + (size_t)(q0 + q) * ptotal + Fr.pbase + ((size_t)C * nch + h) * SB + j, cs[e]);
        }
        stores_done_barrier();
        if (tid == 0) publish_add(S.bgc + fb + C);
        return;
    }
    // IT_BDIAG (column block C) or IT_BSMALL (the whole front, C = 0)
    const int C = it.z, j0 = C * SB;
    const int nb = min(SB, ne - j0), n1 = min(64, nb), n2 = nb - n1;
    // a fused front loads its diagonal block after the chunk products (fewer registers live
    // across them); a diagonal item loads it first, its loads in flight during the waits
    if (it.y == IT_BDIAG) load_diag_block(dl, L, lda, j0, n1, n2);
    double acc[NR];
    const int st = (tid < nb) ? gbl(Fr.sidx)[j0 + tid] : 0;
#pragma unroll
    for (int q = 0; q < NR; ++q)
        acc[q] = (tid < nb && q < nq) ? dsolve_col(Fr, j0 + tid, W + (size_t)(q0 + q) * wtotal + Fr.wbase) : 0.0;
    if (it.y == IT_BSMALL) {
        // L21^T x_rows chunk by chunk, subtracted in chunk order (as the chunk partials are)
        for (int h = 0; h < nch; ++h) {
            const int rb = h * GH, nr = min(GH, r - rb);
            double xv[NR][GH / 64];
            gather_rows_x<NR>(Fr, rb, nr, x, n, q0, nq, xv, S.err);
            tcol_product<NR, GH / 64>(L, lda, ne + rb, nr, 0, ne, xv, nq, cs);
            __syncthreads();
            if (tid < nb) {
#pragma unroll
                for (int q = 0; q < NR; ++q)
                    if (q < nq) acc[q] -= cs[q * SB + tid];
            }
            __syncthreads();
        }
        load_diag_block(dl, L, lda, j0, n1, n2);
    } else if (r > 0) {
        if (tid == 0) wait_at_least(S.bgc + fb + C, nch, S.err);
        __syncthreads();
        if (tid < nb) {
            for (int h = 0; h < nch; ++h) {
#pragma unroll
                for (int q = 0; q < NR; ++q)
                    if (q < nq) acc[q] -= ld_co(Pp + (size_t)(q0 + q) * ptotal + Fr.pbase + ((size_t)C * nch + h) * SB + tid);
            }
        }
    }
    // later blocks of the same front, as they are published (descending)
    for (int Cp = nblk - 1; Cp > C; --Cp) {
        const int i0 = Cp * SB, nr = min(SB, ne - i0);
        // the tile is loaded before block Cp's flag is awaited (for Cp = C + 1 that wait ends the
        // previous chain step; only the products remain after it)
        double lv[4][2][8];
        tcol_load2(L, lda, i0, nr, j0, nb, lv);
        double xv[NR][2];
        __syncthreads();  // cs of the previous block consumed
        if (tid == 0) wait_at_least(S.bflag + fb + Cp, 1, S.err);
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int i = lane + 64 * u;
#pragma unroll
            for (int q = 0; q < NR; ++q)
                xv[q][u] = (i < nr && q < nq) ? ld_co(Xb + (size_t)(q0 + q) * wtotal + Fr.wbase + i0 + i) : 0.0;
        }
        tcol_compute2<NR>(lv, xv, nq, nb, cs);
        __syncthreads();
        if (tid < nb) {
#pragma unroll
            for (int q = 0; q < NR; ++q)
                if (q < nq) acc[q] -= cs[q * SB + tid];
        }
    }
    if (tid < nb) {
#pragma unroll
        for (int q = 0; q < NR; ++q) ws[q * SB + tid] = acc[q];
    }
    __syncthreads();
    bwd_block<NR>(dl, ws, part, n1, n2, nq);
    if (tid < nb) {
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            if (q < nq) {
                const double v = ws[q * SB + tid];
                if (nblk > 1) st_co(Xb + (size_t)(q0 + q) * wtotal + Fr.wbase + j0 + tid, v);
                if ((unsigned)st < (unsigned)n) st_co(x + (size_t)(q0 + q) * n + st, v);
                else atomicOr(S.err, 2);
            }
        }
    }
    stores_done_barrier();
    if (tid == 0) {
        if (nblk > 1) publish_set(S.bflag + fb + C);
        publish_add(S.bdone + (size_t)grp * P.nfront + li);
    }
}

}  // namespace

void NumericGPU::prepare_solve() {
    const Prepared &P = *prep_;
    const int ns = (int)P.sns.size();
    const i64 n = P.n;
    // fronts taking part (ne > 0), in level order
    std::vector<int> flist, li_of(ns, -1);
    for (const auto &lv : P.levels)
        for (int s : lv)
            if (ne_[s] > 0) {
                li_of[s] = (int)flist.size();
                flist.push_back(s);
            }
    const int nf = (int)flist.size();
    // per-front bases
    std::vector<long long> wb(nf), yb(nf), pb(nf);
    std::vector<int> fbase(nf), nrow(nf), nchk(nf);
    std::vector<long long> ybase_sn(ns, 0);
    long long wt = 0, yt = 0, pt = 0;
    int nflag = 0;
    for (int li = 0; li < nf; ++li) {
        const int s = flist[li];
        const int ne = ne_[s], r = m_[s] - ne, nblk = (ne + SB - 1) / SB;
        wb[li] = wt;
        wt += ne;
        yb[li] = yt;
        ybase_sn[s] = yt;
        yt += r;
        fbase[li] = nflag;
        nflag += nblk;
        nrow[li] = (r + SB - 1) / SB;
        nchk[li] = (r + GH - 1) / GH;
        pb[li] = pt;
        pt += (long long)nblk * nchk[li] * SB;
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
            long long p = ybase_sn[s];
            for (int g : delayed_[s]) ent[pos[g]++] = p++;
            for (int g : P.sns[s].pattern) ent[pos[g]++] = p++;
        }
    }
    // unresolved columns (delays left at the roots)
    std::vector<int> unres;
    for (int s = 0; s < ns; ++s)
        if (P.sns[s].parent < 0) unres.insert(unres.end(), delayed_[s].begin(), delayed_[s].end());
    // effective children / parent: nearest descendants (ancestor) with ne > 0
    auto fused = [&](int li) {
        const int s = flist[li];
        return ne_[s] <= SB && m_[s] - ne_[s] <= FUSE_R;
    };
    std::vector<int> ech_ptr(nf + 1, 0), echf, epar(nf, -1);
    for (int li = 0; li < nf; ++li) {
        std::vector<int> stack(P.children[flist[li]].rbegin(), P.children[flist[li]].rend());
        while (!stack.empty()) {
            const int c = stack.back();
            stack.pop_back();
            if (ne_[c] > 0) {
                echf.push_back(li_of[c]);
                epar[li_of[c]] = li;
            } else {
                stack.insert(stack.end(), P.children[c].rbegin(), P.children[c].rend());
            }
        }
        ech_ptr[li + 1] = (int)echf.size();
    }
    // forward done target of a front: its rows items, or 1 for a fused front
    auto ndone = [&](int li) { return fused(li) ? 1 : nrow[li]; };
    std::vector<int2> ech(echf.size());
    for (size_t e = 0; e < echf.size(); ++e) ech[e] = make_int2(echf[e], ndone(echf[e]));
    std::vector<SolveFront> sf(nf);
    for (int li = 0; li < nf; ++li) {
        const int s = flist[li];
        const FrontDev &h = h_fr_[s];
        SolveFront &f = sf[li];
        f.F = h.Lf;
        f.sidx = h.sidx;
        f.dkind = h.dkind;
        f.ddiag = h.ddiag;
        f.doff = h.doff;
        f.wbase = wb[li];
        f.ybase = yb[li];
        f.pbase = pb[li];
        f.m = m_[s];
        f.ne = ne_[s];
        f.fbase = fbase[li];
        f.ndone = ndone(li);
        f.nchk = nchk[li];
        f.ech0 = ech_ptr[li];
        f.ech1 = ech_ptr[li + 1];
        f.epar = epar[li];
        f.epar_nblk = epar[li] >= 0 ? (ne_[flist[epar[li]]] + SB - 1) / SB : 0;
    }
    // items: forward by levels bottom-up, backward top-down.  Within a level: forward, the
    // diagonal (and fused) items by block step across fronts, then the rows items; backward, the
    // fused and chunk items, then the diagonal items by step (last block first).
    std::vector<int4> fit, bit;
    std::vector<std::vector<int>> lvl_fronts(P.levels.size());
    for (size_t lv = 0; lv < P.levels.size(); ++lv)
        for (int s : P.levels[lv])
            if (ne_[s] > 0) lvl_fronts[lv].push_back(li_of[s]);
    auto nblk_of = [&](int li) { return (ne_[flist[li]] + SB - 1) / SB; };
    for (size_t lv = 0; lv < lvl_fronts.size(); ++lv) {
        const auto &fl = lvl_fronts[lv];
        int maxblk = 0;
        for (int li : fl) maxblk = std::max(maxblk, nblk_of(li));
        for (int c = 0; c < maxblk; ++c)
            for (int li : fl)
                if (c < nblk_of(li)) fit.push_back(make_int4(li, fused(li) ? IT_FSMALL : IT_FDIAG, c, 0));
        for (int li : fl)
            if (!fused(li))
                for (int g = 0; g < nrow[li]; ++g) fit.push_back(make_int4(li, IT_FROWS, g, 0));
    }
    for (int lv = (int)lvl_fronts.size() - 1; lv >= 0; --lv) {
        const auto &fl = lvl_fronts[lv];
        int maxblk = 0;
        for (int li : fl) {
            if (fused(li)) {
                bit.push_back(make_int4(li, IT_BSMALL, 0, 0));
                continue;
            }
            maxblk = std::max(maxblk, nblk_of(li));
            for (int C = nblk_of(li) - 1; C >= 0; --C)
                for (int h = 0; h < nchk[li]; ++h) bit.push_back(make_int4(li, IT_BCHUNK, C * nchk[li] + h, 0));
        }
        for (int step = 0; step < maxblk; ++step)
            for (int li : fl)
                if (!fused(li) && step < nblk_of(li)) bit.push_back(make_int4(li, IT_BDIAG, nblk_of(li) - 1 - step, 0));
    }
    // diagonal triangles (64 columns each) of every front, inverted after each factorization;
    // grouped by order class (<= 16, <= 32, <= 64), one launch per class
    std::vector<int4> tcls[3], tris;
    for (int li = 0; li < nf; ++li) {
        const int ne = ne_[flist[li]];
        for (int jb = 0; jb < ne; jb += 64) {
            const int tn = std::min(64, ne - jb);
            tcls[tn <= 16 ? 0 : tn <= 32 ? 1 : 2].push_back(make_int4(li, jb, tn, 0));
        }
    }
    for (int c = 0; c < 3; ++c) {
        sm_.ntri[c] = (int)tcls[c].size();
        tris.insert(tris.end(), tcls[c].begin(), tcls[c].end());
    }
    // device metadata
    std::vector<char> &st = sol_meta_host_;
    st.clear();
    auto push = [&](const void *p, size_t bytes) {
        size_t off = (st.size() + 255) / 256 * 256;
        st.resize(off + std::max<size_t>(bytes, 1));
        if (bytes) std::memcpy(st.data() + off, p, bytes);
        return off;
    };
    const size_t o_ptr = push(cnt.data(), cnt.size() * sizeof(int));
    const size_t o_ent = push(ent.data(), ent.size() * sizeof(long long));
    const size_t o_sf = push(sf.data(), sf.size() * sizeof(SolveFront));
    const size_t o_ec = push(ech.data(), ech.size() * sizeof(int2));
    const size_t o_fit = push(fit.data(), fit.size() * sizeof(int4));
    const size_t o_bit = push(bit.data(), bit.size() * sizeof(int4));
    const size_t o_un = push(unres.data(), unres.size() * sizeof(int));
    const size_t o_tr = push(tris.data(), tris.size() * sizeof(int4));
    if (st.size() > sol_meta_bytes_) {
        if (d_sol_meta_) (void)hipFree(d_sol_meta_);
        d_sol_meta_ = nullptr;
        sol_meta_bytes_ = 0;
        RV_HIP(hipMalloc(&d_sol_meta_, st.size()));
        sol_meta_bytes_ = st.size();
    }
    RV_HIP(hipMemcpy(d_sol_meta_, st.data(), st.size(), hipMemcpyHostToDevice));
    char *b = (char *)d_sol_meta_;
    sm_.ptr = (const int *)(b + o_ptr);
    sm_.ent = (const long long *)(b + o_ent);
    sm_.fronts = (const void *)(b + o_sf);
    sm_.ech = (const int2 *)(b + o_ec);
    sm_.fitems = (const int4 *)(b + o_fit);
    sm_.bitems = (const int4 *)(b + o_bit);
    sm_.nfitems = (int)fit.size();
    sm_.nbitems = (int)bit.size();
    sm_.unres = (const int *)(b + o_un);
    sm_.tris = (const int4 *)(b + o_tr);
    sm_.ntris = (int)tris.size();
    inv_valid_ = false;
    sm_.nunres = (int)unres.size();
    sm_.nfront = nf;
    sm_.nflag = nflag;
    sm_.wtotal = std::max<long long>(wt, 1);
    sm_.ytotal = std::max<long long>(yt, 1);
    sm_.ptotal = std::max<long long>(pt, 1);
    sol_plan_gen_ = plan_gen_;
    solve_ready_ = true;
}

long long NumericGPU::solve_scratch_bytes(int nrhs) const {
    const int nr = solve_group_width(nrhs);
    const long long ngrp = nrhs > 0 ? (nrhs + nr - 1) / nr : 0;
    const long long nsync = 2 + 1 + ngrp * (3LL * sm_.nflag + 2LL * sm_.nfront);
    const long long doubles = ((long long)prep_->n + 2 * sm_.wtotal + sm_.ytotal + sm_.ptotal) * nrhs;
    return ((doubles * 8 + 255) / 256) * 256 + ((nsync * 4 + 255) / 256) * 256;
}

int NumericGPU::solve_group_width(int nrhs) {
    return nrhs >= 8 ? 8 : nrhs >= 4 ? 4 : nrhs >= 2 ? 2 : 1;
}

void NumericGPU::solve(double *d_rhs, i64 ldrhs, int nrhs, const double *d_scaling) {
    if (!solve_ready_) prepare_solve();
    const i64 n = prep_->n;
    if (n == 0 || nrhs <= 0) return;
    void *ws = sol_buf_.get((size_t)solve_scratch_bytes(nrhs));
    solve_with(d_rhs, ldrhs, nrhs, d_scaling, ws);
    int err = 0;
    RV_HIP(hipMemcpyAsync(&err, sol_err_, sizeof(int), hipMemcpyDeviceToHost, stream_));
    RV_HIP(hipStreamSynchronize(stream_));
    if (err) throw std::runtime_error(solve_error_text(err));
}

// The solve on a caller-provided workspace of solve_scratch_bytes(nrhs) bytes (device memory,
// 256-byte aligned), enqueued on the library stream; allocates nothing.  The error word (set
// when a dependency wait exceeded its limit) is left at sol_err_.
void NumericGPU::solve_with(double *d_rhs, i64 ldrhs, int nrhs, const double *d_scaling, void *work) {
    const Prepared &P = *prep_;
    const i64 n = P.n;
    const int NRW = solve_group_width(nrhs);
    const int ngrp = (nrhs + NRW - 1) / NRW;
    const long long wt = sm_.wtotal, yt = sm_.ytotal, pt = sm_.ptotal, nn = n;
    double *xp = (double *)work;
    double *W = xp + (size_t)n * nrhs;
    double *Xb = W + (size_t)wt * nrhs;
    double *Y = Xb + (size_t)wt * nrhs;
    double *Pp = Y + (size_t)yt * nrhs;
    const long long doubles = ((long long)n + 2 * wt + yt + pt) * nrhs;
    int *sync = (int *)((char *)work + ((doubles * 8 + 255) / 256) * 256);
    const long long nsync = 3 + (long long)ngrp * (3LL * sm_.nflag + 2LL * sm_.nfront);
    SyncDev S;
    S.tick = sync;
    S.err = sync + 2;
    S.fflag = sync + 3;
    S.bflag = S.fflag + (size_t)ngrp * sm_.nflag;
    S.bgc = S.bflag + (size_t)ngrp * sm_.nflag;
    S.fdone = S.bgc + (size_t)ngrp * sm_.nflag;
    S.bdone = S.fdone + (size_t)ngrp * sm_.nfront;
    sol_err_ = S.err;
    PlanDev D;
    D.sf = (const SolveFront *)sm_.fronts;
    D.ech = sm_.ech;
    D.ptr = sm_.ptr;
    D.ent = sm_.ent;
    D.nfront = sm_.nfront;
    D.nflag = sm_.nflag;
    if (!inv_valid_) {
        // the first solve after a factorization inverts the diagonal triangles
        const int4 *tp = sm_.tris;
        if (sm_.ntri[0] > 0)
            hipLaunchKernelGGL(k_diag_inverse<16>, dim3((unsigned)sm_.ntri[0]), dim3(64), 0, stream_,
                               (const SolveFront *)sm_.fronts, tp);
        tp += sm_.ntri[0];
        if (sm_.ntri[1] > 0)
            hipLaunchKernelGGL(k_diag_inverse<32>, dim3((unsigned)sm_.ntri[1]), dim3(64), 0, stream_,
                               (const SolveFront *)sm_.fronts, tp);
        tp += sm_.ntri[1];
        if (sm_.ntri[2] > 0)
            hipLaunchKernelGGL(k_diag_inverse<64>, dim3((unsigned)sm_.ntri[2]), dim3(64), 0, stream_,
                               (const SolveFront *)sm_.fronts, tp);
        inv_valid_ = true;
    }
    const long long nstart = std::max<long long>(n, nsync);
    hipLaunchKernelGGL(k_solve_start, dim3((unsigned)((nstart + 255) / 256)), dim3(256), 0, stream_, d_rhs,
                       (long long)ldrhs, xp, nn, nrhs, P.d_perm_fwd, d_scaling, sync, nsync);
    const int nfw = sm_.nfitems * ngrp, nbw = sm_.nbitems * ngrp;
#define RV_SWEEPS(NRV)                                                                                        \
    do {                                                                                                      \
        if (nfw > 0)                                                                                          \
            hipLaunchKernelGGL(k_fwd_sweep<NRV>, dim3((unsigned)nfw), dim3(256), 0, stream_, D, sm_.fitems,   \
                               nfw, ngrp, nrhs, xp, nn, W, wt, Y, yt, S);                                     \
        if (sm_.nunres > 0)                                                                                   \
            hipLaunchKernelGGL(k_pull_rows, dim3((sm_.nunres + 255) / 256, nrhs), dim3(256), 0, stream_,       \
                               sm_.unres, sm_.nunres, sm_.ptr, sm_.ent, (const double *)Y, yt, xp, nn);       \
        if (nbw > 0)                                                                                          \
            hipLaunchKernelGGL(k_bwd_sweep<NRV>, dim3((unsigned)nbw), dim3(256), 0, stream_, D, sm_.bitems,   \
                               nbw, ngrp, nrhs, xp, nn, (const double *)W, Xb, wt, Pp, pt, S);                \
    } while (0)
    switch (NRW) {
        case 8: RV_SWEEPS(8); break;
        case 4: RV_SWEEPS(4); break;
        case 2: RV_SWEEPS(2); break;
        default: RV_SWEEPS(1); break;
    }
#undef RV_SWEEPS
    const dim3 gp((unsigned)((n + 255) / 256), nrhs);
    hipLaunchKernelGGL(k_perm_out, gp, dim3(256), 0, stream_, d_rhs, (long long)ldrhs, (const double *)xp, nn,
                       P.d_perm_fwd, d_scaling);
}


// ===========================================================================
// Fixed workspace of the allocation-free solve entry point: one array per GPU, held until the
// process ends.  Host code never passes the array's name to a kernel (on the host it names a
// stand-in, not device memory): the device address comes from hipGetSymbolAddress, once per
// device, with that device current.
// ===========================================================================
namespace {
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)RV_SOLVE_NOALLOC_BYTES;
__device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[MAX_DEVICES] = {};
}  // namespace

size_t noalloc_scratch_bytes() { return NOALLOC_SCRATCH_BYTES; }

std::string solve_error_text(int err) {
    std::string m = "solve failed, results invalid:";
    if (err & 1) m += " a dependency wait exceeded its limit;";
    if (err & 2) m += " an index fell outside the system (solve plan and fronts disagree);";
    return m;
}

void *noalloc_scratch() {
    int dev = 0;
    RV_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) throw std::runtime_error("device index beyond the supported range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        RV_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw std::runtime_error("hipGetSymbolAddress returned no address for the solve workspace");
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

}  // namespace rv
