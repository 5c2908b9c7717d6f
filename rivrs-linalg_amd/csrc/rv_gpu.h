// rv_gpu.h -- device-side data layout of the numeric phase.
//
// HBM layout: every front of the assembly tree owns an m x m column-major FP64 block
// (lower triangle used) in a chunked arena; the factor L11/D/L21 stays in place after the
// factorization and the contribution block of a front is its trailing (m-ne) x (m-ne)
// sub-block, which the parent's extend-add reads directly (no extraction copies).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace rv {

// Inner block size limit of the device kernels (complete pivoting block lives in LDS).
constexpr int MAX_IB = 32;
constexpr int MAX_OB_IB = 64;  // inner blocks per outer block recorded for the deferred exact update

// Front data is reached through pointers held in structs, which the compiler cannot place in
// an address space: plain dereferences become flat accesses, and those count against the LDS
// wait counter too, so every wait for LDS data would also wait for the outstanding global
// loads.  Device code addresses front data through these global-memory views.
typedef __attribute__((address_space(1))) double gdouble;
typedef __attribute__((address_space(1))) int gint;
typedef __attribute__((address_space(1))) int8_t gi8;
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T *gbl(T *p) {
    return (__attribute__((address_space(1))) T *)p;
}

// Per-front descriptor, one entry per supernode, resident on the device.
struct FrontDev {
    double *F;        // m x m, lda = m (working front)
    double *Lf;       // eliminated columns kept for the solve: m x k, lda = m (F for a root)
    int *rows;        // global permuted index of each local row (length m)
    int *perm;        // APTP column order (length m): local position -> original local index
    int *sidx;        // after factorization: global index of local row i (rows[perm[i]] for i < k)
    int8_t *dkind;    // D pivot kind per eliminated column: 1 = 1x1, 2 = 2x2 (length k)
    double *ddiag;    // D diagonal (length k)
    double *doff;     // D off-diagonal at the first column of a 2x2 (length k)
    int m, k;         // front order, fully-summed count
    int n_owned;      // owned columns (delayed-in columns follow them)
    int ndelay_in;    // delayed columns received from children
    int ne;           // eliminated columns (result)
    int n1, n2;       // 1x1 / 2x2 pivot counts
    int flags;        // FRONT_TPP: whole front by threshold partial pivoting
    unsigned long long max_l_bits;  // max |L| as the bits of a non-negative double
    double uflops;    // algorithmic flops of the outer-block Schur updates
    unsigned long long obfail;  // bit ob (capped at 63): outer block ob ended with failed columns
};

constexpr int FRONT_TPP = 1;
constexpr int FRONT_ALLFS = 2;  // FP64 front whose inner blocks update every fully-summed column
constexpr int FRONT_W32 = 4;    // FP32 front: k_form_ld writes W and a copy of L21 as FP32 pairs

// APTP progress of one front through its outer (nb) and inner (ib) blocks.
struct AptpState {
    int kpos;         // next column to eliminate (absolute)
    int end_pos;      // end of the live region of the current outer block (absolute)
    int ob_start;     // first column of the current outer block
    int ob_cols;      // columns of the current outer block
    int remaining;    // fully-summed columns not yet assigned to an outer block
    int gne;          // columns eliminated by completed outer blocks
    int bs;           // current inner block size
    int block_nelim;  // columns eliminated by complete pivoting in the inner block
    int fail_min;     // first failing pivot of the a-posteriori check (atomicMin target;
                      // INT_MAX between inner blocks)
    int nib;          // inner blocks finished in the current outer block
    int bperm[MAX_IB];
    int8_t bkind[MAX_IB];
    double bdiag[MAX_IB];
    double boff[MAX_IB];
    int8_t ib_ne[MAX_OB_IB];  // columns eliminated by each inner block of the outer block
    // look-ahead: the next inner block factored by the update kernel of a block without
    // failures (nx_*), handed over by k_block_finish; pref = the current block was factored so
    int pref, nx_valid, nx_bs, nx_bn;
    int nx_bperm[MAX_IB];
    int8_t nx_bkind[MAX_IB];
    double nx_bdiag[MAX_IB];
    double nx_boff[MAX_IB];
    double *blkcopy;  // the next inner block's lower ib x ib diagonal block (column-major,
                      // leading dimension MAX_IB), copied before k_block_apply writes the front
    double *backup;   // (m - kpos) x ib backup of the block column, lda = m
    double *swapbuf;  // 2*ib lines of length m for symmetric line permutations
    int *mark;        // length m, zero outside a permutation: 1 + index of a touched line
};

// Work item of a tiled launch: front index plus tile coordinates.
struct TileItem {
    int front;
    int ti, tj;
    int aux;
};

// Extend-add item: one lower tile of one child's contribution block.
struct EaItem {
    int child;        // child front index
    int parent;       // parent front index
    int ti, tj;       // 64-row / 64-col tile of the contribution block
    const unsigned *map;  // contribution row -> parent local row
    int skip_zero;    // reference extend_add (g2l path) skips exact zeros
    int cne;          // child's eliminated count the map was built for
};

// Original-entry scatter entry (amap tuple, numeric.rs:225-408): source index into the CSC
// values, destination row/col in the front assuming no delayed columns, scaling indices.
struct AmapEntry {
    unsigned src, row, col, sr, sc;
};

}  // namespace rv

#define RV_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) throw rv::HipError(e_, #call, __FILE__, __LINE__);      \
    } while (0)
