// rv_numeric.h -- host driver of the GPU numeric phase (factorization + solve).
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <array>
#include <vector>

#include "rv_gpu.h"
#include "rv_host.h"

namespace rv {

struct ChainSync;

struct HipError : std::runtime_error {
    hipError_t code;
    HipError(hipError_t e, const char *what, const char *file, int line)
        : std::runtime_error(std::string("HIP error ") + hipGetErrorString(e) + " in " + what +
                             " at " + file + ":" + std::to_string(line)),
          code(e) {}
};

// Mirrors AptpOptions (sparse/src/symmetric/factor.rs:86-130) plus the update precision.
struct FactorOptions {
    double threshold = 0.01;
    double small = 1e-20;
    int outer_block_size = 256;
    int inner_block_size = 32;
    int failed_pivot_tpp = 1;
    int nemin = 32;
    int small_leaf_threshold = 256;
    int update_precision = 0;  // 0 FP64 (MFMA), 1 FP32, 2 BF16, 3 FP64 in the reference's order
};

// Chunked device arena: allocations never move, everything is freed together.
// Bump allocator over device chunks of at least 256 MiB; chunks are kept until release().
class DevArena {
public:
    struct Mark { size_t chunk = 0, used = 0; };
    ~DevArena() { release(); }
    void *alloc(size_t bytes, size_t align = 256);
    void reset_keep();  // reuse every chunk from the start (no device frees)
    Mark mark() const;
    void reset_to(const Mark &m);  // free everything allocated after the mark (no device frees)
    void release();
    size_t bytes_reserved() const { return reserved_; }

private:
    struct Chunk { char *base; size_t size, used; };
    std::vector<Chunk> chunks_;
    size_t cur_ = 0;
    size_t reserved_ = 0;
};

// Working fronts of the levels: a block is alive from its level until `death` (the level whose
// assembly reads its contribution block and delayed columns) and then handed out again.  Each
// block takes the lowest offset of a chunk not covered by a block alive at its level; chunks never
// move and are freed together.
class FrontPool {
public:
    ~FrontPool() { release(); }
    // unbounded address range without device memory (placement estimates)
    void set_virtual() { virtual_ = true; }
    void reserve(size_t bytes);  // first chunk, if there is none
    void drop_from(int lv);      // forget the blocks of levels >= lv (planned again)
    int place(int lv, int death, size_t bytes);  // block id
    char *ptr(int id) const { return chunks_[blocks_[id].chunk].base + blocks_[id].off; }
    // a block of a later level was placed over this one after it died
    bool clobbered(int id) const { return blocks_[id].clobbered; }
    size_t top() const { return top_; }  // highest end offset placed
    size_t bytes_reserved() const { return reserved_; }
    void release();

private:
    struct Chunk { char *base; size_t size; };
    struct Block { int chunk; size_t off, size; int lv, death; bool clobbered; };
    std::vector<Chunk> chunks_;
    std::vector<Block> blocks_;
    bool virtual_ = false;
    size_t top_ = 0, reserved_ = 0;
};

// Device-resident scratch that grows on demand.
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    void *get(size_t bytes);
    ~DevBuf();
};

struct FactorStats {
    long long total_1x1 = 0, total_2x2 = 0, total_delayed = 0, zero_pivots = 0;
    long long max_front_size = 0, sn_before = 0, sn_after = 0, merges = 0;
    long long small_leaf_subtrees = 0, small_leaf_nodes = 0;
};

struct PhaseTimes {  // milliseconds, measured with HIP events on the library stream
    double factor_ms = 0, solve_ms = 0;
    double schur_ms = 0;          // contribution + outer Schur GEMMs
    double schur_flops = 0;       // algorithmic flops of those GEMMs
    int schur_launches = 0;
    // contribution GEMMs of levels whose largest contribution block has order >= 2048
    double big_ms = 0, big_flops = 0;
    int big_launches = 0;
    double assembly_ms = 0;
    double assembly_bytes = 0;
};

// Prepared structure: amalgamated supernodes, levels and assembly maps for one pattern.
struct Prepared {
    i64 n = 0;
    std::vector<int> perm_fwd, perm_inv;
    std::vector<Supernode> sns;
    std::vector<std::vector<int>> children;
    std::vector<std::vector<int>> levels;   // level 0 = leaves
    std::vector<int> level_of;
    std::vector<i64> amap_off;               // per supernode range into amap
    std::vector<AmapEntry> amap;
    std::vector<std::vector<unsigned>> ea0;  // per supernode: concatenated child maps (no delays)
    std::vector<std::vector<int>> ea0_off;   // per supernode: offset of each child's map
    long long sn_before = 0, merges = 0, small_leaf_subtrees = 0, small_leaf_nodes = 0;
    i64 nnz = 0;
    AmapEntry *d_amap = nullptr;
    int *d_perm_fwd = nullptr;
    void build(const Symbolic &sym, const i64 *col_ptr, const i64 *row_idx, int nemin,
               int small_leaf_threshold);
    ~Prepared();
};

class NumericGPU {
public:
    NumericGPU(const Prepared *prep, const FactorOptions &opts, hipStream_t stream);
    ~NumericGPU();
    // values: full-symmetric CSC values on the device; scaling: device, elimination order
    void factor(const double *d_val, const double *d_scaling);
    // rhs: device, original coordinates, n x nrhs column-major (ld = ldrhs)
    void solve(double *d_rhs, i64 ldrhs, int nrhs, const double *d_scaling);
    void inertia(long long out[3]) const;
    const FactorStats &stats() const { return stats_; }
    const PhaseTimes &times() const { return times_; }
    // update arithmetic per supernode (RV_PREC_* codes; NULL = the options' update_precision for
    // all); takes effect at the next factorization
    void set_front_precision(const signed char *prec);
    const std::vector<signed char> &front_precision() const { return fprec_; }
    // 16 classifier features per supernode (row-major) from the last factorization
    void front_features(float *out) const;
    int num_supernodes() const { return (int)prep_->sns.size(); }
    // per-supernode statistics: m, k, ne, n1, n2, parent, in_small_leaf
    void supernode_info(int s, long long out[8], double *max_l) const;
    // copies of one front's factors (host): see rv_export_front in include/rivrs_ldlt.h
    void export_front(int s, double *l11, double *l21, signed char *dkind, double *ddiag,
                      double *doff, int *col_indices, int *row_indices, int *local_perm) const;
    bool timing = false;
    // refactorizations whose levels all run on unchanged plans replay one captured hipGraph of
    // the whole level sequence (kernels and copies only)
    bool graph_mode = false;
    long long graph_launches() const { return graph_launches_; }
    long long chain_launches() const { return chain_launches_; }
    long long lookahead_launches() const { return la_launches_; }
    size_t device_bytes() const {
        return arena_.bytes_reserved() + scratch_.bytes_reserved() + pool_.bytes_reserved();
    }
    // working fronts on fresh memory that no later level reuses (RV_KEEP_FRONTS=1, read when the
    // handle is made): every contribution block stays readable after the factorization
    bool keep_fronts() const { return keep_fronts_; }
    // the working front of s (its contribution block) still holds this factorization's data
    bool front_held(int s) const;

private:
    // Everything one level's launches need, built on the host from the children's results
    // (eliminated counts, delayed columns) it records; kept across refactorizations and
    // reused while those results repeat.
    struct LevelPlan {
        bool valid = false;
        DevArena::Mark mark;  // arena position before the level's allocations
        std::vector<int> child_ne;
        std::vector<std::vector<int>> child_del;
        double *Fbase = nullptr;
        size_t totF = 0, totM = 0;
        int *ibase = nullptr;
        std::vector<char> stage;  // host copy of the staging blob
        char *dstage = nullptr;   // device copy
        bool uploaded = false;
        char *dres = nullptr;     // k_collect output (ints, max|L|, update flops, delayed columns)
        size_t res_bytes = 0, res_del_off = 0;
        char *hres = nullptr;     // pinned copy of dres for the current run
        AptpState *dS = nullptr;
        char *mark_p = nullptr;
        size_t mark_bytes = 0;
        size_t o_ids = 0, o_lfr = 0, o_rows = 0, o_sch = 0, o_tpp = 0, o_aptp = 0, o_k2 = 0, o_k3 = 0,
               o_wp0 = 0, o_wp1 = 0, o_wp1b = 0, o_f0 = 0, o_f1 = 0, o_doffs = 0, o_hs = 0;
        // copies of the APTP states for outer updates run beside the next outer block (by
        // outer-block parity, as the two W buffers o_wp1 / o_wp1b)
        AptpState *dSnap[2] = {nullptr, nullptr};
        // contribution tiles per update precision (RV_PREC_* index)
        std::array<size_t, 4> o_ct{}, nct{};
        // outer-update tiles per outer block: [0] reference order (k_outer_exact), [1] FP32,
        // [2] BF16 and [3] FP64 rank-nb MFMA updates; the first na_ot of them cover the next
        // outer block's columns, the rest the columns after it
        std::vector<std::array<size_t, 4>> o_ot, n_ot, na_ot;
        // the next outer block's columns of FP32 / FP64-MFMA outer updates in 64 x 64 tiles (one
        // launch; the rest of the columns in 128 x 128 tiles, another)
        bool ota64 = false;
        std::vector<size_t> o_ea, n_ea;
        int nf = 0, na = 0, max_outer = 0;
        unsigned prec_gen = 0;
        size_t nsch = 0, ntpp = 0, nk2 = 0, nk3 = 0, nf0 = 0, nf1 = 0;
        size_t o_zl = 0, nzl = 0;  // lower-triangle zeroing items (front, first column)
        size_t o_kf = 0, nkf = 0;  // factor-column copy items (front, chunk)
        std::vector<int> ninner;
        // persistent outer-block launches (k_aptp_outer): panel rows per workgroup (0: the
        // per-inner-block launches), work items and count per outer block, sync blocks per front
        int chain_ch = 0;
        std::vector<size_t> o_chain, n_chain;
        ChainSync *dC = nullptr;
        std::vector<long long> doffs;
        std::vector<std::pair<hipEvent_t, hipEvent_t>> sev;
        std::vector<char> sev_contrib;  // per timed launch: contribution GEMM (mode 0)
    };
    std::vector<LevelPlan> plans_;
    std::vector<signed char> fprec_;  // update precision per supernode
    unsigned plan_gen_ = 0;           // bumped by every level plan built
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t gexec_ = nullptr;
    unsigned g_gen_ = 0;
    const double *g_val_ = nullptr, *g_scal_ = nullptr;
    long long graph_launches_ = 0;
    long long chain_launches_ = 0;
    void drop_graph();
    unsigned prec_gen_ = 0;           // bumped when fprec_ changes
    bool belief_valid_ = false;  // ne_/delayed_ hold the results of a completed factorization
    // persistent outer-block launches: enabled (RV_APTP_CHAIN=0 turns them off), a launch whose
    // wait ran out (the factorization is repeated without them), resident workgroups per
    // launch for 64- and 128-row panels (-1 until queried)
    bool chain_on_ = true;
    bool chain_err_ = false;
    // outer look-ahead: the trailing update of an outer block that had no failed columns in the
    // previous factorization runs on stream2_ beside the next outer block (RV_OUTER_LOOKAHEAD=0
    // turns it off); an unplanned failure repeats the factorization without it
    bool la_on_ = true;
    // 64 x 64 tiles for the next outer block's columns of MFMA outer updates (RV_OUTER_TILE64=0
    // turns them off: 128 x 128 tiles throughout, the same sums)
    bool outer_t64_ = true;
    bool la_err_ = false;
    bool la_suspend_ = false;
    bool capturing_ = false;
    hipStream_t stream2_ = nullptr;
    hipEvent_t ev_fork_ = nullptr, ev_join_ = nullptr;
    void ensure_stream2();  // stream2_ and its events (default priority), made on first use
    std::vector<unsigned long long> obfail_;  // per supernode: outer blocks with failed columns
    bool obfail_valid_ = false;
    long long la_launches_ = 0;
    int chain_cap_[2] = {-1, -1};
    size_t chain_capacity(int ch);
    bool plan_matches(int lv) const;
    void plan_level(int lv, bool reset_arena);
    void enqueue_level(int lv, const double *d_val, const double *d_scaling);
    void read_level(int lv);
    // pinned host blocks for staging uploads and results of the levels in flight
    std::vector<std::pair<char *, size_t>> pin_blocks_;
    size_t pin_cur_ = 0, pin_used_ = 0;
    char *pin_take(size_t bytes);
    void pin_reset();
    const Prepared *prep_;
    FactorOptions o_;
    hipStream_t stream_;
    DevArena arena_;    // per-level memory that outlives the level: fronts, indices, D, plans
    DevArena scratch_;  // per-level working memory, shared by all levels
    FrontPool pool_;    // working fronts of non-root supernodes (not with keep_fronts_)
    bool keep_fronts_ = false;
    std::vector<int> pool_block_;  // per supernode: its pool block (-1: front in arena_)
    void pool_groups(int lv, const std::vector<int> &m,
                     std::vector<std::pair<std::vector<int>, int>> *groups) const;
    size_t pool_estimate() const;
    FrontDev *d_fr_ = nullptr;
    std::vector<FrontDev> h_fr_;
    std::vector<int> ne_, m_, k_, n1_, n2_, zp_;
    std::vector<std::array<long long, 3>> inert_;  // inertia of each front's D
    std::vector<double> maxl_;
    std::vector<std::vector<int>> delayed_;  // global indices of delayed columns per front
    std::vector<std::vector<int>> rows_;     // host copy of each front's row list
    FactorStats stats_;
    PhaseTimes times_;
    // solve structures (built once per factorization)
    void prepare_solve();
    bool solve_ready_ = false;
    struct SolvePtrs {
        const int *ptr = nullptr;          // pull lists: Y entries per global index (CSR)
        const long long *ent = nullptr;
        const void *fronts = nullptr;      // per-front solve records (solve.hip)
        const int2 *ech = nullptr;         // effective children with their done targets
        const int4 *fitems = nullptr, *bitems = nullptr;  // forward / backward work items
        int nfitems = 0, nbitems = 0;
        const int *unres = nullptr;
        int nunres = 0;
        int nfront = 0, nflag = 0;
        long long wtotal = 1, ytotal = 1, ptotal = 1;
        const int4 *tris = nullptr;        // diagonal triangles inverted for the sweeps
        int ntris = 0;
        int ntri[3] = {0, 0, 0};           // of order <= 16, <= 32, <= 64 (in this order in tris)
    } sm_;
    bool inv_valid_ = false;  // the diagonal triangles' inverses match the current factorization
    DevBuf sol_buf_;
    int *sol_err_ = nullptr;  // error word of the last enqueued solve (inside its workspace)
    unsigned sol_plan_gen_ = ~0u;  // plan_gen_ when the solve plan (front addresses) was built
    std::vector<char> sol_meta_host_;
    void *d_sol_meta_ = nullptr;
    size_t sol_meta_bytes_ = 0;

public:
    // device workspace of a solve of nrhs columns, in bytes (valid once factored)
    long long solve_scratch_bytes(int nrhs) const;
    // right-hand sides per work group of the sweeps (every load of L serves the group)
    static int solve_group_width(int nrhs);
    // the solve on a caller-provided device workspace of solve_scratch_bytes(nrhs) bytes
    // (256-byte aligned): enqueued on the library stream, allocates nothing; the caller
    // synchronizes and then reads the error word (nonzero: see solve_error_text)
    void solve_with(double *d_rhs, i64 ldrhs, int nrhs, const double *d_scaling, void *work);
    const int *solve_error_word() const { return sol_err_; }
    bool solve_ready() const { return solve_ready_; }
    hipStream_t stream() const { return stream_; }
    // raw front after factorization (host copies): F[m*m], perm[m], rows[m]
    void export_raw(int s, double *F, int *perm, int *rows) const;
    int front_m(int s) const { return m_[s]; }
    int front_k(int s) const { return k_[s]; }
    int front_ne(int s) const { return ne_[s]; }
    void front_d(int s, signed char *kind, double *diag, double *off) const;
    int export_contribution(int s, double *data, int *rows, int *num_delayed) const;
};


// Fixed device workspace of the allocation-free solve (solve.hip): its size, and its address on
// the current device (looked up once per device; allocates nothing).
size_t noalloc_scratch_bytes();
void *noalloc_scratch();
// message for a nonzero solve error word (bit 0: a wait exceeded its limit, bit 1: an index
// outside the system)
std::string solve_error_text(int err);

// FP64 iterative refinement around a factorization (refine.hip): d_b, d_x are n doubles in
// original coordinates on the device; A is the full-symmetric CSC (d_cp, d_ri, d_val).
struct RefineResult {
    int iters = 0;                // corrections applied
    double backward_error = 0.0;  // of the returned x
};
RefineResult solve_refine(NumericGPU &gpu, hipStream_t st, i64 n, i64 nnz, const i64 *d_cp, const i64 *d_ri,
                          const double *d_val, const double *d_scal, const double *d_b, double *d_x,
                          int max_iter, double tol, DevBuf &work);

}  // namespace rv
