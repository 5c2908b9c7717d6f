// rv_host.h -- host-side structures shared by the symbolic analysis, the structure
// preparation (amalgamation, small-leaf classification, assembly maps) and the GPU driver.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace rv {

typedef int64_t i64;

struct ColRange { int lo, hi; };

// Supernode descriptor after the host symbolic analysis / amalgamation
// (mirrors SupernodeInfo, sparse/src/symmetric/numeric.rs:431-448).
struct Supernode {
    int col_begin = 0, col_end = 0;
    std::vector<int> pattern;        // sorted global permuted row indices > owned columns
    int parent = -1;                 // -1 = root
    std::vector<ColRange> owned;     // owned column ranges (non-contiguous after merges)
    bool in_small_leaf = false;
    int n_owned() const {
        int s = 0;
        for (const ColRange &r : owned) s += r.hi - r.lo;
        return s;
    }
};

// Symbolic analysis result: permutation + pre-amalgamation supernode partition
// (the data AptpSymbolic exposes to build_supernode_info, numeric.rs:2179-2236).
struct Symbolic {
    i64 n = 0;
    std::vector<int> perm_fwd, perm_inv;  // perm_fwd[new] = old
    std::vector<int> etree;               // parent column in permuted order, -1 root
    std::vector<i64> col_counts;          // nnz per column of L incl. diagonal
    std::vector<Supernode> sn;            // pre-amalgamation supernodes, postorder
    i64 predicted_nnz = 0;
};

// Ordering selector for analyze (mirrors OrderingStrategy, solver.rs:42-55; the
// nested-dissection variants stand in for METIS, which is not available here).
enum OrderingKind {
    ORDER_NATURAL = 0,
    ORDER_USER = 1,
    ORDER_ND_GRAPH = 2,
    ORDER_ND_GRID = 3,
};

// symbolic.cpp
bool analyze(i64 n, const i64 *col_ptr, const i64 *row_idx, int ordering, const int *user_perm,
             const int *grid_dims, Symbolic &out, std::string &err);
void nested_dissection_graph(i64 n, const i64 *col_ptr, const i64 *row_idx, std::vector<int> &perm);
void nested_dissection_grid(int nx, int ny, int nz, std::vector<int> &perm);

// matching.cpp: MC64 matching + symmetric scaling (matching.rs:121-248) and the matching-based
// ordering (metis.rs:118-190, the library's nested dissection in place of METIS)
struct Mc64Result {
    std::vector<int> matching;     // row i matched to column matching[i]
    std::vector<double> scaling;   // A_scaled = S A S
    std::vector<char> is_matched;  // row i has a matching edge
    i64 matched = 0;
};
struct CycleSplit {
    std::vector<int> partner;  // paired index, -1 single, -2 unmatched
    std::vector<int> node_of;  // condensed vertex of each index
    std::vector<int> first;    // first member of each condensed vertex
    i64 singletons = 0, pairs = 0;
};
struct MatchOrderResult {
    std::vector<int> order;  // order[new] = old
    std::vector<double> scaling;
    std::vector<int> matching;
    std::vector<char> is_matched;
    i64 matched = 0, condensed_dim = 0, singletons = 0, two_cycles = 0;
};
bool mc64_matching(i64 n, const i64 *col_ptr, const i64 *row_idx, const double *values, Mc64Result &out,
                   std::string &err);
CycleSplit split_cycles(const std::vector<int> &fwd, const std::vector<char> &is_matched);
bool match_order(i64 n, const i64 *col_ptr, const i64 *row_idx, const double *values,
                 MatchOrderResult &out, std::string &err);

// structure.cpp: amalgamation (amalgamation.rs:115-231), children map, small-leaf
// classification (numeric.rs:2278-2369)
std::vector<Supernode> amalgamate(std::vector<Supernode> sns, int nemin);
std::vector<std::vector<int>> children_map(const std::vector<Supernode> &sns);
int classify_small_leaf(std::vector<Supernode> &sns, const std::vector<std::vector<int>> &ch,
                        int threshold, int *n_nodes);

}  // namespace rv
