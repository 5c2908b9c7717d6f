// structure.cpp -- supernode amalgamation, assembly-tree children and small-leaf subtree
// classification, run on the host at the start of every factorization exactly as
// AptpNumeric::factor does (sparse/src/symmetric/numeric.rs:862-871).
#include <algorithm>

#include "rv_host.h"

namespace rv {

// do_merge predicate (sparse/src/symmetric/amalgamation.rs:38-54): merge when the parent
// is a single column whose count matches the child's minus one (no fill), or when both
// have fewer than nemin eliminated columns.
static bool merge_rule(int parent_ncols, int parent_cc, int child_ncols, int child_cc, int nemin) {
    if (parent_ncols == 1 && parent_cc == std::max(child_cc - 1, 0)) return true;
    return parent_ncols < nemin && child_ncols < nemin;
}

// Sorted merge of two sorted index lists dropping indices in [lo, hi)
// (amalgamation.rs:57-97).
static std::vector<int> merge_patterns(const std::vector<int> &a, const std::vector<int> &b,
                                       int lo, int hi) {
    std::vector<int> out;
    out.reserve(a.size() + b.size());
    size_t x = 0, y = 0;
    while (x < a.size() || y < b.size()) {
        int v;
        if (y == b.size() || (x < a.size() && a[x] < b[y])) v = a[x++];
        else if (x == a.size() || b[y] < a[x]) v = b[y++];
        else { v = a[x]; ++x; ++y; }
        if (v < lo || v >= hi) out.push_back(v);
    }
    return out;
}

// amalgamate (amalgamation.rs:115-231).  Parents are visited in increasing (postorder)
// index; each parent tests its current children in order, and a merged child's own
// children are re-parented to the parent without being re-tested in the same pass.
std::vector<Supernode> amalgamate(std::vector<Supernode> sns, int nemin) {
    const int n = (int)sns.size();
    if (n <= 1) return sns;
    std::vector<char> gone(n, 0);
    std::vector<int> ncols(n);
    for (int s = 0; s < n; ++s) ncols[s] = sns[s].col_end - sns[s].col_begin;
    std::vector<std::vector<int>> kids(n);
    for (int s = 0; s < n; ++s)
        if (sns[s].parent >= 0) kids[sns[s].parent].push_back(s);
    for (int p = 0; p < n; ++p) {
        if (gone[p]) continue;
        std::vector<int> tested;
        tested.swap(kids[p]);
        for (int c : tested) {
            if (gone[c]) continue;
            const int pcc = ncols[p] + (int)sns[p].pattern.size();
            const int ccc = ncols[c] + (int)sns[c].pattern.size();
            if (!merge_rule(ncols[p], pcc, ncols[c], ccc, nemin)) continue;
            Supernode &P = sns[p];
            Supernode &C = sns[c];
            const int lo = std::min(P.col_begin, C.col_begin);
            const int hi = std::max(P.col_end, C.col_end);
            P.pattern = merge_patterns(P.pattern, C.pattern, lo, hi);
            C.pattern.clear();
            P.owned.insert(P.owned.end(), C.owned.begin(), C.owned.end());
            C.owned.clear();
            P.col_begin = lo;
            P.col_end = hi;
            ncols[p] += ncols[c];
            for (int g : kids[c])
                if (!gone[g]) sns[g].parent = p;
            kids[p].insert(kids[p].end(), kids[c].begin(), kids[c].end());
            kids[c].clear();
            gone[c] = 1;
        }
        for (int c : tested)
            if (!gone[c]) kids[p].push_back(c);
    }
    std::vector<int> renum(n, -1);
    int cnt = 0;
    for (int s = 0; s < n; ++s)
        if (!gone[s]) renum[s] = cnt++;
    std::vector<Supernode> out;
    out.reserve(cnt);
    for (int s = 0; s < n; ++s) {
        if (gone[s]) continue;
        out.push_back(std::move(sns[s]));
        Supernode &x = out.back();
        if (x.parent >= 0) x.parent = renum[x.parent];
    }
    return out;
}

std::vector<std::vector<int>> children_map(const std::vector<Supernode> &sns) {
    std::vector<std::vector<int>> ch(sns.size());
    for (size_t s = 0; s < sns.size(); ++s)
        if (sns[s].parent >= 0) ch[sns[s].parent].push_back((int)s);
    return ch;
}

// classify_small_leaf_subtrees (numeric.rs:2278-2369).  A supernode is in a small-leaf
// subtree when its front (owned + pattern) is below the threshold and all its children are;
// subtrees of a single node are dropped.  Returns the subtree count.
int classify_small_leaf(std::vector<Supernode> &sns, const std::vector<std::vector<int>> &ch,
                        int threshold, int *n_nodes) {
    *n_nodes = 0;
    if (threshold == 0) return 0;
    const int ns = (int)sns.size();
    for (int s = 0; s < ns; ++s) {
        const int front = sns[s].n_owned() + (int)sns[s].pattern.size();
        bool small = front < threshold;
        if (small)
            for (int c : ch[s]) small = small && sns[c].in_small_leaf;
        sns[s].in_small_leaf = small;
    }
    int count = 0;
    for (int s = 0; s < ns; ++s) {
        if (!sns[s].in_small_leaf) continue;
        const int p = sns[s].parent;
        if (p >= 0 && sns[p].in_small_leaf) continue;  // not a subtree root
        int size = 0;
        std::vector<int> stack{s};
        while (!stack.empty()) {
            const int v = stack.back();
            stack.pop_back();
            ++size;
            for (int c : ch[v])
                if (sns[c].in_small_leaf) stack.push_back(c);
        }
        if (size < 2) {
            sns[s].in_small_leaf = false;
            continue;
        }
        ++count;
        *n_nodes += size;
    }
    return count;
}

}  // namespace rv
