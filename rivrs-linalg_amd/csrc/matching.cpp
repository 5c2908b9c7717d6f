// MC64 maximum-product matching with symmetric scaling, and the matching-based ordering
// (cycle condensation + nested dissection of the condensed graph + expansion).
//
// Follows sparse/src/ordering/matching.rs (mc64_matching, matching.rs:121-248): the published
// Duff & Koster (2001) algorithm MPD — logarithmic costs c_ij = log(max_k |a_kj|) - log|a_ij|,
// row-minimum duals and a two-pass greedy initial matching, then one shortest augmenting path
// per unmatched column by Dijkstra on reduced costs c_ij - u_i - v_j — and Duff & Pralet's (2005)
// symmetrisation s_i = exp((u_i + v_i - cmax_i) / 2).  The ordering follows metis.rs:118-190
// (match_order_metis, after Hogg & Scott's HSL_MC80): matched 2-cycles become one vertex of a
// condensed graph, which is ordered by the library's nested dissection (METIS is not available)
// and expanded with each pair kept adjacent.
//
// Optimal matchings are not unique: where several attain the maximum product, which one is
// returned depends on the order in which equal-cost paths are explored, so the matching (not its
// product) can differ from the reference's on ties.
#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <string>
#include <vector>

#include "rv_host.h"

namespace rv {

namespace {

constexpr int NONE = -1;
constexpr double INF = std::numeric_limits<double>::infinity();
constexpr double LOG_CLAMP = 500.0;  // exp() stays finite (matching.rs:97)

// Bipartite cost graph: both triangles, rows sorted and unique per column, explicit zeros
// dropped (build_cost_graph, matching.rs:402-475).
struct CostGraph {
    int n = 0;
    std::vector<i64> cp;
    std::vector<int> ri;
    std::vector<double> cost;
    std::vector<double> cmax;  // log max |a_kj| per column, -inf for an empty column
};

CostGraph build_costs(i64 n, const i64 *cp, const i64 *ri, const double *val) {
    CostGraph g;
    g.n = (int)n;
    // counting pass over the mirrored entries, then fill, then sort+dedup each column
    std::vector<i64> cnt(n + 1, 0);
    for (i64 j = 0; j < n; ++j)
        for (i64 p = cp[j]; p < cp[j + 1]; ++p) {
            if (val[p] == 0.0) continue;
            const i64 i = ri[p];
            cnt[j + 1]++;
            if (i != j) cnt[i + 1]++;
        }
    for (i64 j = 0; j < n; ++j) cnt[j + 1] += cnt[j];
    std::vector<std::pair<int, double>> ent(cnt[n]);
    std::vector<i64> pos(cnt.begin(), cnt.end() - 1);
    for (i64 j = 0; j < n; ++j)
        for (i64 p = cp[j]; p < cp[j + 1]; ++p) {
            if (val[p] == 0.0) continue;
            const i64 i = ri[p];
            const double a = std::fabs(val[p]);
            ent[pos[j]++] = {(int)i, a};
            if (i != j) ent[pos[i]++] = {(int)j, a};
        }
    g.cp.assign(n + 1, 0);
    g.cmax.assign(n, -INF);
    for (i64 j = 0; j < n; ++j) {
        auto b = ent.begin() + cnt[j], e = ent.begin() + cnt[j + 1];
        std::sort(b, e, [](const std::pair<int, double> &x, const std::pair<int, double> &y) {
            return x.first != y.first ? x.first < y.first : x.second < y.second;
        });
        int last = -1;
        const size_t start = g.ri.size();
        for (auto it = b; it != e; ++it) {
            if (it->first == last) continue;  // a full-storage input mirrors each entry twice
            last = it->first;
            g.ri.push_back(it->first);
            g.cost.push_back(std::log(it->second));  // log|a| for now
        }
        double m = -INF;
        for (size_t q = start; q < g.ri.size(); ++q) m = std::max(m, g.cost[q]);
        g.cmax[j] = m;
        for (size_t q = start; q < g.ri.size(); ++q) g.cost[q] = m - g.cost[q];
        g.cp[j + 1] = (i64)g.ri.size();
    }
    return g;
}

struct Matching {
    std::vector<int> row_of_col, col_of_row;
    std::vector<i64> col_edge;  // edge index of column j's matched entry
    std::vector<double> u;      // row duals
};

// Initial duals and matching (Duff & Koster §4; greedy_initial_matching, matching.rs:482-583):
// u_i = min_j c_ij; rows take their cheapest column when it is free and not dense; each still
// free column takes its cheapest row by reduced cost, directly or by a length-2 exchange along
// zero-slack edges.
Matching greedy(const CostGraph &g) {
    const int n = g.n;
    Matching M;
    M.row_of_col.assign(n, NONE);
    M.col_of_row.assign(n, NONE);
    M.col_edge.assign(n, -1);
    M.u.assign(n, INF);
    std::vector<int> best_col(n, NONE);
    for (int j = 0; j < n; ++j)
        for (i64 p = g.cp[j]; p < g.cp[j + 1]; ++p) {
            const int i = g.ri[p];
            if (g.cost[p] < M.u[i]) { M.u[i] = g.cost[p]; best_col[i] = j; }
        }
    for (double &x : M.u)
        if (x == INF) x = 0.0;
    const i64 dense = n > 50 ? n / 10 : n;
    for (int i = 0; i < n; ++i) {
        const int j = best_col[i];
        if (j == NONE || M.row_of_col[j] != NONE) continue;
        if (g.cp[j + 1] - g.cp[j] > dense) continue;
        M.col_of_row[i] = j;
        M.row_of_col[j] = i;
    }
    std::vector<double> vcol(n, 0.0);
    std::vector<i64> next_scan(g.cp.begin(), g.cp.end() - 1);
    for (int j = 0; j < n; ++j) {
        if (M.row_of_col[j] != NONE || g.cp[j] == g.cp[j + 1]) continue;
        i64 bk = g.cp[j];
        int bi = g.ri[bk];
        double brc = g.cost[bk] - M.u[bi];
        for (i64 p = bk + 1; p < g.cp[j + 1]; ++p) {
            const int i = g.ri[p];
            const double rc = g.cost[p] - M.u[i];
            if (rc < brc || (rc == brc && M.col_of_row[i] == NONE && M.col_of_row[bi] != NONE)) {
                brc = rc; bi = i; bk = p;
            }
        }
        vcol[j] = brc;
        if (M.col_of_row[bi] == NONE) {
            M.col_of_row[bi] = j;
            M.row_of_col[j] = bi;
            next_scan[j] = bk + 1;
            continue;
        }
        bool done = false;
        for (i64 p = bk; p < g.cp[j + 1] && !done; ++p) {
            const int i = g.ri[p];
            if (g.cost[p] - M.u[i] > brc) continue;
            const int jj = M.col_of_row[i];
            if (jj == NONE) continue;
            for (i64 q = next_scan[jj]; q < g.cp[jj + 1]; ++q) {
                const int ii = g.ri[q];
                if (M.col_of_row[ii] != NONE) continue;
                if (g.cost[q] - M.u[ii] <= vcol[jj]) {
                    M.row_of_col[jj] = ii; M.col_of_row[ii] = jj; next_scan[jj] = q + 1;
                    M.row_of_col[j] = i;   M.col_of_row[i] = j;   next_scan[j] = p + 1;
                    done = true;
                    break;
                }
            }
            if (!done) next_scan[jj] = g.cp[jj + 1];
        }
    }
    for (int j = 0; j < n; ++j) {
        const int i = M.row_of_col[j];
        if (i == NONE) continue;
        for (i64 p = g.cp[j]; p < g.cp[j + 1]; ++p)
            if (g.ri[p] == i) { M.col_edge[j] = p; break; }
    }
    return M;
}

// Indexed binary min-heap of rows keyed by their tentative distance.
struct RowHeap {
    std::vector<int> h;      // heap of rows
    std::vector<int> where;  // position of a row in h, -1 if absent
    const std::vector<double> *d = nullptr;
    explicit RowHeap(int n) : where(n, -1) {}
    bool less(int a, int b) const { return (*d)[a] < (*d)[b]; }
    void up(size_t k) {
        const int r = h[k];
        while (k > 0) {
            const size_t p = (k - 1) / 2;
            if (!less(r, h[p])) break;
            h[k] = h[p]; where[h[k]] = (int)k; k = p;
        }
        h[k] = r; where[r] = (int)k;
    }
    void down(size_t k) {
        const int r = h[k];
        for (;;) {
            size_t c = 2 * k + 1;
            if (c >= h.size()) break;
            if (c + 1 < h.size() && less(h[c + 1], h[c])) ++c;
            if (!less(h[c], r)) break;
            h[k] = h[c]; where[h[k]] = (int)k; k = c;
        }
        h[k] = r; where[r] = (int)k;
    }
    void push_or_decrease(int r) {
        if (where[r] < 0) { h.push_back(r); where[r] = (int)h.size() - 1; }
        up((size_t)where[r]);
    }
    int pop() {
        const int top = h[0];
        where[top] = -1;
        const int last = h.back();
        h.pop_back();
        if (!h.empty()) { h[0] = last; where[last] = 0; down(0); }
        return top;
    }
};

// One shortest augmenting path from free column `root` (Dijkstra over rows on reduced costs;
// dijkstra_augment, matching.rs:672-896).  Rows reached through a matched row's column form the
// search tree; a free row ends a path; the search stops once no tentative distance is below the
// best path found.  Finalised rows' duals move by d_i - csp, keeping every reduced cost >= 0.
struct Dijkstra {
    std::vector<double> d;
    std::vector<char> done;
    std::vector<i64> via_edge;  // edge by which a row was reached
    std::vector<int> via_col;   // column that edge belongs to
    std::vector<int> touched, finals;
    RowHeap heap;
    explicit Dijkstra(int n) : d(n, INF), done(n, 0), via_edge(n, -1), via_col(n, NONE), heap(n) {
        heap.d = &d;
    }
    bool augment(int root, const CostGraph &g, Matching &M) {
        double csp = INF;
        i64 end_edge = -1;
        int end_col = NONE;
        auto relax = [&](int j, double vj) {
            for (i64 p = g.cp[j]; p < g.cp[j + 1]; ++p) {
                const int i = g.ri[p];
                if (done[i]) continue;
                const double dn = vj + g.cost[p] - M.u[i];
                if (dn >= csp) continue;
                if (M.col_of_row[i] == NONE) {
                    csp = dn; end_edge = p; end_col = j;
                } else if (dn < d[i]) {
                    if (d[i] == INF) touched.push_back(i);
                    d[i] = dn; via_edge[i] = p; via_col[i] = j;
                    heap.push_or_decrease(i);
                }
            }
        };
        relax(root, 0.0);
        while (!heap.h.empty()) {
            const int q = heap.h[0];
            if (d[q] >= csp) break;
            heap.pop();
            done[q] = 1;
            finals.push_back(q);
            const int j = M.col_of_row[q];
            relax(j, d[q] - g.cost[M.col_edge[j]] + M.u[q]);
        }
        const bool found = end_col != NONE;
        if (found) {
            for (int i : finals) M.u[i] += d[i] - csp;
            // flip the path: the free row takes end_col; each column on the way takes the row
            // through which the search entered it
            int i = g.ri[end_edge], j = end_col;
            i64 e = end_edge;
            for (;;) {
                const int prev_row = M.row_of_col[j];
                M.row_of_col[j] = i; M.col_of_row[i] = j; M.col_edge[j] = e;
                if (j == root) break;
                i = prev_row;
                e = via_edge[i];
                j = via_col[i];
            }
        }
        for (int r : touched) { d[r] = INF; done[r] = 0; via_edge[r] = -1; via_col[r] = NONE; }
        for (int r : heap.h) heap.where[r] = -1;
        heap.h.clear();
        touched.clear();
        finals.clear();
        return found;
    }
};

}  // namespace

bool mc64_matching(i64 n, const i64 *cp, const i64 *ri, const double *val, Mc64Result &out,
                   std::string &err) {
    out = Mc64Result();
    if (n <= 0) { err = "MC64 requires non-empty matrix"; return false; }
    for (i64 p = 0; p < cp[n]; ++p)
        if (!std::isfinite(val[p])) { err = "MC64 requires finite matrix entries"; return false; }
    if (n == 1) {
        // matching.rs:159-178
        const bool has = cp[1] > cp[0];
        const double a = has ? std::fabs(val[cp[0]]) : 0.0;
        out.matching = {0};
        out.scaling = {a > 0.0 ? 1.0 / std::sqrt(a) : 1.0};
        out.is_matched = {(char)has};
        out.matched = has ? 1 : 0;
        return true;
    }
    const CostGraph g = build_costs(n, cp, ri, val);
    Matching M = greedy(g);
    Dijkstra dj((int)n);
    for (int j = 0; j < (int)n; ++j)
        if (M.row_of_col[j] == NONE) dj.augment(j, g, M);
    i64 matched = 0;
    for (int j = 0; j < (int)n; ++j) matched += M.row_of_col[j] != NONE;
    out.matched = matched;
    // unmatched rows get zero duals (Duff & Koster §4; matching.rs:214-226), then the column
    // duals follow from complementary slackness on the matched edges
    for (int i = 0; i < (int)n; ++i)
        if (M.col_of_row[i] == NONE) M.u[i] = 0.0;
    std::vector<double> v(n, 0.0);
    for (int j = 0; j < (int)n; ++j) {
        const int i = M.row_of_col[j];
        if (i != NONE) v[j] = g.cost[M.col_edge[j]] - M.u[i];
    }
    out.scaling.resize(n);
    for (i64 i = 0; i < n; ++i) {
        double ls = (M.u[i] + v[i] - g.cmax[i]) / 2.0;
        ls = std::min(std::max(ls, -LOG_CLAMP), LOG_CLAMP);
        out.scaling[i] = std::exp(ls);
    }
    out.is_matched.resize(n);
    out.matching.resize(n);
    std::vector<char> used(n, 0);
    for (i64 i = 0; i < n; ++i) {
        out.is_matched[i] = M.col_of_row[i] != NONE;
        if (out.is_matched[i]) { out.matching[i] = M.col_of_row[i]; used[M.col_of_row[i]] = 1; }
    }
    // unmatched rows take the free columns in ascending order (matching.rs:340-370)
    i64 fc = 0;
    for (i64 i = 0; i < n; ++i) {
        if (out.is_matched[i]) continue;
        while (used[fc]) ++fc;
        out.matching[i] = (int)fc;
        used[fc] = 1;
    }
    return true;
}

// Pairs of the matching's cycles (split_matching_cycles, metis.rs:317-400): walking each cycle
// from its smallest index, consecutive elements (j, σ(j)) pair up; an odd cycle leaves its last
// element single; unmatched indices and fixed points are single.
CycleSplit split_cycles(const std::vector<int> &fwd, const std::vector<char> &is_matched) {
    const int n = (int)fwd.size();
    CycleSplit c;
    c.partner.assign(n, -1);
    c.node_of.assign(n, 0);
    std::vector<char> seen(n, 0);
    for (int i = 0; i < n; ++i)
        if (!is_matched[i]) { c.partner[i] = -2; seen[i] = 1; }
    auto single = [&](int i) {
        c.node_of[i] = (int)c.first.size();
        c.first.push_back(i);
        c.singletons++;
        seen[i] = 1;
    };
    for (int i = 0; i < n; ++i) {
        if (c.partner[i] == -2) { c.node_of[i] = (int)c.first.size(); c.first.push_back(i); c.singletons++; continue; }
        if (seen[i]) continue;
        if (fwd[i] == i) { single(i); continue; }
        int j = i;
        for (;;) {
            const int k = fwd[j];
            if (seen[k] || k == i) {
                if (!seen[j]) single(j);
                break;
            }
            c.partner[j] = k;
            c.partner[k] = j;
            c.node_of[j] = c.node_of[k] = (int)c.first.size();
            c.first.push_back(j);
            c.pairs++;
            seen[j] = seen[k] = 1;
            const int nx = fwd[k];
            if (nx == i) break;
            j = nx;
            if (seen[j]) break;
        }
    }
    return c;
}

bool match_order(i64 n, const i64 *cp, const i64 *ri, const double *val, MatchOrderResult &out,
                 std::string &err) {
    out = MatchOrderResult();
    if (n <= 1) {
        // metis.rs:123-134
        out.order.assign(n, 0);
        out.scaling.assign(n, 1.0);
        out.matched = out.condensed_dim = out.singletons = n;
        return true;
    }
    Mc64Result mc;
    if (!mc64_matching(n, cp, ri, val, mc, err)) return false;
    const CycleSplit cs = split_cycles(mc.matching, mc.is_matched);
    const int cdim = (int)cs.first.size();
    // condensed adjacency (build_condensed_adjacency, metis.rs:402-458): an edge between two
    // condensed vertices wherever an entry of A joins their members; no self loops
    std::vector<std::vector<int>> adj(cdim);
    std::vector<int> mark(cdim, -1);
    for (i64 col = 0; col < n; ++col) {
        const int cc = cs.node_of[col];
        mark[cc] = cc;
        for (i64 p = cp[col]; p < cp[col + 1]; ++p) {
            const int cr = cs.node_of[ri[p]];
            if (mark[cr] != cc) {
                mark[cr] = cc;
                adj[cc].push_back(cr);
                adj[cr].push_back(cc);
            }
        }
    }
    std::vector<i64> ccp(cdim + 1, 0);
    std::vector<i64> cri;
    for (int v = 0; v < cdim; ++v) {
        auto &a = adj[v];
        std::sort(a.begin(), a.end());
        a.erase(std::unique(a.begin(), a.end()), a.end());
        // the diagonal keeps the graph reader's vertex-has-entries convention
        cri.push_back(v);
        for (int w : a) cri.push_back(w);
        ccp[v + 1] = (i64)cri.size();
    }
    std::vector<int> corder;
    if (cdim <= 1 || cri.size() == (size_t)cdim) {
        corder.resize(cdim);
        std::iota(corder.begin(), corder.end(), 0);
    } else {
        nested_dissection_graph(cdim, ccp.data(), cri.data(), corder);
    }
    // expansion (expand_ordering, metis.rs:510-535): a pair occupies two consecutive positions
    out.order.clear();
    out.order.reserve(n);
    for (int cnode : corder) {
        const int a = cs.first[cnode];
        out.order.push_back(a);
        if (cs.partner[a] >= 0) out.order.push_back(cs.partner[a]);
    }
    if ((i64)out.order.size() != n) { err = "match ordering expansion lost indices"; return false; }
    out.scaling = std::move(mc.scaling);
    out.matched = mc.matched;
    out.condensed_dim = cdim;
    out.singletons = cs.singletons;
    out.two_cycles = cs.pairs;
    out.matching = std::move(mc.matching);
    out.is_matched = std::move(mc.is_matched);
    return true;
}

}  // namespace rv
