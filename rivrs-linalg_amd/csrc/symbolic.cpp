// symbolic.cpp -- host symbolic analysis producing the supernode partition that the
// numeric phase consumes.
//
// The reference delegates this to faer's factorize_symbolic_cholesky and METIS
// (sparse/src/symmetric/symbolic.rs:248-310, sparse/src/ordering/metis.rs:559-601), neither
// of which exists here.  This file provides an independent implementation of the same
// products: a fill-reducing ordering (nested dissection on the graph, or on a structured
// grid), the elimination tree (Liu's algorithm), a postorder, column counts, fundamental
// supernodes, supernode patterns and the supernode parent derived from the etree exactly as
// AptpSymbolic::supernode_parent does (symbolic.rs:576-598).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "rv_host.h"

namespace rv {

namespace {

struct Graph {
    i64 n = 0;
    std::vector<i64> xadj;
    std::vector<int> adj;
};

// Symmetrized off-diagonal adjacency of a CSC pattern (duplicates removed).
Graph build_graph(i64 n, const i64 *cp, const i64 *ri) {
    std::vector<i64> deg(n, 0);
    for (i64 j = 0; j < n; ++j)
        for (i64 p = cp[j]; p < cp[j + 1]; ++p) {
            const i64 i = ri[p];
            if (i == j) continue;
            deg[i]++;
            deg[j]++;
        }
    Graph g;
    g.n = n;
    g.xadj.assign(n + 1, 0);
    for (i64 j = 0; j < n; ++j) g.xadj[j + 1] = g.xadj[j] + deg[j];
    g.adj.resize(g.xadj[n]);
    std::vector<i64> pos(g.xadj.begin(), g.xadj.end() - 1);
    for (i64 j = 0; j < n; ++j)
        for (i64 p = cp[j]; p < cp[j + 1]; ++p) {
            const i64 i = ri[p];
            if (i == j) continue;
            g.adj[pos[i]++] = (int)j;
            g.adj[pos[j]++] = (int)i;
        }
    // sort + dedup each list
    std::vector<i64> nx(n + 1, 0);
    i64 w = 0;
    for (i64 j = 0; j < n; ++j) {
        int *b = g.adj.data() + g.xadj[j], *e = g.adj.data() + g.xadj[j + 1];
        std::sort(b, e);
        int *u = std::unique(b, e);
        nx[j] = w;
        for (int *q = b; q < u; ++q) g.adj[w++] = *q;
    }
    nx[n] = w;
    g.adj.resize(w);
    g.xadj = nx;
    return g;
}

const int ND_LEAF = 64;

}  // namespace

void relax_supernodes(Symbolic &out);

namespace {

// Minimum-degree order of a small vertex set (<= 64) on its induced subgraph, using a
// bitmask elimination graph.  Ties break to the smallest local index.
void leaf_min_degree(const Graph &g, const std::vector<int> &vs, std::vector<int> &loc,
                     std::vector<int> &order) {
    const int k = (int)vs.size();
    if (k > 64) {
        order.insert(order.end(), vs.begin(), vs.end());
        return;
    }
    for (int t = 0; t < k; ++t) loc[vs[t]] = t;
    std::vector<uint64_t> nb(k, 0);
    for (int t = 0; t < k; ++t) {
        const int v = vs[t];
        for (i64 p = g.xadj[v]; p < g.xadj[v + 1]; ++p) {
            const int u = g.adj[p];
            const int lu = loc[u];
            if (lu >= 0 && lu < k && vs[lu] == u) nb[t] |= (uint64_t)1 << lu;
        }
    }
    uint64_t alive = (k == 64) ? ~(uint64_t)0 : (((uint64_t)1 << k) - 1);
    for (int step = 0; step < k; ++step) {
        int best = -1, bd = 1 << 30;
        for (int t = 0; t < k; ++t)
            if (alive >> t & 1) {
                const int d = __builtin_popcountll(nb[t] & alive);
                if (d < bd) { bd = d; best = t; }
            }
        const uint64_t nbs = nb[best] & alive;
        for (int t = 0; t < k; ++t)
            if (nbs >> t & 1) nb[t] |= nbs & ~((uint64_t)1 << t);
        alive &= ~((uint64_t)1 << best);
        order.push_back(vs[best]);
    }
    for (int t = 0; t < k; ++t) loc[vs[t]] = -1;
}

struct NDWork {
    std::vector<int> tag, lvl, loc, queue;
    int next_tag = 1;
};

// BFS restricted to vertices carrying tag t; returns vertices in BFS order, levels in lvl.
void bfs(const Graph &g, int root, int t, NDWork &w, std::vector<int> &out) {
    out.clear();
    out.push_back(root);
    w.lvl[root] = 0;
    w.tag[root] = -t;  // visited marker
    for (size_t h = 0; h < out.size(); ++h) {
        const int v = out[h];
        for (i64 p = g.xadj[v]; p < g.xadj[v + 1]; ++p) {
            const int u = g.adj[p];
            if (w.tag[u] == t) {
                w.tag[u] = -t;
                w.lvl[u] = w.lvl[v] + 1;
                out.push_back(u);
            }
        }
    }
    for (int v : out) w.tag[v] = t;
}

void nd_rec(const Graph &g, std::vector<int> vs, NDWork &w, std::vector<int> &order) {
    if ((int)vs.size() <= ND_LEAF) {
        leaf_min_degree(g, vs, w.loc, order);
        return;
    }
    const int t = w.next_tag++;
    for (int v : vs) w.tag[v] = t;
    std::vector<int> comp;
    bfs(g, vs[0], t, w, comp);
    if (comp.size() < vs.size()) {
        // split into connected components and order each independently
        std::vector<std::vector<int>> comps;
        const int tc = w.next_tag++;
        for (int v : comp) w.tag[v] = tc;
        comps.push_back(comp);
        for (int v : vs)
            if (w.tag[v] == t) {
                std::vector<int> c2;
                bfs(g, v, t, w, c2);
                for (int u : c2) w.tag[u] = tc;
                comps.push_back(c2);
            }
        for (auto &c : comps) nd_rec(g, c, w, order);
        return;
    }
    // pseudo-peripheral root
    int root = vs[0];
    int ecc = -1;
    std::vector<int> bo;
    for (int it = 0; it < 6; ++it) {
        bfs(g, root, t, w, bo);
        const int e = w.lvl[bo.back()];
        if (e <= ecc) break;
        ecc = e;
        // farthest vertex of minimum degree
        int best = bo.back();
        i64 bd = g.xadj[best + 1] - g.xadj[best];
        for (int q = (int)bo.size() - 1; q >= 0 && w.lvl[bo[q]] == e; --q) {
            const i64 d = g.xadj[bo[q] + 1] - g.xadj[bo[q]];
            if (d < bd) { bd = d; best = bo[q]; }
        }
        root = best;
    }
    bfs(g, root, t, w, bo);
    const int nl = w.lvl[bo.back()] + 1;
    std::vector<i64> cnt(nl, 0);
    for (int v : bo) cnt[w.lvl[v]]++;
    const i64 N = (i64)bo.size();
    int sl = -1;
    i64 bestsz = 0, bestimb = 0;
    i64 before = 0;
    for (int l = 0; l < nl; ++l) {
        const i64 after = N - before - cnt[l];
        if (l > 0 && l < nl - 1 && before * 4 >= N && after * 4 >= N) {
            const i64 imb = before > after ? before - after : after - before;
            if (sl < 0 || cnt[l] < bestsz || (cnt[l] == bestsz && imb < bestimb)) {
                sl = l; bestsz = cnt[l]; bestimb = imb;
            }
        }
        before += cnt[l];
    }
    if (sl < 0) {
        before = 0;
        for (int l = 0; l < nl; ++l) {
            if (before + cnt[l] >= N / 2) { sl = l; break; }
            before += cnt[l];
        }
        if (sl <= 0 || sl >= nl - 1) {
            // no useful separator (e.g. dense): order by level
            order.insert(order.end(), bo.begin(), bo.end());
            return;
        }
    }
    std::vector<int> pa, pb, sep;
    for (int v : bo) {
        if (w.lvl[v] < sl) pa.push_back(v);
        else if (w.lvl[v] > sl) pb.push_back(v);
    }
    // thin the separator: a separator vertex with no neighbour beyond the separator
    // level is not needed to separate and joins part A.
    for (int v : bo)
        if (w.lvl[v] == sl) {
            bool touches_b = false;
            for (i64 p = g.xadj[v]; p < g.xadj[v + 1] && !touches_b; ++p) {
                const int u = g.adj[p];
                if (w.tag[u] == t && w.lvl[u] > sl) touches_b = true;
            }
            if (touches_b) sep.push_back(v);
            else pa.push_back(v);
        }
    for (int v : vs) w.tag[v] = 0;
    nd_rec(g, pa, w, order);
    nd_rec(g, pb, w, order);
    order.insert(order.end(), sep.begin(), sep.end());
}

void grid_rec(int nx, int ny, int x0, int x1, int y0, int y1, int z0, int z1,
              std::vector<int> &order) {
    const int dx = x1 - x0, dy = y1 - y0, dz = z1 - z0;
    if (dx <= 0 || dy <= 0 || dz <= 0) return;
    if ((i64)dx * dy * dz <= 8 || (dx <= 2 && dy <= 2 && dz <= 2)) {
        for (int z = z0; z < z1; ++z)
            for (int y = y0; y < y1; ++y)
                for (int x = x0; x < x1; ++x) order.push_back(x + nx * (y + ny * z));
        return;
    }
    if (dx >= dy && dx >= dz) {
        const int mid = x0 + dx / 2;
        grid_rec(nx, ny, x0, mid, y0, y1, z0, z1, order);
        grid_rec(nx, ny, mid + 1, x1, y0, y1, z0, z1, order);
        for (int z = z0; z < z1; ++z)
            for (int y = y0; y < y1; ++y) order.push_back(mid + nx * (y + ny * z));
    } else if (dy >= dz) {
        const int mid = y0 + dy / 2;
        grid_rec(nx, ny, x0, x1, y0, mid, z0, z1, order);
        grid_rec(nx, ny, x0, x1, mid + 1, y1, z0, z1, order);
        for (int z = z0; z < z1; ++z)
            for (int x = x0; x < x1; ++x) order.push_back(x + nx * (mid + ny * z));
    } else {
        const int mid = z0 + dz / 2;
        grid_rec(nx, ny, x0, x1, y0, y1, z0, mid, order);
        grid_rec(nx, ny, x0, x1, y0, y1, mid + 1, z1, order);
        for (int y = y0; y < y1; ++y)
            for (int x = x0; x < x1; ++x) order.push_back(x + nx * (y + ny * mid));
    }
}

}  // namespace

void nested_dissection_graph(i64 n, const i64 *cp, const i64 *ri, std::vector<int> &perm) {
    Graph g = build_graph(n, cp, ri);
    NDWork w;
    w.tag.assign(n, 0);
    w.lvl.assign(n, 0);
    w.loc.assign(n, -1);
    perm.clear();
    perm.reserve(n);
    // dense rows (degree above 10 sqrt(n), the AMD threshold) would sit in every BFS level
    // and spoil the separators: they are ordered last
    const i64 dense = std::max<i64>(16, (i64)(10.0 * std::sqrt((double)n)));
    std::vector<int> sparse_v, dense_v;
    for (i64 v = 0; v < n; ++v) {
        if (g.xadj[v + 1] - g.xadj[v] > dense) dense_v.push_back((int)v);
        else sparse_v.push_back((int)v);
    }
    nd_rec(g, sparse_v, w, perm);
    perm.insert(perm.end(), dense_v.begin(), dense_v.end());
}

void nested_dissection_grid(int nx, int ny, int nz, std::vector<int> &perm) {
    perm.clear();
    perm.reserve((size_t)nx * ny * nz);
    grid_rec(nx, ny, 0, nx, 0, ny, 0, nz, perm);
}

bool analyze(i64 n, const i64 *cp, const i64 *ri, int ordering, const int *user_perm,
             const int *grid_dims, Symbolic &out, std::string &err) {
    out = Symbolic();
    out.n = n;
    if (n < 0) { err = "negative dimension"; return false; }
    for (i64 j = 0; j < n; ++j)
        if (cp[j + 1] < cp[j]) { err = "column pointers not monotone"; return false; }
    for (i64 p = 0; p < cp[n]; ++p)
        if (ri[p] < 0 || ri[p] >= n) { err = "row index out of range"; return false; }
    std::vector<int> perm;
    switch (ordering) {
        case ORDER_NATURAL:
            perm.resize(n);
            std::iota(perm.begin(), perm.end(), 0);
            break;
        case ORDER_USER: {
            if (!user_perm) { err = "user permutation missing"; return false; }
            perm.assign(user_perm, user_perm + n);
            break;
        }
        case ORDER_ND_GRAPH:
            nested_dissection_graph(n, cp, ri, perm);
            break;
        case ORDER_ND_GRID: {
            if (!grid_dims) { err = "grid dimensions missing"; return false; }
            const i64 cells = (i64)grid_dims[0] * grid_dims[1] * grid_dims[2];
            if (cells > n || cells <= 0) { err = "grid dimensions exceed matrix dimension"; return false; }
            nested_dissection_grid(grid_dims[0], grid_dims[1], grid_dims[2], perm);
            // rows beyond the grid (e.g. constraint rows) are ordered last
            for (i64 i = cells; i < n; ++i) perm.push_back((int)i);
            break;
        }
        default:
            err = "unknown ordering";
            return false;
    }
    {
        std::vector<char> seen(n, 0);
        if ((i64)perm.size() != n) { err = "permutation has wrong length"; return false; }
        for (int v : perm) {
            if (v < 0 || v >= n || seen[v]) { err = "invalid permutation"; return false; }
            seen[v] = 1;
        }
    }
    Graph g = build_graph(n, cp, ri);
    std::vector<int> pinv(n);
    for (i64 i = 0; i < n; ++i) pinv[perm[i]] = (int)i;
    // elimination tree of the permuted pattern (Liu)
    std::vector<int> parent(n, -1), anc(n, -1);
    for (i64 j = 0; j < n; ++j) {
        const int oj = perm[j];
        for (i64 p = g.xadj[oj]; p < g.xadj[oj + 1]; ++p) {
            int r = pinv[g.adj[p]];
            if (r >= j) continue;
            while (anc[r] != -1 && anc[r] != (int)j) {
                const int t = anc[r];
                anc[r] = (int)j;
                r = t;
            }
            if (anc[r] == -1) {
                anc[r] = (int)j;
                parent[r] = (int)j;
            }
        }
    }
    // postorder (children in increasing order)
    std::vector<int> head(n, -1), nxt(n, -1);
    for (i64 j = n - 1; j >= 0; --j)
        if (parent[j] >= 0) { nxt[j] = head[parent[j]]; head[parent[j]] = (int)j; }
    std::vector<int> post;
    post.reserve(n);
    std::vector<int> stack;
    for (i64 r = 0; r < n; ++r) {
        if (parent[r] != -1) continue;
        stack.push_back((int)r);
        while (!stack.empty()) {
            const int v = stack.back();
            if (head[v] != -1) {
                const int c = head[v];
                head[v] = nxt[c];
                stack.push_back(c);
            } else {
                stack.pop_back();
                post.push_back(v);
            }
        }
    }
    std::vector<int> postinv(n);
    for (i64 k = 0; k < n; ++k) postinv[post[k]] = (int)k;
    out.perm_fwd.resize(n);
    out.perm_inv.resize(n);
    out.etree.assign(n, -1);
    for (i64 k = 0; k < n; ++k) {
        out.perm_fwd[k] = perm[post[k]];
        const int pa = parent[post[k]];
        out.etree[k] = pa < 0 ? -1 : postinv[pa];
    }
    for (i64 k = 0; k < n; ++k) out.perm_inv[out.perm_fwd[k]] = (int)k;
    const std::vector<int> &pf = out.perm_fwd, &pi = out.perm_inv;
    const std::vector<int> &et = out.etree;
    // column counts via row subtrees
    out.col_counts.assign(n, 1);
    {
        std::vector<int> mark(n, -1);
        for (i64 i = 0; i < n; ++i) {
            mark[i] = (int)i;
            const int oi = pf[i];
            for (i64 p = g.xadj[oi]; p < g.xadj[oi + 1]; ++p) {
                int k = pi[g.adj[p]];
                if (k >= i) continue;
                while (mark[k] != (int)i) {
                    out.col_counts[k]++;
                    mark[k] = (int)i;
                    k = et[k];
                }
            }
        }
    }
    // fundamental supernodes
    std::vector<int> nchild(n, 0);
    for (i64 j = 0; j < n; ++j)
        if (et[j] >= 0) nchild[et[j]]++;
    std::vector<int> sn_of(n);
    std::vector<int> begins;
    for (i64 j = 0; j < n; ++j) {
        const bool merge = j > 0 && et[j - 1] == (int)j && nchild[j] == 1 &&
                           out.col_counts[j - 1] == out.col_counts[j] + 1;
        if (!merge) begins.push_back((int)j);
        sn_of[j] = (int)begins.size() - 1;
    }
    const int ns = (int)begins.size();
    out.sn.resize(ns);
    for (int s = 0; s < ns; ++s) {
        out.sn[s].col_begin = begins[s];
        out.sn[s].col_end = (s + 1 < ns) ? begins[s + 1] : (int)n;
        out.sn[s].owned = {ColRange{out.sn[s].col_begin, out.sn[s].col_end}};
        const int last = out.sn[s].col_end - 1;
        out.sn[s].parent = et[last] < 0 ? -1 : sn_of[et[last]];
    }
    std::vector<std::vector<int>> sch(ns);
    for (int s = 0; s < ns; ++s)
        if (out.sn[s].parent >= 0) sch[out.sn[s].parent].push_back(s);
    {
        std::vector<int> mark(n, -1);
        for (int s = 0; s < ns; ++s) {
            Supernode &S = out.sn[s];
            const int last = S.col_end - 1;
            std::vector<int> &pat = S.pattern;
            for (int j = S.col_begin; j < S.col_end; ++j) {
                const int oj = pf[j];
                for (i64 p = g.xadj[oj]; p < g.xadj[oj + 1]; ++p) {
                    const int i = pi[g.adj[p]];
                    if (i > last && mark[i] != s) { mark[i] = s; pat.push_back(i); }
                }
            }
            for (int c : sch[s])
                for (int i : out.sn[c].pattern)
                    if (i > last && mark[i] != s) { mark[i] = s; pat.push_back(i); }
            std::sort(pat.begin(), pat.end());
        }
    }
    relax_supernodes(out);
    out.predicted_nnz = 0;
    for (const Supernode &S : out.sn) {
        const i64 nc = S.col_end - S.col_begin;
        out.predicted_nnz += nc * (nc + 1) / 2 + nc * (i64)S.pattern.size();
    }
    return true;
}

// Relaxed supernodes (Ashcraft & Grimes 1989; parameters of CHOLMOD's published defaults:
// nrelax = 4, 16, 48 and zrelax = 0.8, 0.1, 0.05).  A supernode merges into its parent when it
// immediately precedes it in postorder (so the columns stay contiguous) and the merged block's
// fraction of explicitly stored zeros stays under the size-dependent limit.  The parent's pattern
// already contains the child's rows beyond the parent's columns, so the merged pattern is the
// parent's.
void relax_supernodes(Symbolic &out) {
    const int ns = (int)out.sn.size();
    if (ns <= 1) return;
    std::vector<int> rep(ns);
    for (int s = 0; s < ns; ++s) rep[s] = s;
    std::vector<i64> zeros(ns, 0);
    std::vector<char> gone(ns, 0);
    auto stored = [](i64 nc, i64 np) { return nc * (nc + 1) / 2 + nc * np; };
    for (int s = 0; s < ns; ++s) {
        Supernode &S = out.sn[s];
        const int p = S.parent;
        if (p < 0) continue;
        Supernode &P = out.sn[p];
        if (S.col_end != P.col_begin) continue;
        const i64 ncs = S.col_end - S.col_begin, ncp = P.col_end - P.col_begin;
        const i64 np = (i64)P.pattern.size();
        const i64 nc = ncs + ncp;
        const i64 es = stored(ncs, (i64)S.pattern.size()), ep = stored(ncp, np);
        const i64 em = stored(nc, np);
        const i64 z = em - (es - zeros[s]) - (ep - zeros[p]);
        const double zf = em > 0 ? (double)z / (double)em : 0.0;
        const bool merge = nc <= 4 || (nc <= 16 && zf < 0.8) || (nc <= 48 && zf < 0.1) || zf < 0.05;
        if (!merge) continue;
        P.col_begin = S.col_begin;
        P.owned = {ColRange{P.col_begin, P.col_end}};
        zeros[p] = z;
        gone[s] = 1;
        rep[s] = p;
    }
    // renumber survivors, parents through representatives
    auto find = [&](int x) {
        while (rep[x] != x) x = rep[x];
        return x;
    };
    std::vector<int> newid(ns, -1);
    int cnt = 0;
    for (int s = 0; s < ns; ++s)
        if (!gone[s]) newid[s] = cnt++;
    std::vector<Supernode> res;
    res.reserve(cnt);
    for (int s = 0; s < ns; ++s) {
        if (gone[s]) continue;
        Supernode x = std::move(out.sn[s]);
        x.parent = x.parent < 0 ? -1 : newid[find(x.parent)];
        res.push_back(std::move(x));
    }
    out.sn.swap(res);
}

}  // namespace rv
