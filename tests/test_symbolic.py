"""Host structure: symbolic analysis, amalgamation, small-leaf classification.

Amalgamation is pinned by the reference's own known-answer cases (restated as input/expected
data from sparse/src/symmetric/amalgamation.rs:260-657) for both the oracle's and the product's
implementation, and the two implementations agree on structures produced by the analysis.
The symbolic analysis is checked against a brute-force dense symbolic Cholesky: every
supernode pattern equals the nonzero rows of its columns of L.
"""
import numpy as np
import pytest

import oracle
import rivrs_sparse as rs
import workloads


def _sn(b, e, pat, parent):
    return (b, e, list(pat), -1 if parent is None else parent)


def _run(impl, sns, nemin):
    begin = [s[0] for s in sns]
    end = [s[1] for s in sns]
    pp = [0]
    pat = []
    for s in sns:
        pat.extend(s[2])
        pp.append(len(pat))
    parent = [s[3] for s in sns]
    if impl == "oracle":
        return oracle.amalgamate(begin, end, pp, pat, parent, nemin)
    return rs.amalgamate(begin, end, pp, pat, parent, nemin)


# (name, supernodes, nemin, expected count, expected [(col_begin, col_end, pattern)] or None)
KATS = [
    ("no_merges_large", [_sn(0, 40, [200, 201], 4), _sn(40, 80, [200, 202], 4), _sn(80, 120, [200, 203], 4),
                         _sn(120, 160, [200, 204], 4), _sn(160, 200, [], None)], 32, 5, None),
    ("nemin_pair", [_sn(0, 4, [4, 5, 10], 1), _sn(4, 8, [10], None)], 32, 1, [(0, 8, [10])]),
    ("structural_match", [_sn(0, 3, [3, 10, 20], 1), _sn(3, 4, [10, 20, 30, 40], None)], 32, 1,
     [(0, 4, [10, 20, 30, 40])]),
    ("chain", [_sn(0, 2, [2, 3, 4, 5, 6, 7, 8, 9], 1), _sn(2, 4, [4, 5, 6, 7, 8, 9], 2), _sn(4, 6, [6, 7, 8, 9], 3),
               _sn(6, 8, [8, 9], 4), _sn(8, 10, [], None)], 32, 1, [(0, 10, [])]),
    ("bushy", [_sn(0, 2, [8, 9], 4), _sn(2, 4, [8, 9], 4), _sn(4, 6, [8, 9], 4), _sn(6, 8, [8, 9], 4),
               _sn(8, 10, [], None)], 32, 1, [(0, 10, [])]),
    ("reparenting", [_sn(0, 2, [4, 5, 8, 9], 2), _sn(2, 4, [4, 5, 8, 9], 2), _sn(4, 6, [8, 9], 3),
                     _sn(8, 10, [], None)], 32, 1, [(0, 10, [])]),
    ("pattern_union", [_sn(0, 2, [2, 5, 10, 20], 1), _sn(2, 4, [5, 15, 20], None)], 32, 1, [(0, 4, [5, 10, 15, 20])]),
    ("single", [_sn(0, 10, [], None)], 32, 1, [(0, 10, [])]),
    ("nemin1", [_sn(0, 4, [4, 5, 6], 1), _sn(4, 8, [5, 6], None)], 1, 2, [(0, 4, [4, 5, 6]), (4, 8, [5, 6])]),
    ("nemin32_large", [_sn(0, 40, [40, 50, 60], 1), _sn(40, 80, [50, 60], None)], 32, 2, None),
    ("nemin64_merge", [_sn(0, 40, [40, 50, 60], 1), _sn(40, 80, [50, 60], None)], 64, 1, [(0, 80, [])]),
]


@pytest.mark.parametrize("impl", ["oracle", "product"])
@pytest.mark.parametrize("name,sns,nemin,count,expected", KATS, ids=[k[0] for k in KATS])
def test_amalgamation_kats(impl, name, sns, nemin, count, expected):
    out = _run(impl, sns, nemin)
    assert len(out) == count
    if expected is not None:
        got = [(o["col_begin"], o["col_end"], o["pattern"]) for o in out]
        if name == "nemin64_merge":
            got = [(g[0], g[1], []) for g in got]
        assert got == expected
    for i, o in enumerate(out):
        assert o["parent"] == -1 or o["parent"] > i  # postorder preserved


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_amalgamation_partial_mixed_sizes(impl):
    # amalgamation.rs test_partial_merge_mixed_sizes: 2 small children merge, the large stays
    sns = [_sn(0, 4, [100, 140], 3), _sn(4, 8, [100, 140], 3), _sn(8, 48, [100, 140], 3),
           _sn(100, 104, [140], 4), _sn(140, 190, [], None)]
    out = _run(impl, sns, 32)
    assert len(out) == 3
    assert out[0]["col_end"] - out[0]["col_begin"] == 40


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_amalgamation_simplicial_chain(impl):
    n = 100
    sns = [_sn(i, i + 1, range(i + 1, min(n, i + 6)), i + 1 if i + 1 < n else None) for i in range(n)]
    out = _run(impl, sns, 32)
    assert len(out) < 20
    for o in out:
        assert o["col_begin"] < o["col_end"]
        for r in o["pattern"]:
            assert r < o["col_begin"] or r >= o["col_end"]


def _lower_pattern_dense(A, perm):
    Ad = (A.to_scipy().toarray() != 0)[np.ix_(perm, perm)]
    n = A.n
    L = np.tril(Ad)
    # symbolic Cholesky by brute force: column j's rows propagate to its first off-diagonal row
    for j in range(n):
        rows = np.nonzero(L[j + 1:, j])[0] + j + 1
        if len(rows):
            p = rows[0]
            L[rows, p] = True
    return L


@pytest.mark.parametrize("kind", ["kkt", "helm", "fixture"])
def test_supernode_patterns_match_brute_force(kind):
    if kind == "kkt":
        A = workloads.kkt(8, seed=4)[0]
        opts = rs.AnalyzeOptions(rs.OrderingStrategy.Metis)
    elif kind == "helm":
        A, _, info = workloads.helmholtz(6, seed=0)
        opts = rs.AnalyzeOptions(rs.OrderingStrategy.Grid, grid=info["grid"])
    else:
        import gen
        A = rs.load_mtx([c for c in gen.hand_constructed() if c[0] == "bordered-block-20"][0][1])
        opts = rs.AnalyzeOptions(rs.OrderingStrategy.Metis)
    st = rs.SparseLDLT.analyze(A, opts).structure()
    n = A.n
    assert sorted(st.perm_fwd.tolist()) == list(range(n))
    L = _lower_pattern_dense(A, st.perm_fwd)
    covered = 0
    for s in range(len(st.sn_begin)):
        b, e = st.sn_begin[s], st.sn_end[s]
        covered += e - b
        pat = st.pat[st.pat_ptr[s]:st.pat_ptr[s + 1]].tolist()
        rows = sorted(set(np.nonzero(L[e:, b:e].any(axis=1))[0] + e))
        assert pat == rows, s
        # parent = supernode holding the first pattern row (the etree parent of the last column)
        if pat:
            p = st.parent[s]
            assert st.sn_begin[p] <= pat[0] < st.sn_end[p]
        else:
            assert st.parent[s] == -1
    assert covered == n


def test_product_and_oracle_amalgamation_agree():
    A, _, info = workloads.helmholtz(16, seed=1)
    st = rs.SparseLDLT.analyze(A, rs.AnalyzeOptions(rs.OrderingStrategy.Grid, grid=info["grid"])).structure()
    args = (st.sn_begin, st.sn_end, st.pat_ptr, st.pat, st.parent)
    for nemin in (1, 8, 32):
        a = oracle.amalgamate(*args, nemin)
        b = rs.amalgamate(*args, nemin)
        assert a == b


def test_grid_nested_dissection_separators():
    # the root supernode of an N^3 grid ordered by planes contains the middle x-plane
    N = 12
    A, _, info = workloads.helmholtz(N, seed=0)
    st = rs.SparseLDLT.analyze(A, rs.AnalyzeOptions(rs.OrderingStrategy.Grid, grid=info["grid"])).structure()
    last = st.perm_fwd[-N * N:]
    assert np.all(last % N == N // 2)


def test_user_perm_and_errors():
    A = workloads.kkt(6, seed=0)[0]
    perm = np.arange(A.n)[::-1].copy()
    st = rs.SparseLDLT.analyze(A, rs.AnalyzeOptions(rs.OrderingStrategy.UserSupplied, perm=perm)).structure()
    assert len(st.perm_fwd) == A.n
    with pytest.raises(rs.DimensionMismatch):
        rs.SparseLDLT.analyze(A, rs.AnalyzeOptions(rs.OrderingStrategy.UserSupplied, perm=perm[:-1]))
    bad = perm.copy()
    bad[0] = bad[1]
    with pytest.raises(rs.SparseError):
        rs.SparseLDLT.analyze(A, rs.AnalyzeOptions(rs.OrderingStrategy.UserSupplied, perm=bad))
    with pytest.raises(rs.AnalysisFailure):
        rs.SparseLDLT.analyze(A, rs.AnalyzeOptions(rs.OrderingStrategy.MatchOrderMetis))
