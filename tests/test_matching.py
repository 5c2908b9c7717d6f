"""MC64 matching, symmetric scaling and the matching-based ordering (host code, CPU).

Pinned by what fixes the answer independently of the implementation:
* optimality: the product of matched magnitudes equals the maximum that SciPy's Hungarian-type
  solver (min_weight_full_bipartite_matching on -log|a|) finds, on random sparse symmetric
  matrices and on the SuiteSparse-CI fixtures;
* the dual certificate as the reference's tests state it (sparse/src/testing/mc64_validation.rs:
  113-180): |s_i a_ij s_j| <= 1 between matched indices and every nonzero row reaches 1;
* the reference's own unit matrices (matching.rs:1400-1500, tests/mc64_matching.rs) with their
  stated answers (identity matching for diagonal / diagonally dominant matrices, 1/sqrt(a) for 1x1);
* the cycle split (metis.rs:986-1060 cases): pair/singleton counts and pair adjacency in the order.
A plausible mistake — a dual updated with the wrong sign, a path flipped off by one column,
costs not normalised per column — breaks the optimum or the certificate.
"""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import min_weight_full_bipartite_matching

import gen
import rivrs_sparse as rs
import workloads


def _sym(n, entries):
    r, c, v = [], [], []
    for i, j, x in entries:
        r.append(i); c.append(j); v.append(x)
        if i != j:
            r.append(j); c.append(i); v.append(x)
    A = sp.csc_matrix((v, (r, c)), shape=(n, n))
    A.sum_duplicates()
    A.sort_indices()
    return rs.SparseColMat(n, A.indptr, A.indices, A.data), A


def _certificate(A, res, full=True):
    S = sp.diags(res.scaling) @ A @ sp.diags(res.scaling)
    S = abs(S.tocoo())
    m = res.is_matched
    both = m[S.row] & m[S.col]
    if both.any():
        assert S.data[both].max() <= 1.0 + 1e-10
    if full:
        rowmax = np.zeros(A.shape[0])
        np.maximum.at(rowmax, S.row, S.data)
        nz = rowmax > 0
        assert rowmax[nz].min() >= 1.0 - 1e-12


def _log_product(A, matching):
    A = A.tocsc()
    return sum(np.log(abs(A[i, j])) for i, j in enumerate(matching))


def _scipy_best(A):
    A = A.tocsr().copy()
    A.data = np.abs(A.data)
    A.eliminate_zeros()
    W = A.copy()
    W.data = -np.log(W.data) + (np.log(W.data).max() + 1.0)  # positive weights, same argmin
    r, c = min_weight_full_bipartite_matching(W)
    return sum(np.log(A[i, j]) for i, j in zip(r, c))


UNIT = [
    ("diagonal", 3, [(0, 0, 4.0), (1, 1, 9.0), (2, 2, 1.0)]),
    ("tridiagonal_indefinite", 4, [(0, 0, 2.0), (0, 1, -1.0), (1, 1, -3.0), (1, 2, 2.0), (2, 2, 1.0),
                                   (2, 3, -1.0), (3, 3, -4.0)]),
    ("arrow_indefinite", 5, [(0, 0, 10.0), (0, 1, 1.0), (0, 2, 1.0), (0, 3, 1.0), (0, 4, 1.0), (1, 1, -3.0),
                             (2, 2, 5.0), (3, 3, -2.0), (4, 4, 4.0)]),
    ("trivial_2x2", 2, [(0, 0, 3.0), (0, 1, 1.0), (1, 1, 5.0)]),
    ("badly_scaled", 4, [(0, 0, 1e10), (0, 1, 1e5), (1, 1, 1e0), (1, 2, 1e-5), (2, 2, 1e-10), (2, 3, 1e-3),
                         (3, 3, 1e3)]),
    ("well_conditioned_pd", 4, [(0, 0, 4.0), (0, 1, 1.0), (0, 2, 0.5), (1, 1, 4.0), (1, 2, 1.0), (1, 3, 0.5),
                                (2, 2, 4.0), (2, 3, 1.0), (3, 3, 4.0)]),
    ("6x6_indefinite", 6, [(0, 0, 4.0), (0, 1, 1.0), (0, 4, 1.0), (1, 1, -3.0), (1, 2, 2.0), (2, 2, 5.0),
                           (2, 3, 1.0), (3, 3, -2.0), (3, 4, 3.0), (4, 4, 6.0), (4, 5, 1.0), (5, 5, -1.0)]),
]


@pytest.mark.parametrize("name,n,entries", UNIT, ids=[u[0] for u in UNIT])
def test_reference_unit_matrices(name, n, entries):
    M, A = _sym(n, entries)
    r = rs.mc64_matching(M)
    assert r.matched == n
    assert np.all(r.scaling > 0) and np.all(np.isfinite(r.scaling))
    _certificate(A, r)
    assert _log_product(A, r.matching) == pytest.approx(_scipy_best(A), abs=1e-9)
    if name in ("diagonal", "well_conditioned_pd"):
        assert list(r.matching) == list(range(n))
    if name == "diagonal":
        np.testing.assert_allclose(r.scaling, [0.5, 1 / 3, 1.0], rtol=1e-14)


def test_trivial_and_errors():
    M, _ = _sym(1, [(0, 0, 7.0)])
    r = rs.mc64_matching(M)
    assert r.matched == 1 and r.scaling[0] == pytest.approx(1 / np.sqrt(7.0), rel=1e-15)
    with pytest.raises(rs.InvalidInput):
        rs.mc64_matching(rs.SparseColMat(0, [0], [], []))
    M, _ = _sym(2, [(0, 0, np.inf), (1, 1, 1.0)])
    with pytest.raises(rs.InvalidInput):
        rs.mc64_matching(M)


def test_structurally_singular():
    # zero diagonal path 0-1-2: rows 0 and 1 pair up, one index stays unmatched
    M, A = _sym(3, [(0, 1, 5.0), (1, 2, 3.0)])
    r = rs.mc64_matching(M)
    assert r.matched == 2
    assert sorted(r.matching) == [0, 1, 2]
    assert np.all(r.scaling > 0) and np.all(np.isfinite(r.scaling))
    _certificate(A, r, full=False)
    # a row without entries is unmatched and scaled finitely
    M, A = _sym(4, [(0, 0, 4.0), (0, 1, 2.0), (1, 1, 5.0), (1, 2, 1.0), (2, 2, 3.0)])
    r = rs.mc64_matching(M)
    assert r.matched == 3 and not r.is_matched[3]
    assert np.all(np.isfinite(r.scaling))


@pytest.mark.parametrize("seed", range(6))
def test_random_sparse_optimal(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(30, 300))
    B = sp.random(n, n, density=min(1.0, 4.0 / n), random_state=seed, format="coo")
    B.data = rng.standard_normal(B.nnz) * np.exp(rng.uniform(-8, 8, B.nnz))
    d = rng.standard_normal(n) * np.exp(rng.uniform(-8, 8, n))
    d[rng.random(n) < 0.3] = 0.0  # zero diagonals force off-diagonal matches
    A = (B + B.T + sp.diags(d)).tocsc()
    A.eliminate_zeros()
    A.sort_indices()
    M = rs.SparseColMat(n, A.indptr, A.indices, A.data)
    r = rs.mc64_matching(M)
    try:
        best = _scipy_best(A)
    except ValueError:  # structurally singular: no perfect matching exists
        assert r.matched < n
        return
    assert r.matched == n
    assert _log_product(A, r.matching) == pytest.approx(best, abs=1e-8 * max(1.0, abs(best)))
    _certificate(A, r)


@pytest.mark.parametrize("name,path", gen.suitesparse_ci(), ids=[c[0] for c in gen.suitesparse_ci()])
def test_suitesparse_ci_optimal(name, path):
    A = rs.load_mtx(path)
    S = sp.csc_matrix((A.values, A.row_idx, A.col_ptr), shape=(A.n, A.n))
    r = rs.mc64_matching(A)
    assert r.matched == A.n
    _certificate(S, r)
    assert _log_product(S, r.matching) == pytest.approx(_scipy_best(S), rel=1e-10, abs=1e-8)


def _check_order(M, res):
    n = M.n
    o = np.asarray(res.ordering)
    assert sorted(o.tolist()) == list(range(n))
    return o


@pytest.mark.parametrize("fwd,pairs,singles", [
    ([0, 1, 2, 3], 0, 4), ([1, 0, 3, 2], 2, 0), ([1, 2, 0], 1, 1), ([1, 2, 3, 0], 2, 0), ([1, 2, 3, 4, 0], 2, 1),
])
def test_cycle_split_counts(fwd, pairs, singles):
    # a matrix whose only entries are those of the permutation forces that matching (or its inverse)
    n = len(fwd)
    M, A = _sym(n, [(i, j, 1.0 + 0.1 * i) for i, j in enumerate(fwd)])
    res = rs.match_order_metis(M)
    assert (res.two_cycles, res.singletons, res.condensed_dim) == (pairs, singles, pairs + singles)
    o = _check_order(M, res)
    # each pair is a matching edge (j, sigma(j)) of a cycle and sits at two adjacent positions
    m = rs.mc64_matching(M).matching
    edge = lambda a, b: a != b and (m[a] == b or m[b] == a)
    found, p = 0, 0
    while p + 1 < n:
        if edge(o[p], o[p + 1]):
            found += 1
            p += 2
        else:
            p += 1
    assert found >= pairs


def test_match_order_edge_cases():
    r = rs.match_order_metis(rs.SparseColMat(0, [0], [], []))
    assert r.ordering.size == 0 and r.condensed_dim == 0
    M, _ = _sym(1, [(0, 0, 3.0)])
    r = rs.match_order_metis(M)
    assert list(r.ordering) == [0] and r.condensed_dim == 1 and r.scaling[0] == 1.0
    M, _ = _sym(4, [(i, i, 2.0 + i) for i in range(4)])
    r = rs.match_order_metis(M)
    assert r.matched == 4 and r.two_cycles == 0
    _check_order(M, r)


def test_kkt_pairs_and_analysis():
    A, _b, _x, info = workloads.kkt(16, seed=0)
    r = rs.match_order_metis(A)
    assert r.matched == A.n
    _check_order(A, r)
    s = rs.SparseLDLT.analyze_with_matrix(A, rs.AnalyzeOptions(rs.OrderingStrategy.MatchOrderMetis))
    st = s.structure()
    sc = s._scaling
    assert sc is not None and sc.shape == (A.n,)
    # elimination-order scaling is the MC64 scaling permuted by perm_fwd (solver.rs:276-279)
    np.testing.assert_array_equal(sc, r.scaling[np.asarray(st.perm_fwd)])
    with pytest.raises(rs.AnalysisFailure):
        rs.SparseLDLT.analyze(A, rs.AnalyzeOptions(rs.OrderingStrategy.MatchOrderMetis))
