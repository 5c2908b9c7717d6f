"""Pins of the CPU oracle that do not go through its own formulas:

  * sparse_backward_error (validate.rs:209-255) against an independent SciPy/NumPy evaluation of
    ||A x - b||_2 / (||A||_F ||x||_2 + ||b||_2), plus its closed forms;
  * the hand-constructed fixtures' L and D factors (sparse/test-data/hand-constructed/*.json,
    read as sparse/tests/aptp_data_structures.rs:42-79 reads them): for a fixed elimination order
    a block LDL^T factorization is unique up to the block partition, so wherever the oracle
    reproduces the fixture's permutation, every stretch between block boundaries common to both
    factorizations must give the same L[:, seg] D[seg, seg] L[seg, seg]^T -- blocks of equal kind
    compare D entry by entry;
  * the a-posteriori threshold bound |L| < 1/u of the APTP path (apply_and_check,
    factor.rs:1295-1408) on the reference's delay-causing LCG matrices (factor.rs:4886-4944), with
    columns actually delayed;
  * extend_add_mapped against the reference's known answers (numeric.rs:3470-3713).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import gen
import oracle
import rivrs_sparse as rs

EPS = np.finfo(np.float64).eps


# ---------------------------------------------------------------------------
# sparse_backward_error
# ---------------------------------------------------------------------------

def _sym_random(n, density, seed):
    rng = np.random.default_rng(seed)
    M = sp.random(n, n, density=density, random_state=rng, format="csc")
    M = (M + M.T + sp.diags(rng.standard_normal(n))).tocsc()
    M.sum_duplicates()
    M.sort_indices()
    return M


def _be_numpy(M, x, b):
    r = M @ x - b
    den = sp.linalg.norm(M, "fro") * np.linalg.norm(x) + np.linalg.norm(b)
    return 0.0 if den == 0 else np.linalg.norm(r) / den


@pytest.mark.parametrize("n,density,seed", [(1, 1.0, 0), (7, 0.5, 1), (50, 0.1, 2), (300, 0.02, 3), (2000, 0.002, 4)])
def test_backward_error_matches_numpy(n, density, seed):
    M = _sym_random(n, density, seed)
    A = rs.SparseColMat.from_scipy(M)
    rng = np.random.default_rng(seed + 100)
    for _ in range(3):
        x = rng.standard_normal(n)
        b = rng.standard_normal(n) * rng.uniform(0.1, 10.0)
        got = oracle.backward_error(A, x, b)
        ref = _be_numpy(M, x, b)
        assert abs(got - ref) <= 1e-13 * ref, (got, ref)
        # the library's own host-side evaluation (rivrs_sparse) agrees as well
        assert abs(rs.sparse_backward_error(A, x, b) - ref) <= 1e-13 * ref


def test_backward_error_closed_forms():
    n = 40
    d = np.linspace(-3.0, 5.0, n)
    d[d == 0] = 1.0
    A = rs.SparseColMat.from_scipy(sp.diags(d).tocsc())
    b = np.arange(1.0, n + 1.0)
    # exact solution of a diagonal system: residual is one rounding of b_i / d_i per entry
    x = b / d
    assert oracle.backward_error(A, x, b) <= 2 * EPS
    # x = 0: residual = -b, denominator = ||b||  =>  exactly 1
    assert oracle.backward_error(A, np.zeros(n), b) == 1.0
    # x = 0, b = 0: zero denominator is defined as 0 (validate.rs:250-252)
    assert oracle.backward_error(A, np.zeros(n), np.zeros(n)) == 0.0
    # invariant under scaling the matrix and right-hand side together, and x and b together
    M = _sym_random(60, 0.1, 9)
    rng = np.random.default_rng(9)
    x = rng.standard_normal(60)
    b = rng.standard_normal(60)
    A1 = rs.SparseColMat.from_scipy(M)
    A2 = rs.SparseColMat.from_scipy((4.0 * M).tocsc())
    e1 = oracle.backward_error(A1, x, b)
    assert abs(oracle.backward_error(A2, x, 4.0 * b) - e1) <= 1e-14 * e1
    assert abs(oracle.backward_error(A1, 8.0 * x, 8.0 * b) - e1) <= 1e-14 * e1
    # both triangles count towards ||A||_F: a single off-diagonal pair [[0, a], [a, 0]]
    P = rs.SparseColMat.from_scipy(sp.csc_matrix(np.array([[0.0, 3.0], [3.0, 0.0]])))
    assert abs(oracle.backward_error(P, np.array([1.0, 0.0]), np.zeros(2)) - 3.0 / np.sqrt(18.0)) <= 2 * EPS


# ---------------------------------------------------------------------------
# hand-constructed fixtures: L and D
# ---------------------------------------------------------------------------

def _fixture_factors(data, n):
    L = np.eye(n)
    for e in data["l_entries"]:
        L[e["row"], e["col"]] = e["value"]
    D = np.zeros((n, n))
    bounds = [0]
    kinds = []
    c = 0
    for blk in data["d_blocks"]:
        s = blk["size"]
        D[c:c + s, c:c + s] = np.array(blk["values"], dtype=float).reshape(s, s)
        kinds.append(s)
        c += s
        bounds.append(c)
    return np.array(data["permutation"]), L, D, bounds, kinds


def _block_inertia(blocks):
    """Sign counts of D from its blocks by the trace/determinant table of
    MixedDiagonal::compute_inertia (diagonal.rs:367-390), as the reference's fixture test
    computes them (aptp_data_structures.rs:42-79)."""
    pos = neg = zero = 0
    for blk in blocks:
        v = np.array(blk["values"], dtype=float).reshape(blk["size"], blk["size"])
        if blk["size"] == 1:
            d = v[0, 0]
            pos, neg, zero = pos + (d > 0), neg + (d < 0), zero + (d == 0)
            continue
        det = v[0, 0] * v[1, 1] - v[0, 1] * v[1, 0]
        tr = v[0, 0] + v[1, 1]
        if det < 0:
            pos, neg = pos + 1, neg + 1
        elif det > 0:
            pos, neg = (pos + 2, neg) if tr > 0 else (pos, neg + 2)
        else:
            if tr > 0:
                pos, zero = pos + 1, zero + 1
            elif tr < 0:
                neg, zero = neg + 1, zero + 1
            else:
                zero += 2
    return pos, neg, zero


def _oracle_bounds(dkind, ne):
    b, c = [0], 0
    while c < ne:
        c += 2 if dkind[c] == 2 else 1
        b.append(c)
    return b


def _segments_agree(Lf, Df, bf, Lo, Do, bo, scale):
    """Compare L[:, seg] D[seg, seg] L[seg, seg]^T over the stretches between common boundaries;
    returns the largest difference relative to scale."""
    common = sorted(set(bf) & set(bo))
    worst = 0.0
    for a, z in zip(common[:-1], common[1:]):
        Mf = Lf[a:, a:z] @ Df[a:z, a:z] @ Lf[a:z, a:z].T
        Mo = Lo[a:, a:z] @ Do[a:z, a:z] @ Lo[a:z, a:z].T
        worst = max(worst, np.abs(Mf - Mo).max() / scale)
    return worst, len(common) - 1


# fixtures whose permutation the complete-pivoting path (ib = 2, nb = 4) reproduces
APTP_SAME_ORDER = {"arrow-5-pd", "singular-3", "stress-delayed-pivots", "stress-fill-in", "stress-ill-cond",
                   "tridiag-10-indef", "tridiag-20-indef", "trivial-1x1", "trivial-2x2", "zero-diagonal-3"}


@pytest.mark.parametrize("name,path,data", gen.hand_constructed(), ids=[c[0] for c in gen.hand_constructed()])
def test_fixture_factors_match_oracle(name, path, data):
    A = rs.load_mtx(path)
    n = A.n
    Ad = A.to_scipy().toarray()
    P, Lf, Df, bf, kinds = _fixture_factors(data, n)
    scale = max(np.abs(Ad).max(), 1e-300)
    # the fixture itself: P^T A P = L D L^T, and the inertia of its D as stated
    assert np.abs(Ad[np.ix_(P, P)] - Lf @ Df @ Lf.T).max() <= 1e-12 * scale
    inert = data["inertia"]
    assert _block_inertia(data["d_blocks"]) == (inert["positive"], inert["negative"], inert["zero"])
    # threshold partial pivoting (the dense kernel's path for p < ib) keeps the natural order
    # on every fixture; complete pivoting (ib = 2) on the listed ones
    for label, opts in (("tpp", oracle.options()), ("aptp", oracle.options(ib=2, nb=4))):
        r = oracle.dense_factor(Ad, n, opts)
        ne = r["num_eliminated"]
        assert ne == n, (label, ne)
        same = list(r["perm"]) == list(P)
        if label == "tpp" or name in APTP_SAME_ORDER:
            assert same, (label, list(r["perm"]))
        Lo = gen.extract_l(r["a"], r["dkind"], ne)
        Do = gen.d_matrix(r["dkind"], r["ddiag"], r["doff"], ne)
        if not same:
            # another order: the factorization still reproduces A in that order
            pr = r["perm"]
            assert np.abs(Ad[np.ix_(pr, pr)] - Lo @ Do @ Lo.T).max() <= 1e-12 * scale
            continue
        worst, nseg = _segments_agree(Lf, Df, bf, Lo, Do, _oracle_bounds(r["dkind"], ne), scale)
        assert nseg >= 1
        assert worst <= 1e-12, (label, worst)
        # where the block kinds coincide, D is compared entry by entry
        if _oracle_bounds(r["dkind"], ne) == bf:
            assert np.abs(Do - Df).max() <= 1e-12 * scale, label
            assert np.abs(Lo - Lf).max() <= 1e-12 * max(1.0, np.abs(Lf).max()), label


def test_fixture_block_kinds_stress_cases():
    # stress-delayed-pivots: "All pivots must be 2x2" (zero diagonal); zero-diagonal-3 and
    # trivial-2x2 start with a 2x2 block -- the oracle's pivot kinds equal the fixture's there
    for name, path, data in gen.hand_constructed():
        if name not in ("stress-delayed-pivots", "zero-diagonal-3", "trivial-2x2"):
            continue
        A = rs.load_mtx(path)
        r = oracle.dense_factor(A.to_scipy().toarray(), A.n)
        _, _, _, bf, kinds = _fixture_factors(data, A.n)
        assert _oracle_bounds(r["dkind"], r["num_eliminated"]) == bf, name
        assert r["num_2x2"] == sum(1 for k in kinds if k == 2), name


# ---------------------------------------------------------------------------
# APTP a-posteriori threshold bound
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("n,seed,ib,nb", [(64, 1, 32, 256), (96, 2, 16, 64), (128, 3, 8, 32), (257, 4, 32, 256),
                                          (300, 5, 32, 128)])
def test_aptp_threshold_bound(n, seed, ib, nb):
    a = gen.cause_delays(gen.deterministic_indefinite_matrix_fast(n, seed, 0.1), seed, 1000.0)
    u = 0.01
    delayed_somewhere = False
    for tpp in (False, True):
        r = oracle.dense_factor(a, n, oracle.options(threshold=u, ib=ib, nb=nb, tpp=tpp))
        ne = r["num_eliminated"]
        assert r["num_1x1"] + 2 * r["num_2x2"] + r["num_delayed"] == n
        L = gen.extract_l(r["a"], r["dkind"], ne)[:, :ne]
        off = np.tril(L, -1)
        lmax = np.abs(off).max() if ne else 0.0
        if not tpp:
            # columns eliminated by complete pivoting + apply_and_check: strictly below 1/u
            assert lmax < 1.0 / u, lmax
            delayed_somewhere |= r["num_delayed"] > 0
        else:
            # the TPP fallback keeps |L| <= 1/u up to rounding (factor.rs:6919-6929)
            assert lmax <= (1.0 / u) * (1 + 1e-10), lmax
        assert abs(r["max_l"] - lmax) <= 1e-15 * max(lmax, 1.0) or r["max_l"] >= lmax
    assert delayed_somewhere, "the case must exercise the threshold test"


# ---------------------------------------------------------------------------
# extend-add known answers (numeric.rs:3470-3713)
# ---------------------------------------------------------------------------

def _naive_extend_add(P, child, mp):
    P = P.copy()
    sz = child.shape[0]
    for j in range(sz):
        for i in range(j, sz):
            li, lj = mp[i], mp[j]
            if li >= lj:
                P[li, lj] += child[i, j]
            else:
                P[lj, li] += child[i, j]
    return P


def test_extend_add_mapped_hand_constructed():
    # numeric.rs:3470-3536
    child = np.zeros((3, 3))
    child[0, 0], child[1, 0], child[1, 1], child[2, 0], child[2, 1], child[2, 2] = 1, 2, 3, 4, 5, 6
    p = oracle.extend_add_mapped(np.zeros((5, 5)), child, [2, 0, 3])
    expect = {(0, 0): 3.0, (2, 0): 2.0, (2, 2): 1.0, (3, 0): 5.0, (3, 2): 4.0, (3, 3): 6.0}
    for (i, j), v in expect.items():
        assert p[i, j] == v
    for (i, j) in ((1, 0), (1, 1), (4, 0), (4, 4)):
        assert p[i, j] == 0.0
    assert np.count_nonzero(p) == 6


def test_extend_add_mapped_10x10_and_non_monotonic():
    # numeric.rs:3538-3600 (monotonic map) and 3601-3662 (non-monotonic map)
    c10 = np.zeros((10, 10))
    for j in range(10):
        for i in range(j, 10):
            c10[i, j] = (i + 1) * 100.0 + (j + 1)
    mp = [1, 3, 4, 5, 7, 8, 10, 11, 12, 14]
    assert np.array_equal(oracle.extend_add_mapped(np.zeros((15, 15)), c10, mp),
                          _naive_extend_add(np.zeros((15, 15)), c10, mp))
    c4 = np.zeros((4, 4))
    for j in range(4):
        for i in range(j, 4):
            c4[i, j] = i * 10 + j + 1.0
    mp = [5, 2, 7, 1]
    got = oracle.extend_add_mapped(np.zeros((8, 8)), c4, mp)
    assert np.array_equal(got, _naive_extend_add(np.zeros((8, 8)), c4, mp))
    assert np.count_nonzero(np.triu(got, 1)) == 0  # only the lower triangle is written


def test_extend_add_mapped_accumulates():
    # numeric.rs:3663-3713
    c1 = np.array([[10.0, 0.0], [20.0, 30.0]])
    c2 = np.array([[5.0, 0.0], [7.0, 9.0]])
    p = oracle.extend_add_mapped(np.zeros((4, 4)), c1, [1, 3])
    p = oracle.extend_add_mapped(p, c2, [1, 3])
    assert (p[1, 1], p[3, 1], p[3, 3]) == (15.0, 27.0, 39.0)


def test_threads_do_not_change_results():
    # level-parallel waves and column-split products give the same bits as one thread
    import workloads
    A, b, info = workloads.helmholtz(12, seed=4)
    st = rs.SparseLDLT.analyze(A, rs.AnalyzeOptions(rs.OrderingStrategy.Grid, grid=info["grid"])).structure()
    nt = oracle.max_threads()
    try:
        oracle.set_threads(1)
        f1 = oracle.OracleFactor(A, st)
        x1 = f1.solve(b)
        oracle.set_threads(max(nt, 4))
        f2 = oracle.OracleFactor(A, st)
        x2 = f2.solve(b)
    finally:
        oracle.set_threads(nt)
    assert np.array_equal(x1, x2)
    for s in range(f1.num_supernodes()):
        a1, a2 = f1.front(s), f2.front(s)
        assert np.array_equal(a1["l21"], a2["l21"]) and np.array_equal(a1["ddiag"], a2["ddiag"])
