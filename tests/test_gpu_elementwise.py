"""Element-wise GPU parity of the default FP64 path (contribution GEMM on FP64 MFMA), the plain
FP64 drop-in at the 40^3 Helmholtz size, and regressions of the C ABI's ownership rules.

Bars:
  * contribution blocks (ContributionBlock, numeric.rs:516-523, after compute_contribution_gemm,
    factor.rs:1590-1652): the pivot decisions and L21, D of a front are bitwise those of the
    oracle, so the two contribution blocks differ only by the summation order of
    C - sum_c W[:, c] L21[:, c]^T.  Each entry obeys the classical bound of two K-term sums,
    |dC| <= 2 gamma_K (|W| |L21|^T) + 2 eps |C|  (gamma_K = K eps / (1 - K eps)), taken here
    with a further factor 2 for the MFMA's partial sums;
  * multifrontal factors: every front's pivot kinds, row/column indices and permutation equal
    the oracle's; L11, L21 within 1e-12 sqrt(m) max(1, max|L|) and D within 1e-12 sqrt(m) max|D|
    of the front, m the order of the largest front of the factorization (summation-order
    differences of the contribution GEMMs reach every front through its descendants, and a
    K-term sum's rounding error grows as sqrt(K) on average);
  * Helmholtz 40^3 (n = 64 000): inertia, delayed count, per-supernode (m, k, ne) identical,
    backward error <= 1e-14, solution within 1e-9 of the oracle's.
"""
import ctypes as C

import numpy as np
import pytest

import batched
import gen
import oracle
import rivrs_sparse as rs
import workloads

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps


def say(msg):
    print(msg, flush=True)


def _dense_two_front(a, p):
    """Two supernodes over a dense m x m matrix: front 0 owns columns [0, p) with pattern
    [p, m) (its contribution is the (m-p)^2 Schur complement), front 1 is the root."""
    m = a.shape[0]
    M = rs.SparseColMat(m, np.arange(0, m * m + 1, m, dtype=np.int64), np.tile(np.arange(m, dtype=np.int64), m),
                        np.asfortranarray(a).ravel(order="F"))
    st = rs.Structure(m, np.array([0, p], np.int32), np.array([p, m], np.int32),
                      np.array([0, m - p, m - p], np.int64), np.arange(p, m, dtype=np.int32),
                      np.array([1, -1], np.int32), np.arange(m, dtype=np.int32))
    return M, st


def _d_block_abs(dkind, ddiag, doff):
    ne = len(dkind)
    D = np.zeros((ne, ne))
    c = 0
    while c < ne:
        if dkind[c] == 2:
            D[c, c], D[c + 1, c + 1] = ddiag[c], ddiag[c + 1]
            D[c, c + 1] = D[c + 1, c] = doff[c]
            c += 2
        else:
            D[c, c] = ddiag[c]
            c += 1
    return D


SCHUR_CASES = [(m, f) for m in (256, 1024, 2048, 4096) for f in (0.25, 0.5, 0.75)]


@pytest.mark.parametrize("m,frac", SCHUR_CASES, ids=["m%d-nfs%.2f" % c for c in SCHUR_CASES])
def test_contribution_gemm_elementwise(m, frac):
    nfs = int(m * frac)
    p = m - nfs
    a = gen.deterministic_indefinite_matrix_fast(m, 1000 + m + nfs, 0.1)
    M, st = _dense_two_front(a, p)
    say("contribution gemm m=%d p=%d nfs=%d" % (m, p, nfs))
    opts = rs.FactorOptions(nemin=1)
    s = rs.SparseLDLT.from_structure(st)
    s.factor(M, opts)
    ref = oracle.OracleFactor(M, st, opts=oracle.options(nemin=1), keep_contributions=True)
    g_f, r_f = s.export_front(0), ref.front(0)
    # the APTP part is bitwise the oracle's
    for key in ("dkind", "ddiag", "doff", "l21", "l11", "row_indices", "local_perm"):
        assert np.array_equal(g_f[key], r_f[key]), key
    cg = s.export_contribution(0)
    cr = ref.contribution(0)
    assert cg is not None and cr is not None
    Cg, rows_g, nd_g = cg
    Cr, rows_r, nd_r = cr
    assert nd_g == nd_r and np.array_equal(rows_g, rows_r)
    size = Cr.shape[0]
    low = np.tril(np.ones((size, size), dtype=bool))
    # delayed rows/columns are copies: bitwise
    if nd_r:
        dl = low.copy()
        dl[nd_r:, nd_r:] = False
        assert np.array_equal(Cg[dl], Cr[dl])
    ne = r_f["num_eliminated"]
    L = r_f["l21"][nd_r:, :]  # NFS rows of L21
    W = L @ _d_block_abs(r_f["dkind"], r_f["ddiag"], r_f["doff"])
    bound_terms = np.abs(W) @ np.abs(L).T
    gamma = ne * EPS / (1 - ne * EPS)
    Cn_g, Cn_r = Cg[nd_r:, nd_r:], Cr[nd_r:, nd_r:]
    lown = np.tril(np.ones((size - nd_r, size - nd_r), dtype=bool))
    diff = np.abs(Cn_g - Cn_r)[lown]
    bound = (4 * gamma * bound_terms + 4 * EPS * np.abs(Cn_r))[lown]
    worst = (diff / np.maximum(bound, 1e-300)).max() if diff.size else 0.0
    say("  max |dC| %.3e, largest fraction of the bound %.3f, identical entries %.1f%%"
        % (diff.max() if diff.size else 0.0, worst, 100.0 * np.mean(diff == 0) if diff.size else 100.0))
    assert np.all(diff <= bound), worst
    # the strictly upper triangle is not part of the block and the lower one is fully written
    assert np.all(np.isfinite(Cg[low]))


def _front_close(g, r, q):
    for key in ("dkind", "col_indices", "row_indices", "local_perm"):
        assert np.array_equal(g[key], r[key]), (q, key)
    lscale = max(1.0, np.abs(r["l21"]).max() if r["l21"].size else 0.0, np.abs(r["l11"]).max() if r["l11"].size else 0.0)
    dscale = max(np.abs(r["ddiag"]).max() if r["ddiag"].size else 0.0, np.abs(r["doff"]).max() if r["doff"].size else 0.0,
                 1e-300)
    dl = max(np.abs(g["l11"] - r["l11"]).max() if r["l11"].size else 0.0,
             np.abs(g["l21"] - r["l21"]).max() if r["l21"].size else 0.0) / lscale
    dd = max(np.abs(g["ddiag"] - r["ddiag"]).max() if r["ddiag"].size else 0.0,
             np.abs(g["doff"] - r["doff"]).max() if r["doff"].size else 0.0) / dscale
    return dl, dd


def _factor_tol(ref):
    m = max(ref.supernode(q)["front_size"] for q in range(ref.num_supernodes()))
    return 1e-12 * np.sqrt(m)


def _structure(case):
    if case.startswith("helmholtz"):
        N = int(case[len("helmholtz"):])
        A, b, info = workloads.helmholtz(N, seed=1)
        st = rs.SparseLDLT.analyze(A, rs.AnalyzeOptions(rs.OrderingStrategy.Grid, grid=info["grid"])).structure()
        return A, b, st, None
    side = int(case[len("kkt"):])
    A, b, x, info = workloads.kkt(side, seed=1)
    s = rs.SparseLDLT.analyze_with_matrix(A, rs.AnalyzeOptions(rs.OrderingStrategy.MatchOrderMetis))
    return A, b, s.structure(), s._scaling


@pytest.mark.parametrize("case", ["helmholtz12", "helmholtz20", "kkt32"])
def test_multifrontal_factors_elementwise(case):
    # default options: contribution GEMM on FP64 MFMA (not the reference's summation order)
    A, b, st, scaling = _structure(case)
    say("elementwise factors %s n=%d" % (case, A.n))
    s = rs.SparseLDLT.from_structure(st, scaling)
    s.factor(A)
    ref = oracle.OracleFactor(A, st, scaling)
    worst_l = worst_d = 0.0
    for q in range(ref.num_supernodes()):
        dl, dd = _front_close(s.export_front(q), ref.front(q), q)
        worst_l, worst_d = max(worst_l, dl), max(worst_d, dd)
    tol = _factor_tol(ref)
    say("  largest relative deviation: L %.2e, D %.2e (bar %.2e)" % (worst_l, worst_d, tol))
    assert worst_l <= tol and worst_d <= tol, (worst_l, worst_d, tol)


def test_helmholtz40_matches_oracle():
    # the plain FP64 drop-in configuration: 40^3 indefinite Helmholtz (n = 64 000), sigma at the
    # 5 % spectral gap, +-1 % jitter, one N(0,1) right-hand side
    A, b, info = workloads.helmholtz(40, seed=0)
    st = rs.SparseLDLT.analyze(A, rs.AnalyzeOptions(rs.OrderingStrategy.Grid, grid=info["grid"])).structure()
    say("helmholtz40 n=%d: oracle" % A.n)
    ref = oracle.OracleFactor(A, st)
    say("helmholtz40: gpu")
    s = rs.SparseLDLT.from_structure(st)
    s.factor(A)
    inert = s.inertia()
    assert (inert.positive, inert.negative, inert.zero) == ref.inertia()
    assert inert.negative == info["expected_negative"]
    gs, rst = s.stats(), ref.stats()
    assert gs.total_delayed == rst["total_delayed"]
    assert (gs.total_1x1_pivots, gs.total_2x2_pivots, gs.zero_pivots) == (
        rst["total_1x1_pivots"], rst["total_2x2_pivots"], rst["zero_pivots"])
    per = s.per_supernode_stats()
    assert len(per) == ref.num_supernodes()
    for q in per:
        r = ref.supernode(q.snode_id)
        assert (q.front_size, q.num_fully_summed, q.num_eliminated, q.num_delayed) == (
            r["front_size"], r["num_fully_summed"], r["num_eliminated"], r["num_delayed"]), q.snode_id
    x = s.solve(b)
    xr = ref.solve(b)
    be = rs.sparse_backward_error(A, x, b)
    say("  inertia %s, delayed %d, backward error %.2e" % (ref.inertia(), rst["total_delayed"], be))
    assert be <= 1e-14, be
    assert np.abs(x - xr).max() <= 1e-9 * np.abs(xr).max()
    # element-wise factors on the largest fronts and a sample of the rest
    order = sorted(range(len(per)), key=lambda q: -per[q].front_size)
    rng = np.random.default_rng(0)
    sample = order[:6] + list(rng.choice(order[6:], size=24, replace=False))
    worst_l = worst_d = 0.0
    for q in sample:
        dl, dd = _front_close(s.export_front(int(q)), ref.front(int(q)), q)
        worst_l, worst_d = max(worst_l, dl), max(worst_d, dd)
    tol = _factor_tol(ref)
    say("  sampled fronts: largest relative deviation L %.2e, D %.2e (bar %.2e)" % (worst_l, worst_d, tol))
    assert worst_l <= tol and worst_d <= tol, (worst_l, worst_d, tol)


# ---------------------------------------------------------------------------
# C ABI ownership regressions
# ---------------------------------------------------------------------------

def test_refactor_device_refines_against_new_values():
    # factor A1 from host memory, refactor from device memory with A2's values: refinement must
    # use A2 (the residual of the values last factored)
    A1, b, info = workloads.helmholtz(10, seed=1)
    A2, _, _ = workloads.helmholtz(10, seed=2)  # same pattern, other jitter
    assert np.array_equal(A1.row_idx, A2.row_idx)
    s = rs.SparseLDLT.analyze(A1, rs.AnalyzeOptions(rs.OrderingStrategy.Grid, grid=info["grid"]))
    s.factor(A1)
    lib = rs.load_library()
    d2 = batched.DeviceArray(A2.values)
    say("refactor_device helmholtz10")
    rs._check(lib.rv_refactor_device(s._num, d2.ptr))
    x, it, be = s.solve_refine(b, tol=1e-12, max_iter=3)
    be_np = rs.sparse_backward_error(A2, x, b)
    assert be_np <= 1e-12, be_np
    assert abs(be - be_np) <= 1e-3 * be_np + 1e-300, (be, be_np)
    # and a device-owned factorization adopts the new pointer
    d1 = batched.DeviceArray(A1.values)
    h = C.c_void_p()
    o = rs.FactorOptions().to_c()
    st = s.structure()
    s2 = rs.SparseLDLT.from_structure(st)
    rs._check(lib.rv_factor_device(s2._sym, rs._ptr(A1.col_ptr, C.c_int64), rs._ptr(A1.row_idx, C.c_int64),
                                   d1.ptr, None, C.byref(o), C.byref(h)))
    s2._num, s2._pattern, s2._options = h, A1, rs.FactorOptions()
    rs._check(lib.rv_refactor_device(h, d2.ptr))
    x2, it2, be2 = s2.solve_refine(b, tol=1e-12, max_iter=3)
    assert rs.sparse_backward_error(A2, x2, b) <= 1e-12


def test_numeric_handles_outlive_symbolic_changes():
    # two numeric handles on one symbolic handle with different nemin: the second factorization
    # replaces the symbolic handle's prepared structure; the first handle keeps its own
    A, b, info = workloads.helmholtz(12, seed=3)
    s = rs.SparseLDLT.analyze(A, rs.AnalyzeOptions(rs.OrderingStrategy.Grid, grid=info["grid"]))
    lib = rs.load_library()
    hs = []
    for nemin in (32, 8):
        h = C.c_void_p()
        o = rs.FactorOptions(nemin=nemin).to_c()
        rs._check(lib.rv_factor(s._sym, rs._ptr(A.col_ptr, C.c_int64), rs._ptr(A.row_idx, C.c_int64),
                                rs._ptr(A.values, C.c_double), None, C.byref(o), C.byref(h)))
        hs.append(h)
    assert lib.rv_num_supernodes(hs[0]) != lib.rv_num_supernodes(hs[1])
    say("two numerics on one symbolic")
    for h in hs:
        x = np.array(b, copy=True)
        rs._check(lib.rv_solve(h, rs._ptr(x, C.c_double), A.n, 1))
        assert rs.sparse_backward_error(A, x, b) <= 1e-14
        out = np.zeros(2, np.int64)
        rs._check(lib.rv_supernode_structure(h, 0, None, None, rs._ptr(out, C.c_int64)))
    # the symbolic handle can go first
    lib.rv_symbolic_free(s._sym)
    s._sym = None
    x = np.array(b, copy=True)
    rs._check(lib.rv_solve(hs[0], rs._ptr(x, C.c_double), A.n, 1))
    assert rs.sparse_backward_error(A, x, b) <= 1e-14
    for h in hs:
        lib.rv_numeric_free(h)


def test_solve_scratch_known_before_first_solve():
    A, b, info = workloads.helmholtz(10, seed=0)
    s = rs.SparseLDLT.analyze(A, rs.AnalyzeOptions(rs.OrderingStrategy.Grid, grid=info["grid"]))
    s.factor(A)
    before = [s.solve_scratch(k) for k in (1, 4)]
    s.solve(b)
    after = [s.solve_scratch(k) for k in (1, 4)]
    assert before == after and before[0] > 8 * A.n
