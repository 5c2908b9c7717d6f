"""Working fronts in the level pool: a front's memory is handed to later levels once its parent's
level has assembled it, and its eliminated columns are kept apart for the solve.

Against the same factorizations with every front on memory of its own (RV_KEEP_FRONTS=1, read
when the numeric handle is made):
  * the factors, inertia and solution are bitwise equal, on a Helmholtz system refactored with
    values of another shift (delays change between the two factorizations, so the second runs the
    per-level pass over the levels after the first changed one, which reads contributions that
    the first pass's later levels may have overwritten) and on a KKT system;
  * the device memory of a 48^3 system is less than 0.7 of the kept fronts';
  * a contribution block whose memory a later level took is reported as unavailable, and every
    one still held equals the kept copy;
  * a front's contribution block (FP64 and FP32 Schur updates) is within the GEMM's error bound
    of the Schur complement formed from its exported factors.
"""
import os

import numpy as np
import pytest

import rivrs_sparse as rs
import workloads

pytestmark = pytest.mark.gpu

KEYS = ("dkind", "ddiag", "doff", "l11", "l21", "col_indices", "row_indices", "local_perm")


def say(msg):
    print(msg, flush=True)


def _make(make, keep):
    old = os.environ.get("RV_KEEP_FRONTS")
    os.environ["RV_KEEP_FRONTS"] = "1" if keep else "0"
    try:
        return make()
    finally:
        if old is None:
            del os.environ["RV_KEEP_FRONTS"]
        else:
            os.environ["RV_KEEP_FRONTS"] = old


def _same(a, b):
    ns = a.num_supernodes()
    assert ns == b.num_supernodes()
    for q in range(ns):
        fa, fb = a.export_front(q), b.export_front(q)
        assert fa["num_eliminated"] == fb["num_eliminated"], q
        for key in KEYS:
            assert np.array_equal(fa[key], fb[key]), (q, key)
    assert a.inertia() == b.inertia()


def test_helmholtz_refactor_with_changed_delays():
    g = 28
    A1, b1, info = workloads.helmholtz(g, seed=3, frac=0.05)
    A2, b2, _ = workloads.helmholtz(g, seed=4, frac=0.12)
    opts = rs.FactorOptions()

    def make():
        s = rs.SparseLDLT.analyze(A1, rs.AnalyzeOptions(rs.OrderingStrategy.Grid, grid=info["grid"]))
        s.factor(A1, opts)
        return s

    sp, sk = _make(make, False), _make(make, True)
    d1 = [p.num_delayed for p in sp.per_supernode_stats()]
    _same(sp, sk)
    sp.refactor(A2, opts)
    sk.refactor(A2, opts)
    d2 = [p.num_delayed for p in sp.per_supernode_stats()]
    say("supernodes %d, delays changed on %d; device %.3f GB pooled, %.3f GB kept"
        % (len(d1), sum(x != y for x, y in zip(d1, d2)), sp.device_bytes() / 1e9, sk.device_bytes() / 1e9))
    assert d1 != d2
    _same(sp, sk)
    xp, xk = sp.solve(b2), sk.solve(b2)
    assert np.array_equal(xp, xk)
    assert rs.sparse_backward_error(A2, xp, b2) < 1e-12
    # contribution blocks: unavailable once taken by a later level, else equal to the kept copy
    held = taken = 0
    for q in range(min(sp.num_supernodes(), 300)):
        ck = sk.export_contribution(q)
        try:
            cp = sp.export_contribution(q)
        except rs.SparseError:
            taken += 1
            continue
        held += 1
        if ck is None:
            assert cp is None
            continue
        assert np.array_equal(np.tril(cp[0]), np.tril(ck[0])) and np.array_equal(cp[1], ck[1]) and cp[2] == ck[2]
    say("contribution blocks: %d held, %d taken by later levels" % (held, taken))
    assert taken > 0


def test_kkt_pooled_equals_kept():
    A, b, _x, info = workloads.kkt(48, seed=11)
    opts = rs.FactorOptions()

    def make():
        s = rs.SparseLDLT.analyze_with_matrix(A, rs.AnalyzeOptions(rs.OrderingStrategy.MatchOrderMetis))
        s.factor(A, opts)
        return s

    sp, sk = _make(make, False), _make(make, True)
    _same(sp, sk)
    sp.refactor(A, opts)
    _same(sp, sk)
    assert np.array_equal(sp.solve(b), sk.solve(b))


def test_device_memory_helmholtz48():
    """Fronts of a 48^3 system: 2.4 GB as full squares, about 1.1 GB as kept columns plus the
    fronts alive at once (device memory comes in chunks of at least 256 MiB)."""
    A, b, info = workloads.helmholtz(48, seed=5)
    opts = rs.FactorOptions()

    def make():
        s = rs.SparseLDLT.analyze(A, rs.AnalyzeOptions(rs.OrderingStrategy.Grid, grid=info["grid"]))
        s.factor(A, opts)
        return s

    sp, sk = _make(make, False), _make(make, True)
    say("device %.3f GB pooled, %.3f GB kept" % (sp.device_bytes() / 1e9, sk.device_bytes() / 1e9))
    assert sp.inertia() == sk.inertia()
    assert np.array_equal(sp.solve(b), sk.solve(b))
    assert sp.device_bytes() < 0.7 * sk.device_bytes()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_contribution_block_vs_factors(prec):
    """The contribution block equals -L21 D L21^T (the front holds only its own columns'
    entries) from the exported factors within the error bound of an FP32 GEMM (FP32 Schur
    updates; FP64 ones lie far inside it)."""
    import gen
    m, p = 1500, 1100
    a = gen.deterministic_indefinite_matrix_fast(m, 4242, 0.02)
    M = rs.SparseColMat(m, np.arange(0, m * m + 1, m, dtype=np.int64), np.tile(np.arange(m, dtype=np.int64), m),
                        np.asfortranarray(a).ravel(order="F"))
    st = rs.Structure(m, np.array([0, p], np.int32), np.array([p, m], np.int32),
                      np.array([0, m - p, m - p], np.int64), np.arange(p, m, dtype=np.int32),
                      np.array([1, -1], np.int32), np.arange(m, dtype=np.int32))
    s = rs.SparseLDLT.from_structure(st)
    s.factor(M, rs.FactorOptions(nemin=1, update_precision=prec))
    f = s.export_front(0)
    ne = f["num_eliminated"]
    C, rows, nd = s.export_contribution(0)
    # delayed columns (if any) lead the block and stay FP64; the pattern rows follow
    assert nd == p - ne and np.array_equal(rows[nd:], np.arange(p, m))
    C = C[nd:, nd:]
    L = f["l21"][nd:, :]
    D = np.zeros((ne, ne))
    c = 0
    while c < ne:
        if f["dkind"][c] == 2:
            D[c, c], D[c + 1, c + 1] = f["ddiag"][c], f["ddiag"][c + 1]
            D[c, c + 1] = D[c + 1, c] = f["doff"][c]
            c += 2
        else:
            D[c, c] = f["ddiag"][c]
            c += 1
    # front 0 assembles only its own columns' entries: its NFS block starts from zero
    ref = -(L @ D @ L.T)
    eps32 = np.finfo(np.float32).eps
    bound = 4 * ne * eps32 * (np.abs(L) @ np.abs(D) @ np.abs(L).T) + 2 * eps32 * np.abs(ref)
    low = np.tril(np.ones(ref.shape, dtype=bool))
    diff = np.abs(C - ref)[low]
    say("%s contribution (%d delayed): max |dC| %.3e, largest fraction of the bound %.3f"
        % (prec, nd, diff.max(), (diff / bound[low]).max()))
    full = np.where(low, np.abs(C - ref), 0.0)
    bad = np.argwhere(full > np.where(low, bound, np.inf))
    if len(bad):
        say("  %d entries over the bound, first %s: gpu %r ref %r" % (len(bad), bad[:5].tolist(), C[tuple(bad[0])], ref[tuple(bad[0])]))
    assert np.all(diff <= bound[low])
    # the same system in FP64: the solve through the FP32 extend-add refines to the FP64 one
    x, it, be = s.solve_refine(a @ np.ones(m), tol=1e-12, max_iter=10)
    assert be <= 1e-12, (it, be)
