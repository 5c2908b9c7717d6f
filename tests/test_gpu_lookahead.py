"""Outer look-ahead: the trailing update of an outer block on a second stream beside the next
outer block, against the same factorization in stream order.

Each outer block's trailing update touches the columns after the next outer block only, and those
columns receive the blocks' updates in the same order either way, so the factors must be bitwise
equal to the stream-order ones:
  * dense two-front systems of several outer blocks (APTP on the first front), refactored on the
    same handle (the first factorization plans the second: blocks without failed columns in it run
    with look-ahead), with and without rows and columns scaled down to force failed pivots and
    delays (cause_delays, factor.rs:4914-4944);
  * a refactorization whose values fail pivots where the previous one failed none: the planned
    look-ahead meets failed columns, and the factorization is repeated in stream order;
  * Helmholtz and KKT systems (several APTP levels, FP64 exact-order, FP64 MFMA and FP32 outer
    updates).
"""
import os

import numpy as np
import pytest

import gen
import rivrs_sparse as rs
import workloads

pytestmark = pytest.mark.gpu

KEYS = ("dkind", "ddiag", "doff", "l11", "l21", "col_indices", "row_indices", "local_perm")


def say(msg):
    print(msg, flush=True)


def _handle(make, lookahead, chain=True):
    """A handle created with the outer look-ahead on or off, and the persistent outer-block
    launches on or off (both read when the numeric handle is made)."""
    env = {"RV_OUTER_LOOKAHEAD": "1" if lookahead else "0", "RV_APTP_CHAIN": "1" if chain else "0"}
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return make()
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


def _factor_twice(analyze, A1, A2, opts, lookahead, chain=True):
    def make():
        s = analyze()
        s.factor(A1, opts)
        return s
    s = _handle(make, lookahead, chain)
    s.refactor(A2, opts)
    return s


def _same_fronts(a, b):
    ns = a.num_supernodes()
    assert ns == b.num_supernodes()
    for q in range(ns):
        fa, fb = a.export_front(q), b.export_front(q)
        assert fa["num_eliminated"] == fb["num_eliminated"], q
        for key in KEYS:
            assert np.array_equal(fa[key], fb[key]), (q, key)


def _dense_two_front(a, p):
    m = a.shape[0]
    M = rs.SparseColMat(m, np.arange(0, m * m + 1, m, dtype=np.int64), np.tile(np.arange(m, dtype=np.int64), m),
                        np.asfortranarray(a).ravel(order="F"))
    st = rs.Structure(m, np.array([0, p], np.int32), np.array([p, m], np.int32),
                      np.array([0, m - p, m - p], np.int64), np.arange(p, m, dtype=np.int32),
                      np.array([1, -1], np.int32), np.arange(m, dtype=np.int32))
    return M, st


DENSE = [(1200, 1000, 0.02, None, "fp64"), (1500, 1300, 0.1, 1e-3, "fp64"), (1200, 1000, 0.02, None, "fp32"),
         (1500, 1300, 0.1, 1e-3, "fp32"), (1200, 1000, 0.02, None, "fp64_mfma"), (1500, 1300, 0.1, 1e-3, "fp64_mfma")]


@pytest.mark.parametrize("m,p,dscale,delay,prec", DENSE,
                         ids=["m%d-p%d-%s-%s" % (c[0], c[1], "delays" if c[3] else "plain", c[4]) for c in DENSE])
def test_dense_lookahead_bitwise(m, p, dscale, delay, prec):
    a = gen.deterministic_indefinite_matrix_fast(m, 5151 + m, dscale)
    if delay:
        a = gen.cause_delays(a, 91 + m, delay)
    M, st = _dense_two_front(a, p)
    opts = rs.FactorOptions(nemin=1, update_precision=prec)
    say("dense m=%d p=%d delays=%s %s" % (m, p, bool(delay), prec))
    analyze = lambda: rs.SparseLDLT.from_structure(st)
    sl = _factor_twice(analyze, M, M, opts, True)
    ss = _factor_twice(analyze, M, M, opts, False)
    say("  look-ahead launches %d" % sl.lookahead_launches())
    assert ss.lookahead_launches() == 0
    if not delay:
        assert sl.lookahead_launches() > 0
    _same_fronts(sl, ss)
    assert sl.inertia() == ss.inertia()


def test_lookahead_with_inner_block_launches():
    """The same overlap beside the three launches per inner block (persistent launches off)."""
    m, p = 1200, 1000
    M, st = _dense_two_front(gen.deterministic_indefinite_matrix_fast(m, 5151 + m, 0.02), p)
    opts = rs.FactorOptions(nemin=1)
    analyze = lambda: rs.SparseLDLT.from_structure(st)
    sl = _factor_twice(analyze, M, M, opts, True, chain=False)
    ss = _factor_twice(analyze, M, M, opts, False, chain=False)
    assert sl.chain_launches() == 0
    say("  look-ahead launches %d" % sl.lookahead_launches())
    assert sl.lookahead_launches() > 0
    _same_fronts(sl, ss)


def test_unplanned_failures_repeat_in_stream_order():
    """Refactor values that fail pivots after a factorization that failed none: the look-ahead
    planned from the first factorization meets failed columns; the result equals stream order."""
    m, p = 1500, 1300
    a1 = gen.deterministic_indefinite_matrix_fast(m, 777, 0.02)
    a2 = gen.cause_delays(gen.deterministic_indefinite_matrix_fast(m, 778, 0.1), 79, 1e-3)
    M1, st = _dense_two_front(a1, p)
    M2, _ = _dense_two_front(a2, p)
    opts = rs.FactorOptions(nemin=1)
    analyze = lambda: rs.SparseLDLT.from_structure(st)
    def make(M):
        s = analyze()
        s.factor(M, opts)
        return s
    s1 = _handle(lambda: make(M1), True)
    assert s1.export_front(0)["num_eliminated"] == p
    s1.refactor(M1, opts)
    n_la = s1.lookahead_launches()
    assert n_la > 0
    s1.refactor(M2, opts)
    g = s1.export_front(0)
    say("  eliminated %d of %d after the failing refactor" % (g["num_eliminated"], p))
    ref = _handle(lambda: make(M2), False)
    _same_fronts(s1, ref)
    # the next refactorization of the same values plans from the failures just seen
    s1.refactor(M2, opts)
    _same_fronts(s1, ref)


def _grid(info):
    return rs.AnalyzeOptions(rs.OrderingStrategy.Grid, grid=info["grid"])


@pytest.mark.parametrize("case", ["helmholtz28-fp64", "helmholtz28-fp32", "helmholtz28-fp64_mfma", "kkt48-fp64"])
def test_sparse_lookahead_bitwise(case):
    name, prec = case.split("-")
    if name.startswith("helmholtz"):
        A, b, info = workloads.helmholtz(int(name[9:]), seed=4)
        analyze = lambda: rs.SparseLDLT.analyze(A, _grid(info))
    else:
        A, b, _x, info = workloads.kkt(int(name[3:]), seed=6)
        analyze = lambda: rs.SparseLDLT.analyze_with_matrix(A, rs.AnalyzeOptions(rs.OrderingStrategy.MatchOrderMetis))
    opts = rs.FactorOptions(update_precision=prec)
    say("sparse %s n=%d" % (case, A.n))
    sl = _factor_twice(analyze, A, A, opts, True)
    ss = _factor_twice(analyze, A, A, opts, False)
    say("  look-ahead launches %d" % sl.lookahead_launches())
    assert ss.lookahead_launches() == 0
    assert sl.inertia() == ss.inertia()
    _same_fronts(sl, ss)
    x = sl.solve(b)
    if prec in ("fp64", "fp64_mfma"):
        assert rs.sparse_backward_error(A, x, b) < 1e-12
    if name.startswith("helmholtz"):
        # the closed-form count of negative eigenvalues (every update arithmetic keeps it)
        assert sl.inertia().negative == info["expected_negative"]


def _with_env(env, make):
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return make()
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


T64 = [("dense", "fp32", True), ("dense", "fp64_mfma", True), ("dense", "fp32", False),
       ("helmholtz28", "fp32", True), ("helmholtz28", "fp64_mfma", False)]


@pytest.mark.parametrize("case,prec,la", T64, ids=["%s-%s-%s" % (c[0], c[1], "la" if c[2] else "nola") for c in T64])
def test_outer_tile64_bitwise(case, prec, la):
    """The next outer block's columns of MFMA outer updates in 64 x 64 tiles (default) against
    128 x 128 tiles (RV_OUTER_TILE64=0): every entry is summed in the same order, so the factors
    are bitwise equal, with the outer look-ahead on and off."""
    opts = rs.FactorOptions(nemin=1 if case == "dense" else 32, update_precision=prec)
    if case == "dense":
        m, p = 1500, 1300
        M, st = _dense_two_front(gen.cause_delays(gen.deterministic_indefinite_matrix_fast(m, 6262, 0.1), 17, 1e-3), p)
        analyze = lambda: rs.SparseLDLT.from_structure(st)
    else:
        M, b, info = workloads.helmholtz(28, seed=8)
        analyze = lambda: rs.SparseLDLT.analyze(M, _grid(info))

    def make():
        s = analyze()
        s.factor(M, opts)
        s.refactor(M, opts)
        return s
    base = {"RV_OUTER_LOOKAHEAD": "1" if la else "0"}
    s64 = _with_env(dict(base, RV_OUTER_TILE64="1"), make)
    s128 = _with_env(dict(base, RV_OUTER_TILE64="0"), make)
    say("%s %s look-ahead launches %d / %d" % (case, prec, s64.lookahead_launches(), s128.lookahead_launches()))
    _same_fronts(s64, s128)
    assert s64.inertia() == s128.inertia()
