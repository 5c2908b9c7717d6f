"""Persistent outer-block launches (k_aptp_outer) against the three-launch inner-block path and
the oracle.

The persistent launch reorders when the work of an outer block happens (the next block is
factored ahead by a leader workgroup while panel workgroups apply the current one) but not the
arithmetic of any entry, so its factors must equal the three-launch path's bit for bit:
  * dense two-front systems (APTP on the first front, every outer block in one launch), with and
    without rows and columns scaled down to force failed pivots and delays (cause_delays,
    factor.rs:4914-4944): L11, L21, D, pivot kinds and permutation bitwise equal to the
    oracle's and to the three-launch path's;
  * Helmholtz and KKT systems (several APTP levels, 2x2 pivots, delays): every front bitwise
    equal between the two paths, the same inertia and the same delayed columns;
  * the counter of persistent launches shows which path ran.
"""
import os

import numpy as np
import pytest

import gen
import oracle
import rivrs_sparse as rs
import workloads

pytestmark = pytest.mark.gpu

KEYS = ("dkind", "ddiag", "doff", "l11", "l21", "col_indices", "row_indices", "local_perm")


def say(msg):
    print(msg, flush=True)


def _factor(A, analyze_or_structure, opts, chain):
    """A fresh handle factored with the persistent launches on or off (the switch is read when
    the numeric handle is created)."""
    old = os.environ.get("RV_APTP_CHAIN")
    os.environ["RV_APTP_CHAIN"] = "1" if chain else "0"
    try:
        s = analyze_or_structure()
        s.factor(A, opts)
    finally:
        if old is None:
            del os.environ["RV_APTP_CHAIN"]
        else:
            os.environ["RV_APTP_CHAIN"] = old
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


DENSE = [(300, 260, 0.1, None), (700, 600, 0.1, None), (900, 640, 0.5, 1e-3), (1200, 1000, 0.02, 1e-4)]


@pytest.mark.parametrize("m,p,dscale,delay", DENSE, ids=["m%d-p%d-%s" % (c[0], c[1], "delays" if c[3] else "plain")
                                                          for c in DENSE])
def test_dense_front_chain_bitwise(m, p, dscale, delay):
    a = gen.deterministic_indefinite_matrix_fast(m, 4242 + m, dscale)
    if delay:
        a = gen.cause_delays(a, 77 + m, delay)
    M, st = _dense_two_front(a, p)
    opts = rs.FactorOptions(nemin=1)
    say("dense m=%d p=%d delays=%s" % (m, p, bool(delay)))
    sc = _factor(M, lambda: rs.SparseLDLT.from_structure(st), opts, True)
    sp = _factor(M, lambda: rs.SparseLDLT.from_structure(st), opts, False)
    assert sc.chain_launches() > 0
    assert sp.chain_launches() == 0
    ref = oracle.OracleFactor(M, st, opts=oracle.options(nemin=1))
    g, r = sc.export_front(0), ref.front(0)
    say("  eliminated %d of %d, 2x2 pivots %d" % (g["num_eliminated"], p, int(np.sum(g["dkind"] == 2))))
    for key in ("dkind", "ddiag", "doff", "l11", "l21", "row_indices", "local_perm"):
        assert np.array_equal(g[key], r[key]), key
    _same_fronts(sc, sp)
    if delay:
        assert g["num_eliminated"] < p or np.any(g["dkind"] == 2)


def _grid(info):
    return rs.AnalyzeOptions(rs.OrderingStrategy.Grid, grid=info["grid"])


@pytest.mark.parametrize("case", ["helmholtz20", "helmholtz28", "kkt40"])
def test_sparse_chain_bitwise(case):
    if case.startswith("helmholtz"):
        A, b, info = workloads.helmholtz(int(case[9:]), seed=3)
        analyze = lambda: rs.SparseLDLT.analyze(A, _grid(info))
    else:
        A, b, _x, info = workloads.kkt(int(case[3:]), seed=5)
        analyze = lambda: rs.SparseLDLT.analyze_with_matrix(A, rs.AnalyzeOptions(rs.OrderingStrategy.MatchOrderMetis))
    opts = rs.FactorOptions()
    say("sparse %s n=%d" % (case, A.n))
    sc = _factor(A, analyze, opts, True)
    sp = _factor(A, analyze, opts, False)
    say("  persistent launches %d" % sc.chain_launches())
    assert sc.chain_launches() > 0 and sp.chain_launches() == 0
    assert sc.inertia() == sp.inertia()
    _same_fronts(sc, sp)
    x = sc.solve(b)
    assert rs.sparse_backward_error(A, x, b) < 1e-12


def test_refactor_reuses_chain_plans():
    """A second factorization of new values on the same handle (plans reused) matches a fresh
    three-launch handle."""
    A, b, info = workloads.helmholtz(20, seed=11)
    A2, _, _ = workloads.helmholtz(20, seed=12)
    analyze = lambda: rs.SparseLDLT.analyze(A, _grid(info))
    opts = rs.FactorOptions()
    sc = _factor(A, analyze, opts, True)
    n0 = sc.chain_launches()
    sc.refactor(A2, opts)
    assert sc.chain_launches() > n0
    sp = _factor(A2, analyze, opts, False)
    _same_fronts(sc, sp)
