"""The precision and learned configurations at their own shapes, on the GPU through the C ABI.

* learned per-front precision on a held-out 64^3 indefinite Helmholtz system (n = 262 144, a seed
  the network never saw): refinement reaches backward error <= 1e-12 within 3 steps without the
  FP64 fallback, the inertia equals the closed-form count of negative eigenvalues of the
  unjittered operator (sigma sits in a spectral gap, the +-1 % jitter keeps every eigenvalue on
  its side), and the forward error against the FP64 solution is reported;
* the same for uniform FP32 updates (the mixed-precision configuration's FP32 form), with its
  refinement steps and forward error;
* learned precision on a held-out 128 x 128-grid KKT system (n = 20 480) against the oracle:
  identical inertia (also the Haynsworth count), backward error <= 1e-12.
"""
import os
import sys

import numpy as np
import pytest

import oracle
import precision as pr
import rivrs_sparse as rs
import workloads

pytestmark = pytest.mark.gpu
MODEL = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "rivrs-linalg_amd", "models",
                     "precision_mlp.json")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def helmholtz64():
    A, b, info = workloads.helmholtz(64, seed=17)
    s = rs.SparseLDLT.analyze(A, rs.AnalyzeOptions(rs.OrderingStrategy.Grid, grid=info["grid"]))
    say("helmholtz64 seed 17 n=%d: FP64" % A.n)
    s.factor(A)
    x64 = s.solve(b)
    i0 = s.inertia()
    assert rs.sparse_backward_error(A, x64, b) <= 1e-14
    assert i0.negative == info["expected_negative"] and i0.zero == 0
    return A, b, info, s, x64, (i0.positive, i0.negative, i0.zero)


def _fwd(x, xref):
    return float(np.linalg.norm(x - xref) / np.linalg.norm(xref))


@pytest.mark.skipif(not os.path.exists(MODEL), reason="no trained network committed")
def test_learned_helmholtz64_held_out(helmholtz64):
    A, b, info, s, x64, iref = helmholtz64
    model = pr.PrecisionMLP.load(MODEL)
    prec = pr.assign(s, model)
    res = pr.solve_learned(s, A, b, model, prec=prec, inertia_ref=iref)
    i1 = s.inertia()
    fe = _fwd(res.x, x64)
    say("  learned: corrections %d, BE %.2e, fell back %s, forward error vs FP64 %.2e, classes %s" % (
        res.corrections, res.backward_error, res.fell_back, fe, np.bincount(prec, minlength=3).tolist()))
    assert res.converged and not res.fell_back and res.corrections <= 3
    assert res.backward_error <= 1e-12
    assert (i1.positive, i1.negative, i1.zero) == iref
    assert i1.negative == info["expected_negative"]
    assert fe <= 1e-6


def test_fp32_helmholtz64_refines(helmholtz64):
    A, b, info, s, x64, iref = helmholtz64
    s.set_front_precision(np.full(s.num_supernodes(), 1, dtype=np.int8))  # RV_PREC_FP32
    s.refactor(A)
    x, it, be = s.solve_refine(b, tol=1e-12, max_iter=10)
    i1 = s.inertia()
    fe = _fwd(x, x64)
    say("  fp32: corrections %d, BE %.2e, forward error vs FP64 %.2e" % (it, be, fe))
    assert be <= 1e-12 and it <= 3
    assert (i1.positive, i1.negative, i1.zero) == iref
    assert fe <= 1e-6
    s.set_front_precision(None)


@pytest.mark.skipif(not os.path.exists(MODEL), reason="no trained network committed")
def test_learned_kkt128_matches_oracle():
    A, b, _x, info = workloads.kkt(128, seed=19)
    s = rs.SparseLDLT.analyze_with_matrix(A, rs.AnalyzeOptions(rs.OrderingStrategy.MatchOrderMetis))
    say("kkt128 seed 19 n=%d" % A.n)
    s.factor(A)
    ref = oracle.OracleFactor(A, s.structure(), s._scaling)
    xr = ref.solve(b)
    i0 = s.inertia()
    assert (i0.positive, i0.negative, i0.zero) == ref.inertia() == tuple(info["expected_inertia"])
    assert s.stats().total_delayed == ref.stats()["total_delayed"]
    model = pr.PrecisionMLP.load(MODEL)
    prec = pr.assign(s, model)
    res = pr.solve_learned(s, A, b, model, prec=prec, inertia_ref=ref.inertia())
    i1 = s.inertia()
    say("  learned: corrections %d, BE %.2e, fell back %s, forward error vs oracle %.2e" % (
        res.corrections, res.backward_error, res.fell_back, _fwd(res.x, xr)))
    assert res.converged and not res.fell_back
    assert res.backward_error <= 1e-12
    assert (i1.positive, i1.negative, i1.zero) == ref.inertia()
