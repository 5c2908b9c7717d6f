"""Per-front update precision and the precision classifier on the GPU, through the C ABI.

* the classifier's forward pass and loss gradient (rv_mlp_forward / rv_mlp_loss_grad) against the
  CPU restatement oracle/mlp.py (itself pinned by central differences), FP32 tolerances;
* training (rv_mlp_train) fits a separable problem;
* mixed per-front precision: an all-FP64 assignment reproduces the uniform FP64 factorization bit
  for bit; a random BF16/FP32/FP64 mix refines to 1e-12 with the FP64 inertia;
* the committed trained network on held-out matrices: the assignment it picks refines within 3
  steps with the FP64 inertia, and moves most contribution flops off FP64 — what the uniform forms
  cannot both do (BF16 does not converge on Helmholtz; FP64 moves nothing).
"""
import os
import sys

import numpy as np
import pytest

import mlp as mlp_oracle
import precision as pr
import rivrs_sparse as rs
import workloads

pytestmark = pytest.mark.gpu
MODEL = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "rivrs-linalg_amd", "models",
                     "precision_mlp.json")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _data(seed, n=700):
    rng = np.random.default_rng(seed)
    params = (rng.standard_normal(pr.N_PARAMS) * 0.2).astype(np.float32)
    x = rng.standard_normal((n, pr.N_FEATURES)).astype(np.float32)
    y = rng.integers(0, 3, n).astype(np.int32)
    return params, x, y


@pytest.mark.parametrize("seed", [0, 1])
def test_mlp_forward_and_gradient_match_oracle(seed):
    say("mlp oracle seed %d" % seed)
    params, x, y = _data(seed)
    model = pr.PrecisionMLP(params, np.zeros(16), np.ones(16))
    p = model.probabilities(x)
    pref, _ = mlp_oracle.forward(params.astype(np.float64), x.astype(np.float64))
    np.testing.assert_allclose(p, pref, atol=2e-6)
    cw = np.array([0.5, 1.0, 2.0], dtype=np.float32)
    loss, g = model.loss_grad(x, y, cw)
    lref, gref = mlp_oracle.loss_grad(params.astype(np.float64), x.astype(np.float64), y, cw.astype(np.float64))
    assert loss == pytest.approx(lref, rel=1e-5)
    assert np.abs(g - gref).max() <= 1e-4 * max(np.abs(gref).max(), 1e-6)


def test_mlp_training_fits_separable_classes():
    say("mlp training")
    rng = np.random.default_rng(5)
    x = rng.standard_normal((2000, 16)).astype(np.float32)
    y = np.where(x[:, 0] + 0.5 * x[:, 3] > 0.7, 2, np.where(x[:, 1] - x[:, 2] > 0.0, 1, 0)).astype(np.int32)
    model = pr.PrecisionMLP.initial(seed=1)
    l0, _ = model.loss_grad(x, y)
    loss = model.fit(x, y, epochs=600, lr=3e-3)
    acc = float(np.mean(model.predict(x) == y))
    say("  loss %.4f -> %.4f, accuracy %.3f" % (l0, loss, acc))
    assert loss < 0.25 * l0 and acc > 0.95
    # deterministic: the same training gives the same parameters
    m2 = pr.PrecisionMLP.initial(seed=1)
    m2.fit(x, y, epochs=600, lr=3e-3)
    assert np.array_equal(m2.params, model.params)


def _helmholtz(N, seed):
    A, b, info = workloads.helmholtz(N, seed=seed)
    s = rs.SparseLDLT.analyze(A, rs.AnalyzeOptions(rs.OrderingStrategy.Grid, grid=info["grid"]))
    return A, b, s


def test_front_precision_all_fp64_is_uniform_fp64():
    say("front precision fp64 vs uniform")
    A, b, s = _helmholtz(16, 0)
    s.factor(A)
    x0 = s.solve(b)
    s.set_front_precision(np.zeros(s.num_supernodes(), dtype=np.int8))
    s.refactor(A)
    assert np.array_equal(s.solve(b), x0)
    assert list(s.front_precision()) == [0] * s.num_supernodes()


@pytest.mark.parametrize("case", ["helmholtz16", "kkt32"])
def test_mixed_front_precision_refines(case):
    say("mixed front precision %s" % case)
    if case == "helmholtz16":
        A, b, s = _helmholtz(16, 0)
    else:
        A, b, _x, _ = workloads.kkt(32, seed=0)
        s = rs.SparseLDLT.analyze_with_matrix(A, rs.AnalyzeOptions(rs.OrderingStrategy.MatchOrderMetis))
    s.factor(A)
    i0 = s.inertia()
    ns = s.num_supernodes()
    rng = np.random.default_rng(3)
    prec = rng.choice(np.array([0, 1, 2], dtype=np.int8), ns, p=[0.3, 0.5, 0.2])
    if case == "helmholtz16":
        prec[prec == 2] = 1  # BF16 on the near-singular Helmholtz fronts does not converge
    s.set_front_precision(prec)
    s.refactor(A)
    assert np.array_equal(s.front_precision(), prec)
    x, it, be = s.solve_refine(b, tol=1e-12, max_iter=10)
    i1 = s.inertia()
    say("  %d corrections, BE %.2e" % (it, be))
    assert be <= 1e-12
    assert (i1.positive, i1.negative, i1.zero) == (i0.positive, i0.negative, i0.zero)
    with pytest.raises(rs.InvalidInput):
        s.set_front_precision(np.full(ns, 7, dtype=np.int8))


@pytest.mark.skipif(not os.path.exists(MODEL), reason="no trained network committed")
@pytest.mark.parametrize("case", [("helmholtz", 28, 9), ("kkt", 56, 9)])
def test_trained_network_on_held_out_matrices(case):
    kind, size, seed = case
    say("learned precision %s %d seed %d" % case)
    model = pr.PrecisionMLP.load(MODEL)
    if kind == "helmholtz":
        A, b, s = _helmholtz(size, seed)
    else:
        A, b, _x, _ = workloads.kkt(size, seed=seed)
        s = rs.SparseLDLT.analyze_with_matrix(A, rs.AnalyzeOptions(rs.OrderingStrategy.MatchOrderMetis))
    s.factor(A)
    i0 = s.inertia()
    iref = (i0.positive, i0.negative, i0.zero)
    prec = pr.assign(s, model)
    res = pr.solve_learned(s, A, b, model, prec=prec, inertia_ref=iref)
    per = s.per_supernode_stats()
    fl = np.array([q.num_eliminated * (q.front_size - q.num_fully_summed) * (q.front_size - q.num_fully_summed + 1.0)
                   for q in per])
    reduced = float(fl[prec != 0].sum() / max(fl.sum(), 1.0))
    say("  corrections %d, BE %.2e, fell back %s, reduced-precision share of contribution flops %.2f" % (
        res.corrections, res.backward_error, res.fell_back, reduced))
    assert res.converged and not res.fell_back
    assert res.backward_error <= 1e-12
    assert reduced >= 0.5


@pytest.mark.parametrize("case", ["helmholtz14", "kkt40"])
def test_solve_after_replanned_refactor(case):
    # a precision change re-plans every level (the fronts may move in device memory); the next
    # solve must use the new fronts: same solution as a handle that factored that way directly
    say("solve after re-planned refactor %s" % case)
    if case == "helmholtz14":
        A, b, info = workloads.helmholtz(14, seed=4)
        opts = rs.AnalyzeOptions(rs.OrderingStrategy.Grid, grid=info["grid"])
        s = rs.SparseLDLT.analyze(A, opts)
        t = rs.SparseLDLT.analyze(A, opts)
    else:
        A, b, _x, _ = workloads.kkt(40, seed=4)
        s = rs.SparseLDLT.analyze_with_matrix(A, rs.AnalyzeOptions(rs.OrderingStrategy.MatchOrderMetis))
        t = rs.SparseLDLT.analyze_with_matrix(A, rs.AnalyzeOptions(rs.OrderingStrategy.MatchOrderMetis))
    s.factor(A)
    x0 = s.solve(b)
    assert rs.sparse_backward_error(A, x0, b) <= 1e-14
    rng = np.random.default_rng(11)
    prec = rng.choice(np.array([0, 1], dtype=np.int8), s.num_supernodes())
    for _ in range(2):
        s.set_front_precision(prec)
        s.refactor(A)
        x1, it1, be1 = s.solve_refine(b, tol=1e-12, max_iter=10)
        t.factor(A)
        t.set_front_precision(prec)
        t.refactor(A)
        x2, it2, be2 = t.solve_refine(b, tol=1e-12, max_iter=10)
        assert be1 <= 1e-12 and np.array_equal(x1, x2) and it1 == it2
        s.set_front_precision(None)
        s.refactor(A)
        assert np.array_equal(s.solve(b), x0)
