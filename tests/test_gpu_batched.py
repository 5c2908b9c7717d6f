"""Graph replay of refactorizations and the batched (block-diagonal) form, on the GPU.

* a refactorization replayed from the captured hipGraph gives the same factors and solution, bit
  for bit, as the eager enqueue, over several replays with eager work between them; new values
  that keep every pivot decision replay, values that change the delays fall back to re-planning;
* many KKT systems factored and solved as one block-diagonal batch give, system by system, the
  solutions each gives alone (bit for bit), the summed inertia of the Haynsworth formula and
  backward errors at the FP64 bar.
"""
import sys

import numpy as np
import pytest

import batched
import rivrs_sparse as rs
import workloads

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_graph_replay_is_bitwise_eager():
    say("graph replay helmholtz16")
    A, b, info = workloads.helmholtz(16, seed=0)
    opts = rs.AnalyzeOptions(rs.OrderingStrategy.Grid, grid=info["grid"])
    eager = rs.SparseLDLT.analyze(A, opts)
    eager.factor(A)
    x0 = eager.solve(b)
    s = rs.SparseLDLT.analyze(A, opts)
    s.factor(A)
    s.set_graph_mode(True)
    for i in range(3):
        s.refactor(A)  # the first captures, the next replay
        eager.refactor(A)  # eager work between the replays
        assert np.array_equal(s.solve(b), x0), i
    assert s.graph_launches() == 3
    # scaled values: same decisions, the replay stays valid
    A2 = rs.SparseColMat(A.n, A.col_ptr, A.row_idx, 2.0 * A.values)
    eager.refactor(A2)
    s.refactor(A2)
    assert np.array_equal(s.solve(b), eager.solve(b))


def test_graph_replay_falls_back_when_delays_change():
    say("graph replay kkt32 delay change")
    A, b, x, info = workloads.kkt(32, seed=1)
    nh = info["side"] ** 2
    vals = A.values.copy()
    for j in range(nh, A.n):
        lo, hi = A.col_ptr[j], A.col_ptr[j + 1]
        rows = A.row_idx[lo:hi]
        vals[lo:hi][rows == j] = -1.0
    B = rs.SparseColMat(A.n, A.col_ptr, A.row_idx, vals)
    opts = rs.AnalyzeOptions(rs.OrderingStrategy.Metis)
    s = rs.SparseLDLT.analyze(A, opts)
    s.factor(A)
    s.set_graph_mode(True)
    s.refactor(A)
    s.refactor(A)
    d0 = s.stats().total_delayed
    s.refactor(B)  # different delays: re-planned, not replayed
    ref = rs.SparseLDLT.analyze(A, opts)
    ref.factor(B)
    assert s.stats().total_delayed == ref.stats().total_delayed != d0
    assert np.array_equal(s.solve(b), ref.solve(b))
    s.refactor(A)
    s.refactor(A)
    assert s.stats().total_delayed == d0


def test_batched_matches_individual_systems():
    say("batched kkt x10")
    batch = batched.kkt_batch(10, seed=3, lo=14, hi=24)
    systems = [A for A, _ in batch]
    rng = np.random.default_rng(0)
    rhs = [rng.standard_normal((A.n, 4)) for A in systems]
    B = batched.BatchedLDLT(systems)
    B.factor(systems)
    X = B.solve(rhs)
    B.factor(systems)  # captured
    B.factor(systems)  # replayed
    X2 = B.solve(rhs)
    assert B.solver.graph_launches() >= 1
    inert = B.solver.inertia()
    exp = np.sum([info["expected_inertia"] for _, info in batch], axis=0)
    assert (inert.positive, inert.negative, inert.zero) == tuple(exp)
    for i, A in enumerate(systems):
        s = rs.SparseLDLT.analyze_with_matrix(A, rs.AnalyzeOptions(rs.OrderingStrategy.MatchOrderMetis))
        s.factor(A)
        xi = s.solve(rhs[i])
        assert np.array_equal(X[i], xi), i
        assert np.array_equal(X2[i], xi), i
        for c in range(rhs[i].shape[1]):
            assert rs.sparse_backward_error(A, X[i][:, c], rhs[i][:, c]) <= 1e-14


def test_batched_config_shape_matches_oracle():
    # the batched configuration's shape: KKT systems of grid side in [14, 70] (both ends present)
    # with 64 N(0,1) right-hand sides each, as one block-diagonal factorization and one 64-column
    # solve (right-hand-side groups of 8 share every load of L)
    import oracle

    rng = np.random.default_rng(21)
    sides = [14, 70] + [int(v) for v in rng.integers(14, 71, 6)]
    batch = [workloads.kkt(sd, seed=1000 + i) for i, sd in enumerate(sides)]
    systems = [t[0] for t in batch]
    infos = [t[3] for t in batch]
    say("batched config shape: sides %s, 64 rhs" % sides)
    rhs = [rng.standard_normal((A.n, 64)) for A in systems]
    B = batched.BatchedLDLT(systems)
    B.factor(systems)
    X = B.solve(rhs)
    inert = B.solver.inertia()
    exp = np.sum([info["expected_inertia"] for info in infos], axis=0)
    assert (inert.positive, inert.negative, inert.zero) == tuple(exp)
    for i, A in enumerate(systems):
        for c in range(64):
            assert rs.sparse_backward_error(A, X[i][:, c], rhs[i][:, c]) <= 1e-14, (i, c)
    # the smallest, the largest and one more system against the oracle, all 64 columns
    for i in (0, 1, 2):
        A = systems[i]
        s = rs.SparseLDLT.analyze_with_matrix(A, rs.AnalyzeOptions(rs.OrderingStrategy.MatchOrderMetis))
        s.factor(A)
        ref = oracle.OracleFactor(A, s.structure(), s._scaling)
        xr = ref.solve(rhs[i])
        si = s.inertia()
        assert (si.positive, si.negative, si.zero) == ref.inertia() == tuple(infos[i]["expected_inertia"])
        assert s.stats().total_delayed == ref.stats()["total_delayed"]
        assert np.abs(X[i] - xr).max() <= 1e-9 * np.abs(xr).max(), i
        # the system alone gives the same solution as inside the batch, bit for bit
        assert np.array_equal(s.solve(rhs[i]), X[i]), i
