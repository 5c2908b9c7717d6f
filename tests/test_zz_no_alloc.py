"""rv_solve_device_noalloc (no counterpart in the reference: the solve for callers that cannot
allocate) against rv_solve_device through the C ABI: bitwise equal solutions on two calls in a
row, the second with a smaller system and fewer right-hand sides, and a clear error beyond the
fixed workspace."""
import numpy as np
import pytest

import batched
import rivrs_sparse as rs
import workloads

pytestmark = pytest.mark.gpu


def _factor(N, seed):
    A, b, info = workloads.helmholtz(N, seed=seed)
    s = rs.SparseLDLT.analyze(A, rs.AnalyzeOptions(rs.OrderingStrategy.Grid, grid=info["grid"]))
    s.factor(A)
    return A, s


def _both(lib, s, B):
    n, nrhs = B.shape
    d1 = batched.DeviceArray(B)
    d2 = batched.DeviceArray(B)
    rs._check(lib.rv_solve_device(s._num, d1.ptr, n, nrhs))
    rs._check(lib.rv_solve_device_noalloc(s._num, d2.ptr, n, nrhs))
    return d1.download(), d2.download()


def test_noalloc_solve_matches_solve_device():
    lib = rs.load_library()
    rng = np.random.default_rng(7)
    # call 1: 16^3 Helmholtz, 5 right-hand sides (a group of 4 and a group of 1)
    A1, s1 = _factor(16, 1)
    B1 = rng.standard_normal((A1.n, 5))
    print("noalloc: helmholtz16 n=%d nrhs=5 scratch=%d" % (A1.n, s1.solve_scratch(5)), flush=True)
    assert s1.solve_scratch(5) <= rs.RV_SOLVE_NOALLOC_BYTES
    X1, Y1 = _both(lib, s1, B1)
    assert np.array_equal(X1, Y1)
    for q in range(B1.shape[1]):
        assert rs.sparse_backward_error(A1, Y1[:, q], B1[:, q]) <= 1e-14
    # call 2: a smaller system and one right-hand side, on the array the first call left behind
    A2, s2 = _factor(10, 2)
    B2 = rng.standard_normal((A2.n, 1))
    print("noalloc: helmholtz10 n=%d nrhs=1" % A2.n, flush=True)
    X2, Y2 = _both(lib, s2, B2)
    assert np.array_equal(X2, Y2)
    assert rs.sparse_backward_error(A2, Y2[:, 0], B2[:, 0]) <= 1e-14
    # and the first handle again, after the second call
    X3, Y3 = _both(lib, s1, B1[:, :2])
    assert np.array_equal(X3, Y3) and np.array_equal(Y3, X1[:, :2])


def test_noalloc_solve_rejects_workspace_beyond_limit():
    lib = rs.load_library()
    A, s = _factor(10, 3)
    nrhs = 1
    while s.solve_scratch(nrhs) <= rs.RV_SOLVE_NOALLOC_BYTES:
        nrhs *= 2
    d = batched.DeviceArray(np.zeros((A.n, 1)))
    print("noalloc: limit check nrhs=%d scratch=%d" % (nrhs, s.solve_scratch(nrhs)), flush=True)
    # rejected before any device work: the right-hand-side buffer is never read
    code = lib.rv_solve_device_noalloc(s._num, d.ptr, A.n, nrhs)
    assert code == 3, code  # RV_INVALID_INPUT
    assert b"fixed" in lib.rv_last_error()
