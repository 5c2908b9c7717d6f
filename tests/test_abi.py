"""C ABI: the library loads, exports every symbol include/rivrs_ldlt.h declares, and reports
errors the way SparseError does (no numeric work without a GPU)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import rivrs_sparse as rs
import workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rivrs_ldlt.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rv_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = rs.load_library()
    syms = declared_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(lib, s), s
    # the Python binding covers the same set
    assert set(syms) == set(rs.ABI)


def test_defaults_mirror_aptp_options():
    o = rs.FactorOptionsC()
    rs.load_library().rv_factor_options_default(C.byref(o))
    assert (o.threshold, o.small, o.outer_block_size, o.inner_block_size, o.failed_pivot_tpp, o.nemin,
            o.small_leaf_threshold) == (0.01, 1e-20, 256, 32, 1, 32, 256)


def _no_device():
    try:
        lib = C.CDLL("libamdhip64.so")
        n = C.c_int(0)
        return lib.hipGetDeviceCount(C.byref(n)) != 0 or n.value == 0
    except OSError:
        return True


def test_invalid_options_are_invalid_input():
    A = workloads.kkt(4, seed=0)[0]
    s = rs.SparseLDLT.analyze(A)
    for bad in (rs.FactorOptions(threshold=0.0), rs.FactorOptions(threshold=2.0), rs.FactorOptions(small=-1.0),
                rs.FactorOptions(inner_block_size=64, outer_block_size=256),
                rs.FactorOptions(inner_block_size=64, outer_block_size=32)):
        with pytest.raises(rs.InvalidInput):
            s.factor(A, bad)


def test_solve_before_factor():
    A = workloads.kkt(4, seed=0)[0]
    s = rs.SparseLDLT.analyze(A)
    with pytest.raises(rs.SolveBeforeFactor):
        s.solve(np.ones(A.n))
    assert s.inertia() is None and s.stats() is None


def test_dimension_mismatch():
    A = workloads.kkt(4, seed=0)[0]
    B = workloads.kkt(5, seed=0)[0]
    s = rs.SparseLDLT.analyze(A)
    with pytest.raises(rs.DimensionMismatch):
        s.factor(B)


@pytest.mark.skipif(not _no_device(), reason="a GPU is present")
def test_factor_without_gpu_fails_loudly():
    A = workloads.kkt(4, seed=0)[0]
    s = rs.SparseLDLT.analyze(A)
    with pytest.raises(rs.NoDevice):
        s.factor(A)
    with pytest.raises(rs.NoDevice):
        rs.dense_aptp_factor(np.eye(3), 3)


def test_structure_import_validation():
    A = workloads.kkt(4, seed=0)[0]
    st = rs.SparseLDLT.analyze(A).structure()
    ok = rs.SparseLDLT.from_structure(st)
    assert ok.n() == A.n
    bad = rs.Structure(st.n, st.sn_begin, st.sn_end, st.pat_ptr, st.pat, st.parent.copy(), st.perm_fwd)
    bad.parent[0] = 0  # parent must come after the child
    with pytest.raises(rs.InvalidInput):
        rs.SparseLDLT.from_structure(bad)
