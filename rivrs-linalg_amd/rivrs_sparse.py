"""Python mirror of rivrs-sparse's SparseLDLT API over the MI355X numeric phase.

Names, argument meaning and error behaviour follow sparse/src/symmetric/solver.rs
(SparseLDLT, AnalyzeOptions, FactorOptions, SolverOptions, OrderingStrategy) and
sparse/src/error.rs (SparseError).  Every numeric call goes through the C ABI of
librivrs_ldlt.so (include/rivrs_ldlt.h); there is no CPU fallback: without a GPU the
factorization raises SparseError with status RV_NO_DEVICE.
"""
from __future__ import annotations

import ctypes as C
import enum
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librivrs_ldlt.so")

# ---------------------------------------------------------------------------
# errors (sparse/src/error.rs:20-98)
# ---------------------------------------------------------------------------
RV_OK = 0
RV_DIMENSION_MISMATCH = 1
RV_NOT_SQUARE = 2
RV_INVALID_INPUT = 3
RV_ANALYSIS_FAILURE = 6
RV_SOLVE_BEFORE_FACTOR = 9
RV_DEVICE_ERROR = 100
RV_NO_DEVICE = 101
RV_SOLVE_NOALLOC_BYTES = 64 << 20  # workspace of rv_solve_device_noalloc


class SparseError(Exception):
    """Base of the error hierarchy; `status` is the C ABI rv_status code."""

    status = RV_DEVICE_ERROR

    def __init__(self, message: str, status: Optional[int] = None):
        super().__init__(message)
        if status is not None:
            self.status = status


class DimensionMismatch(SparseError):
    status = RV_DIMENSION_MISMATCH


class NotSquare(SparseError):
    status = RV_NOT_SQUARE


class InvalidInput(SparseError):
    status = RV_INVALID_INPUT


class AnalysisFailure(SparseError):
    status = RV_ANALYSIS_FAILURE


class SolveBeforeFactor(SparseError):
    status = RV_SOLVE_BEFORE_FACTOR


class DeviceError(SparseError):
    status = RV_DEVICE_ERROR


class NoDevice(DeviceError):
    status = RV_NO_DEVICE


_ERRORS = {
    RV_DIMENSION_MISMATCH: DimensionMismatch,
    RV_NOT_SQUARE: NotSquare,
    RV_INVALID_INPUT: InvalidInput,
    RV_ANALYSIS_FAILURE: AnalysisFailure,
    RV_SOLVE_BEFORE_FACTOR: SolveBeforeFactor,
    RV_DEVICE_ERROR: DeviceError,
    RV_NO_DEVICE: NoDevice,
}

# ---------------------------------------------------------------------------
# C ABI binding
# ---------------------------------------------------------------------------


class FactorOptionsC(C.Structure):
    _fields_ = [
        ("threshold", C.c_double),
        ("small", C.c_double),
        ("outer_block_size", C.c_int32),
        ("inner_block_size", C.c_int32),
        ("failed_pivot_tpp", C.c_int32),
        ("nemin", C.c_int32),
        ("small_leaf_threshold", C.c_int32),
        ("update_precision", C.c_int32),
    ]


_P = C.c_void_p
_I64P = C.POINTER(C.c_int64)
_I32P = C.POINTER(C.c_int32)
_I8P = C.POINTER(C.c_int8)
_DP = C.POINTER(C.c_double)

# name -> (restype, argtypes); every symbol include/rivrs_ldlt.h declares
ABI = {
    "rv_last_error": (C.c_char_p, []),
    "rv_factor_options_default": (None, [C.POINTER(FactorOptionsC)]),
    "rv_analyze": (C.c_int, [C.c_int64, _I64P, _I64P, C.c_int32, _I32P, _I32P, C.POINTER(_P)]),
    "rv_analyze_with_matrix": (
        C.c_int, [C.c_int64, _I64P, _I64P, _DP, C.c_int32, _I32P, _I32P, C.POINTER(_P), _DP, _I32P]),
    "rv_mc64_matching": (C.c_int, [C.c_int64, _I64P, _I64P, _DP, _I32P, _DP, _I32P, _I64P]),
    "rv_match_order": (C.c_int, [C.c_int64, _I64P, _I64P, _DP, _I32P, _DP, _I64P]),
    "rv_symbolic_from_structure": (
        C.c_int, [C.c_int64, C.c_int32, _I32P, _I32P, _I64P, _I32P, _I32P, _I32P, C.POINTER(_P)]),
    "rv_symbolic_info": (C.c_int, [_P, _I64P]),
    "rv_symbolic_export": (C.c_int, [_P, _I32P, _I32P, _I64P, _I32P, _I32P, _I32P]),
    "rv_symbolic_free": (None, [_P]),
    "rv_amalgamate": (
        C.c_int32, [C.c_int32, _I32P, _I32P, _I64P, _I32P, _I32P, C.c_int32, _I32P, _I32P, _I32P, _I64P,
                    _I32P, _I32P, _I64P, _I32P]),
    "rv_factor": (C.c_int, [_P, _I64P, _I64P, _DP, _DP, C.POINTER(FactorOptionsC), C.POINTER(_P)]),
    "rv_factor_device": (
        C.c_int, [_P, _I64P, _I64P, C.c_void_p, C.c_void_p, C.POINTER(FactorOptionsC), C.POINTER(_P)]),
    "rv_refactor": (C.c_int, [_P, _DP]),
    "rv_solve_refine": (C.c_int, [_P, _DP, _DP, C.c_int32, C.c_double, _I32P, _DP]),
    "rv_solve_refine_device": (C.c_int, [_P, C.c_void_p, C.c_void_p, C.c_int32, C.c_double, _I32P, _DP]),
    "rv_refactor_device": (C.c_int, [_P, C.c_void_p]),
    "rv_solve": (C.c_int, [_P, _DP, C.c_int64, C.c_int32]),
    "rv_solve_device": (C.c_int, [_P, C.c_void_p, C.c_int64, C.c_int32]),
    "rv_solve_device_noalloc": (C.c_int, [_P, C.c_void_p, C.c_int64, C.c_int32]),
    "rv_solve_scratch": (C.c_int64, [_P, C.c_int32]),
    "rv_inertia": (C.c_int, [_P, _I64P]),
    "rv_stats": (C.c_int, [_P, _I64P]),
    "rv_num_supernodes": (C.c_int32, [_P]),
    "rv_set_front_precision": (C.c_int, [_P, _I8P]),
    "rv_set_graph_mode": (C.c_int, [_P, C.c_int32]),
    "rv_graph_launches": (C.c_int64, [_P]),
    "rv_chain_launches": (C.c_int64, [_P]),
    "rv_lookahead_launches": (C.c_int64, [_P]),
    "rv_device_bytes": (C.c_int64, [_P]),
    "rv_front_precision": (C.c_int, [_P, _I8P]),
    "rv_front_features": (C.c_int, [_P, C.POINTER(C.c_float)]),
    "rv_mlp_forward": (C.c_int, [C.POINTER(C.c_float), C.c_int64, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "rv_mlp_loss_grad": (C.c_int, [C.POINTER(C.c_float), C.c_int64, C.POINTER(C.c_float), _I32P,
                                   C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "rv_mlp_train": (C.c_int, [C.POINTER(C.c_float), C.c_int64, C.POINTER(C.c_float), _I32P,
                               C.POINTER(C.c_float), C.c_int32, C.c_float, C.POINTER(C.c_float)]),
    "rv_supernode_stats": (C.c_int, [_P, C.c_int32, _I64P, _DP]),
    "rv_supernode_structure": (C.c_int, [_P, C.c_int32, _I32P, _I32P, _I64P]),
    "rv_export_front": (C.c_int, [_P, C.c_int32, _DP, _DP, _I8P, _DP, _DP, _I32P, _I32P, _I32P]),
    "rv_export_contribution": (C.c_int, [_P, C.c_int32, _DP, _I32P, _I64P]),
    "rv_dense_aptp_factor": (
        C.c_int, [_DP, C.c_int64, C.c_int32, C.c_int32, C.POINTER(FactorOptionsC), _I32P, _I8P, _DP,
                  _DP, _I64P, _DP]),
    "rv_set_timing": (C.c_int, [_P, C.c_int32]),
    "rv_timings": (C.c_int, [_P, _DP]),
    "rv_numeric_free": (None, [_P]),
}

_lib = None


def load_library(path: str = LIB_PATH):
    """Load librivrs_ldlt.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise DeviceError("librivrs_ldlt.so not built: run __graft_entry__.build() (%s)" % path)
    lib = C.CDLL(path)
    for name, (res, args) in ABI.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _check(code: int):
    if code != RV_OK:
        msg = load_library().rv_last_error().decode(errors="replace")
        raise _ERRORS.get(code, SparseError)(msg, code)


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


# ---------------------------------------------------------------------------
# matrices
# ---------------------------------------------------------------------------


class SparseColMat:
    """Full-symmetric CSC matrix (both triangles stored), as SparseColMat<usize, f64>."""

    def __init__(self, n: int, col_ptr, row_idx, values):
        self.n = int(n)
        self.col_ptr = np.ascontiguousarray(col_ptr, dtype=np.int64)
        self.row_idx = np.ascontiguousarray(row_idx, dtype=np.int64)
        self.values = np.ascontiguousarray(values, dtype=np.float64)
        if self.col_ptr.shape != (self.n + 1,):
            raise DimensionMismatch("col_ptr must have n+1 entries")
        if self.row_idx.shape[0] != self.col_ptr[-1] or self.values.shape[0] != self.col_ptr[-1]:
            raise DimensionMismatch("row_idx/values length must equal col_ptr[n]")

    @property
    def nnz(self) -> int:
        return int(self.col_ptr[-1])

    def nrows(self) -> int:
        return self.n

    def ncols(self) -> int:
        return self.n

    @staticmethod
    def from_scipy(a) -> "SparseColMat":
        import scipy.sparse as sp

        a = sp.csc_matrix(a)
        if a.shape[0] != a.shape[1]:
            raise NotSquare("Matrix must be square, but has dimensions %dx%d" % a.shape)
        a.sum_duplicates()
        a.sort_indices()
        return SparseColMat(a.shape[0], a.indptr, a.indices, a.data)

    @staticmethod
    def from_lower_triplets(n: int, rows, cols, vals) -> "SparseColMat":
        """Mirror lower-triangle triplets into full symmetric storage."""
        import scipy.sparse as sp

        rows = np.asarray(rows)
        cols = np.asarray(cols)
        vals = np.asarray(vals, dtype=np.float64)
        off = rows != cols
        r = np.concatenate([rows, cols[off]])
        c = np.concatenate([cols, rows[off]])
        v = np.concatenate([vals, vals[off]])
        return SparseColMat.from_scipy(sp.coo_matrix((v, (r, c)), shape=(n, n)))

    def to_scipy(self):
        import scipy.sparse as sp

        return sp.csc_matrix((self.values, self.row_idx, self.col_ptr), shape=(self.n, self.n))

    def matvec(self, x: np.ndarray) -> np.ndarray:
        return self.to_scipy() @ x


def load_mtx(path: str) -> SparseColMat:
    """Matrix Market reader (io/mtx.rs:27): symmetric files are mirrored to full storage."""
    rows, cols, vals = [], [], []
    symmetric = False
    import gzip

    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "rt") as f:
        header = f.readline().lower()
        if not header.startswith("%%matrixmarket"):
            raise SparseError("not a Matrix Market file: " + path)
        symmetric = "symmetric" in header
        pattern = "pattern" in header
        line = f.readline()
        while line.startswith("%"):
            line = f.readline()
        nr, nc, nnz = (int(t) for t in line.split()[:3])
        if nr != nc:
            raise NotSquare("Matrix must be square, but has dimensions %dx%d" % (nr, nc))
        data = np.loadtxt(f, ndmin=2) if nnz else np.zeros((0, 3))
    i = data[:, 0].astype(np.int64) - 1
    j = data[:, 1].astype(np.int64) - 1
    v = np.ones(len(i)) if pattern else data[:, 2]
    if symmetric:
        lo = np.maximum(i, j)
        hi = np.minimum(i, j)
        return SparseColMat.from_lower_triplets(nr, lo, hi, v)
    import scipy.sparse as sp

    return SparseColMat.from_scipy(sp.coo_matrix((v, (i, j)), shape=(nr, nr)))


def sparse_backward_error(a: SparseColMat, x: np.ndarray, b: np.ndarray) -> float:
    """||Ax-b||_2 / (||A||_F ||x||_2 + ||b||_2) from full CSC storage (validate.rs:209-255)."""
    r = a.matvec(x) - b
    denom = np.sqrt(np.sum(a.values * a.values)) * np.linalg.norm(x) + np.linalg.norm(b)
    if denom == 0.0:
        return 0.0
    return float(np.linalg.norm(r) / denom)


# ---------------------------------------------------------------------------
# options (solver.rs:42-148)
# ---------------------------------------------------------------------------


class OrderingStrategy(enum.Enum):
    """OrderingStrategy (solver.rs:42-55).  Metis maps to the library's nested dissection
    (METIS itself is not available); Grid is nested dissection by planes for grid matrices."""

    Natural = "natural"
    Metis = "metis"
    MatchOrderMetis = "match_order_metis"
    UserSupplied = "user"
    Grid = "grid"


@dataclass
class AnalyzeOptions:
    ordering: OrderingStrategy = OrderingStrategy.Metis
    perm: Optional[np.ndarray] = None          # OrderingStrategy.UserSupplied: perm_fwd[new] = old
    grid: Optional[Sequence[int]] = None       # OrderingStrategy.Grid: (nx, ny, nz)


@dataclass
class FactorOptions:
    threshold: float = 0.01
    outer_block_size: int = 256
    inner_block_size: int = 32
    nemin: int = 32
    small_leaf_threshold: int = 256
    small: float = 1e-20
    failed_pivot_tpp: bool = True
    # arithmetic of the Schur-complement updates: "fp64" (FP64 MFMA contribution GEMM, outer-block
    # updates in the reference's order), "fp64_mfma" (both on FP64 MFMA), "fp64_exact" (FP64 in
    # the reference's summation order), "fp32" or "bf16" (MFMA, FP32 accumulate)
    update_precision: str = "fp64"

    def to_c(self) -> FactorOptionsC:
        o = FactorOptionsC()
        o.threshold = self.threshold
        o.small = self.small
        o.outer_block_size = self.outer_block_size
        o.inner_block_size = self.inner_block_size
        o.failed_pivot_tpp = 1 if self.failed_pivot_tpp else 0
        o.nemin = self.nemin
        o.small_leaf_threshold = self.small_leaf_threshold
        o.update_precision = {"fp64": 0, "fp32": 1, "bf16": 2, "fp64_exact": 3, "fp64_mfma": 4}[self.update_precision]
        return o


@dataclass
class SolverOptions:
    ordering: OrderingStrategy = OrderingStrategy.Metis
    threshold: float = 0.01
    nemin: int = 32
    small_leaf_threshold: int = 256
    perm: Optional[np.ndarray] = None
    grid: Optional[Sequence[int]] = None


@dataclass
class Inertia:
    positive: int
    negative: int
    zero: int

    def dimension(self) -> int:
        return self.positive + self.negative + self.zero


@dataclass
class FactorizationStats:
    total_1x1_pivots: int
    total_2x2_pivots: int
    total_delayed: int
    zero_pivots: int
    max_front_size: int
    supernodes_before_amalgamation: int
    supernodes_after_amalgamation: int
    merges_performed: int
    small_leaf_subtrees: int
    small_leaf_nodes: int


@dataclass
class PerSupernodeStats:
    snode_id: int
    front_size: int
    num_fully_summed: int
    num_eliminated: int
    num_delayed: int
    num_1x1: int
    num_2x2: int
    max_l_entry: float
    parent: int
    in_small_leaf: bool


@dataclass
class Structure:
    """Pre-amalgamation supernode partition (the data build_supernode_info reads)."""

    n: int
    sn_begin: np.ndarray
    sn_end: np.ndarray
    pat_ptr: np.ndarray
    pat: np.ndarray
    parent: np.ndarray
    perm_fwd: np.ndarray
    predicted_nnz: int = 0


_ORDER_CODE = {
    OrderingStrategy.Natural: 0,
    OrderingStrategy.UserSupplied: 1,
    OrderingStrategy.Metis: 2,
    OrderingStrategy.Grid: 3,
    OrderingStrategy.MatchOrderMetis: 4,
}


@dataclass
class Mc64Result:
    """Mc64Result (matching.rs:31-58): matching[i] = column of row i, symmetric scaling."""
    matching: np.ndarray
    scaling: np.ndarray
    matched: int
    is_matched: np.ndarray


@dataclass
class MatchOrderResult:
    """MatchOrderResult (metis.rs:60-80): ordering (order[new] = old), scaling, diagnostics."""
    ordering: np.ndarray
    scaling: np.ndarray
    matched: int
    condensed_dim: int
    singletons: int
    two_cycles: int


def mc64_matching(matrix: SparseColMat) -> Mc64Result:
    """mc64_matching (matching.rs:121-248), job MaximumProduct."""
    lib = load_library()
    n = matrix.n
    mt = np.zeros(max(n, 1), dtype=np.int32)
    sc = np.zeros(max(n, 1), dtype=np.float64)
    im = np.zeros(max(n, 1), dtype=np.int32)
    cnt = C.c_int64(0)
    _check(lib.rv_mc64_matching(n, _ptr(matrix.col_ptr, C.c_int64), _ptr(matrix.row_idx, C.c_int64),
                                _ptr(np.ascontiguousarray(matrix.values, dtype=np.float64), C.c_double),
                                _ptr(mt, C.c_int32), _ptr(sc, C.c_double), _ptr(im, C.c_int32), C.byref(cnt)))
    return Mc64Result(mt[:n].copy(), sc[:n].copy(), int(cnt.value), im[:n].astype(bool))


def match_order_metis(matrix: SparseColMat) -> MatchOrderResult:
    """match_order_metis (metis.rs:118-190); the library's nested dissection replaces METIS."""
    lib = load_library()
    n = matrix.n
    order = np.zeros(max(n, 1), dtype=np.int32)
    sc = np.zeros(max(n, 1), dtype=np.float64)
    info = np.zeros(4, dtype=np.int64)
    vals = np.ascontiguousarray(matrix.values, dtype=np.float64)
    if vals.size == 0:
        vals = np.zeros(1)
    _check(lib.rv_match_order(n, _ptr(matrix.col_ptr, C.c_int64), _ptr(matrix.row_idx, C.c_int64),
                              _ptr(vals, C.c_double), _ptr(order, C.c_int32), _ptr(sc, C.c_double),
                              _ptr(info, C.c_int64)))
    return MatchOrderResult(order[:n].copy(), sc[:n].copy(), int(info[0]), int(info[1]), int(info[2]),
                            int(info[3]))


class SparseLDLT:
    """SparseLDLT (solver.rs:178-548): analyze -> factor -> solve, numeric phase on the GPU."""

    def __init__(self, sym_handle, n: int, scaling: Optional[np.ndarray]):
        self._lib = load_library()
        self._sym = sym_handle
        self._n = n
        self._num = None
        self._scaling = None if scaling is None else np.ascontiguousarray(scaling, dtype=np.float64)
        self._pattern = None

    def __del__(self):
        lib = getattr(self, "_lib", None)
        if lib is None:
            return
        if getattr(self, "_num", None):
            lib.rv_numeric_free(self._num)
            self._num = None
        if getattr(self, "_sym", None):
            lib.rv_symbolic_free(self._sym)
            self._sym = None

    # -- analyze ------------------------------------------------------------
    @staticmethod
    def analyze(matrix: SparseColMat, options: Optional[AnalyzeOptions] = None) -> "SparseLDLT":
        """SparseLDLT::analyze (solver.rs:195-259): symbolic only."""
        options = options or AnalyzeOptions()
        if options.ordering == OrderingStrategy.MatchOrderMetis:
            raise AnalysisFailure(
                "MatchOrderMetis requires numeric matrix values; use analyze_with_matrix() instead",
                RV_ANALYSIS_FAILURE)
        lib = load_library()
        code = _ORDER_CODE[options.ordering]
        perm = grid = None
        if options.ordering == OrderingStrategy.UserSupplied:
            if options.perm is None:
                raise InvalidInput("UserSupplied ordering needs a permutation")
            perm = np.ascontiguousarray(options.perm, dtype=np.int32)
            if perm.shape[0] != matrix.n:
                raise DimensionMismatch(
                    "Custom permutation has dimension %d but matrix is %dx%d"
                    % (perm.shape[0], matrix.n, matrix.n))
        if options.ordering == OrderingStrategy.Grid:
            grid = np.ascontiguousarray(options.grid, dtype=np.int32)
        h = C.c_void_p()
        _check(lib.rv_analyze(matrix.n, _ptr(matrix.col_ptr, C.c_int64), _ptr(matrix.row_idx, C.c_int64),
                              code, None if perm is None else _ptr(perm, C.c_int32),
                              None if grid is None else _ptr(grid, C.c_int32), C.byref(h)))
        return SparseLDLT(h, matrix.n, None)

    @staticmethod
    def analyze_with_matrix(matrix: SparseColMat, options: Optional[AnalyzeOptions] = None) -> "SparseLDLT":
        """SparseLDLT::analyze_with_matrix (solver.rs:261-285)."""
        options = options or AnalyzeOptions()
        if options.ordering != OrderingStrategy.MatchOrderMetis:
            return SparseLDLT.analyze(matrix, options)
        lib = load_library()
        n = matrix.n
        sc = np.zeros(max(n, 1), dtype=np.float64)
        has = C.c_int32(0)
        h = C.c_void_p()
        vals = np.ascontiguousarray(matrix.values, dtype=np.float64)
        if vals.size == 0:
            vals = np.zeros(1)
        _check(lib.rv_analyze_with_matrix(n, _ptr(matrix.col_ptr, C.c_int64), _ptr(matrix.row_idx, C.c_int64),
                                          _ptr(vals, C.c_double), _ORDER_CODE[options.ordering], None, None,
                                          C.byref(h), _ptr(sc, C.c_double), C.byref(has)))
        return SparseLDLT(h, n, sc[:n].copy() if has.value else None)

    @staticmethod
    def from_structure(st: Structure, scaling: Optional[np.ndarray] = None) -> "SparseLDLT":
        """Adopt an existing supernode partition (e.g. the reference's AptpSymbolic)."""
        lib = load_library()
        h = C.c_void_p()
        b = np.ascontiguousarray(st.sn_begin, dtype=np.int32)
        e = np.ascontiguousarray(st.sn_end, dtype=np.int32)
        pp = np.ascontiguousarray(st.pat_ptr, dtype=np.int64)
        pt = np.ascontiguousarray(st.pat, dtype=np.int32)
        pa = np.ascontiguousarray(st.parent, dtype=np.int32)
        pf = np.ascontiguousarray(st.perm_fwd, dtype=np.int32)
        if pt.size == 0:
            pt = np.zeros(1, dtype=np.int32)
        _check(lib.rv_symbolic_from_structure(st.n, len(b), _ptr(b, C.c_int32), _ptr(e, C.c_int32),
                                              _ptr(pp, C.c_int64), _ptr(pt, C.c_int32),
                                              _ptr(pa, C.c_int32), _ptr(pf, C.c_int32), C.byref(h)))
        return SparseLDLT(h, st.n, scaling)

    def structure(self) -> Structure:
        info = np.zeros(3, dtype=np.int64)
        _check(self._lib.rv_symbolic_info(self._sym, _ptr(info, C.c_int64)))
        n, ns = int(info[0]), int(info[1])
        b = np.zeros(ns, np.int32)
        e = np.zeros(ns, np.int32)
        pp = np.zeros(ns + 1, np.int64)
        pa = np.zeros(ns, np.int32)
        pf = np.zeros(max(n, 1), np.int32)
        _check(self._lib.rv_symbolic_export(self._sym, _ptr(b, C.c_int32), _ptr(e, C.c_int32),
                                            _ptr(pp, C.c_int64), None, _ptr(pa, C.c_int32),
                                            _ptr(pf, C.c_int32)))
        pt = np.zeros(max(int(pp[-1]), 1), np.int32)
        _check(self._lib.rv_symbolic_export(self._sym, None, None, None, _ptr(pt, C.c_int32), None, None))
        return Structure(n, b, e, pp, pt[: int(pp[-1])], pa, pf[:n], int(info[2]))

    def set_scaling(self, scaling: Optional[np.ndarray]):
        """Scaling in elimination order, as stored by analyze_with_matrix (solver.rs:276-278)."""
        self._scaling = None if scaling is None else np.ascontiguousarray(scaling, dtype=np.float64)

    # -- factor -------------------------------------------------------------
    def factor(self, matrix: SparseColMat, options: Optional[FactorOptions] = None):
        """SparseLDLT::factor (solver.rs:296-324)."""
        options = options or FactorOptions()
        if matrix.n != self._n:
            raise DimensionMismatch("Matrix dimensions must match symbolic analysis")
        if self._num:
            self._lib.rv_numeric_free(self._num)  # drop the old numeric first (solver.rs:312-315)
            self._num = None
        h = C.c_void_p()
        o = options.to_c()
        sc = None if self._scaling is None else _ptr(self._scaling, C.c_double)
        _check(self._lib.rv_factor(self._sym, _ptr(matrix.col_ptr, C.c_int64), _ptr(matrix.row_idx, C.c_int64),
                                   _ptr(matrix.values, C.c_double), sc, C.byref(o), C.byref(h)))
        self._num = h
        self._pattern = matrix
        self._options = options

    def refactor(self, matrix: SparseColMat, options: Optional[FactorOptions] = None):
        """SparseLDLT::refactor (solver.rs:335-341): same results as factor().  With the options
        of the previous factorization the numeric handle is kept (rv_refactor), so its level
        plans and device memory are reused."""
        options = options or FactorOptions()
        if not self._num or options != getattr(self, "_options", None):
            self.factor(matrix, options)
            return
        if matrix.n != self._n:
            raise DimensionMismatch("Matrix dimensions must match symbolic analysis")
        if matrix.nnz != self._pattern.nnz:
            raise DimensionMismatch("refactor needs the pattern of the previous factorization")
        _check(self._lib.rv_refactor(self._num, _ptr(matrix.values, C.c_double)))
        self._pattern = matrix

    # -- solve --------------------------------------------------------------
    def _need_numeric(self):
        if not self._num:
            raise SolveBeforeFactor("factor() must be called before solve()", RV_SOLVE_BEFORE_FACTOR)

    def solve_in_place(self, rhs: np.ndarray) -> np.ndarray:
        """SparseLDLT::solve_in_place (solver.rs:370-428); rhs of shape (n,) or (n, nrhs)."""
        self._need_numeric()
        if rhs.dtype != np.float64:
            raise InvalidInput("rhs must be float64")
        if rhs.shape[0] != self._n:
            raise DimensionMismatch("RHS length must match matrix dimension")
        x = np.asfortranarray(rhs.reshape(self._n, -1))
        nrhs = x.shape[1]
        _check(self._lib.rv_solve(self._num, _ptr(x, C.c_double), self._n, nrhs))
        rhs[...] = x.reshape(rhs.shape)
        return rhs

    def solve(self, rhs: np.ndarray) -> np.ndarray:
        """SparseLDLT::solve (solver.rs:351-360)."""
        out = np.array(rhs, dtype=np.float64, copy=True)
        return self.solve_in_place(out)

    def solve_refine(self, rhs: np.ndarray, tol: float = 1e-12, max_iter: int = 10):
        """Solve with FP64 iterative refinement (rv_solve_refine; the reference has none):
        returns (x, corrections applied, scaled backward error of x)."""
        self._need_numeric()
        b = np.ascontiguousarray(rhs, dtype=np.float64)
        if b.shape != (self._n,):
            raise DimensionMismatch("RHS length must equal matrix dimension")
        x = np.empty_like(b)
        it = C.c_int32(0)
        be = C.c_double(0.0)
        _check(self._lib.rv_solve_refine(self._num, _ptr(b, C.c_double), _ptr(x, C.c_double), int(max_iter),
                                         float(tol), C.byref(it), C.byref(be)))
        return x, int(it.value), float(be.value)

    # -- per-front update precision (extension; the reference is FP64 only) --------------
    def num_supernodes(self) -> int:
        self._need_numeric()
        return int(self._lib.rv_num_supernodes(self._num))

    def front_features(self) -> np.ndarray:
        """rv_front_features: 16 classifier features per supernode from the last factorization."""
        self._need_numeric()
        out = np.zeros((self.num_supernodes(), 16), dtype=np.float32)
        _check(self._lib.rv_front_features(self._num, _ptr(out, C.c_float)))
        return out

    def set_front_precision(self, prec: Optional[np.ndarray]):
        """Per-supernode RV_PREC_* codes for the next refactorization (None = uniform option)."""
        self._need_numeric()
        if prec is None:
            _check(self._lib.rv_set_front_precision(self._num, None))
            return
        p = np.ascontiguousarray(prec, dtype=np.int8)
        if p.shape != (self.num_supernodes(),):
            raise DimensionMismatch("one precision code per supernode required")
        _check(self._lib.rv_set_front_precision(self._num, _ptr(p, C.c_int8)))

    def set_graph_mode(self, on: bool = True):
        """rv_set_graph_mode: replay refactorizations on unchanged plans as one hipGraph launch."""
        self._need_numeric()
        _check(self._lib.rv_set_graph_mode(self._num, 1 if on else 0))

    def device_bytes(self) -> int:
        self._need_numeric()
        return int(self._lib.rv_device_bytes(self._num))

    def graph_launches(self) -> int:
        self._need_numeric()
        return int(self._lib.rv_graph_launches(self._num))

    def chain_launches(self) -> int:
        """Persistent outer-block launches enqueued so far (rv_chain_launches)."""
        self._need_numeric()
        return int(self._lib.rv_chain_launches(self._num))

    def lookahead_launches(self) -> int:
        """Outer-block trailing updates run beside the next outer block (rv_lookahead_launches)."""
        self._need_numeric()
        return int(self._lib.rv_lookahead_launches(self._num))

    def front_precision(self) -> np.ndarray:
        self._need_numeric()
        out = np.zeros(self.num_supernodes(), dtype=np.int8)
        _check(self._lib.rv_front_precision(self._num, _ptr(out, C.c_int8)))
        return out

    @staticmethod
    def solve_full(matrix: SparseColMat, rhs: np.ndarray, options: Optional[SolverOptions] = None) -> np.ndarray:
        """SparseLDLT::solve_full (solver.rs:470-494)."""
        options = options or SolverOptions()
        s = SparseLDLT.analyze_with_matrix(matrix, AnalyzeOptions(options.ordering, options.perm, options.grid))
        s.factor(matrix, FactorOptions(threshold=options.threshold, nemin=options.nemin,
                                       small_leaf_threshold=options.small_leaf_threshold))
        return s.solve(rhs)

    def solve_scratch(self, rhs_ncols: int) -> int:
        """aptp_solve_scratch (solve.rs:150), in bytes of device memory."""
        if not self._num:
            return 0
        return int(self._lib.rv_solve_scratch(self._num, rhs_ncols))

    # -- accessors ----------------------------------------------------------
    def inertia(self) -> Optional[Inertia]:
        if not self._num:
            return None
        v = np.zeros(3, dtype=np.int64)
        _check(self._lib.rv_inertia(self._num, _ptr(v, C.c_int64)))
        return Inertia(int(v[0]), int(v[1]), int(v[2]))

    def stats(self) -> Optional[FactorizationStats]:
        if not self._num:
            return None
        v = np.zeros(10, dtype=np.int64)
        _check(self._lib.rv_stats(self._num, _ptr(v, C.c_int64)))
        return FactorizationStats(*(int(t) for t in v))

    def per_supernode_stats(self) -> Optional[List[PerSupernodeStats]]:
        if not self._num:
            return None
        out = []
        v = np.zeros(8, dtype=np.int64)
        ml = C.c_double()
        for s in range(self._lib.rv_num_supernodes(self._num)):
            _check(self._lib.rv_supernode_stats(self._num, s, _ptr(v, C.c_int64), C.byref(ml)))
            out.append(PerSupernodeStats(s, int(v[0]), int(v[1]), int(v[2]), int(v[3]), int(v[4]),
                                         int(v[5]), float(ml.value), int(v[6]), bool(v[7])))
        return out

    def supernode_structure(self, s: int):
        out = np.zeros(2, dtype=np.int64)
        _check(self._lib.rv_supernode_structure(self._num, s, None, None, _ptr(out, C.c_int64)))
        owned = np.zeros(max(int(out[0]), 1), np.int32)
        pat = np.zeros(max(int(out[1]), 1), np.int32)
        _check(self._lib.rv_supernode_structure(self._num, s, _ptr(owned, C.c_int32), _ptr(pat, C.c_int32),
                                                _ptr(out, C.c_int64)))
        return owned[: int(out[0])], pat[: int(out[1])]

    def export_front(self, s: int) -> dict:
        """FrontFactors of supernode s (numeric.rs:543-596)."""
        self._need_numeric()
        v = np.zeros(8, dtype=np.int64)
        _check(self._lib.rv_supernode_stats(self._num, s, _ptr(v, C.c_int64), None))
        m, k, ne = int(v[0]), int(v[1]), int(v[2])
        r = m - ne
        l11 = np.zeros((ne, ne), order="F")
        l21 = np.zeros((r, ne), order="F")
        dk = np.zeros(max(ne, 1), np.int8)
        dg = np.zeros(max(ne, 1))
        of = np.zeros(max(ne, 1))
        ci = np.zeros(max(ne, 1), np.int32)
        ri = np.zeros(max(r, 1), np.int32)
        lp = np.zeros(max(k, 1), np.int32)
        _check(self._lib.rv_export_front(
            self._num, s, _ptr(l11, C.c_double) if ne else None, _ptr(l21, C.c_double) if ne * r else None,
            _ptr(dk, C.c_int8), _ptr(dg, C.c_double), _ptr(of, C.c_double), _ptr(ci, C.c_int32),
            _ptr(ri, C.c_int32), _ptr(lp, C.c_int32)))
        return dict(l11=l11, l21=l21, dkind=dk[:ne], ddiag=dg[:ne], doff=of[:ne], col_indices=ci[:ne],
                    row_indices=ri[:r], local_perm=lp[:k], num_eliminated=ne)

    def export_contribution(self, s: int):
        """ContributionBlock of supernode s (numeric.rs:516-523): (lower-triangle block, rows,
        num_delayed), or None for a root or a fully eliminated front."""
        self._need_numeric()
        out = np.zeros(2, dtype=np.int64)
        _check(self._lib.rv_export_contribution(self._num, s, None, None, _ptr(out, C.c_int64)))
        size = int(out[0])
        if size == 0:
            return None
        data = np.zeros((size, size), order="F")
        rows = np.zeros(size, np.int32)
        _check(self._lib.rv_export_contribution(self._num, s, _ptr(data, C.c_double), _ptr(rows, C.c_int32),
                                                _ptr(out, C.c_int64)))
        return data, rows, int(out[1])

    def set_timing(self, on: bool):
        self._need_numeric()
        _check(self._lib.rv_set_timing(self._num, 1 if on else 0))

    def timings(self) -> dict:
        self._need_numeric()
        v = np.zeros(6)
        _check(self._lib.rv_timings(self._num, _ptr(v, C.c_double)))
        return dict(schur_ms=float(v[0]), schur_flops=float(v[1]), schur_launches=int(v[2]),
                    big_ms=float(v[3]), big_flops=float(v[4]), big_launches=int(v[5]))

    def n(self) -> int:
        return self._n


def amalgamate(sn_begin, sn_end, pat_ptr, pat, parent, nemin: int = 32) -> list:
    """amalgamate (amalgamation.rs:115-231) on a bare partition: list of dicts with col_begin,
    col_end, owned (list of ranges), pattern, parent."""
    lib = load_library()
    ns = len(sn_begin)
    b = np.ascontiguousarray(sn_begin, np.int32)
    e = np.ascontiguousarray(sn_end, np.int32)
    pp = np.ascontiguousarray(pat_ptr, np.int64)
    pt = np.ascontiguousarray(pat if len(pat) else [0], np.int32)
    pa = np.ascontiguousarray(parent, np.int32)
    bo, eo, po = (np.zeros(max(ns, 1), np.int32) for _ in range(3))
    rp = np.zeros(ns + 1, np.int64)
    rl, rh = np.zeros(max(ns, 1), np.int32), np.zeros(max(ns, 1), np.int32)
    ppo = np.zeros(ns + 1, np.int64)
    pto = np.zeros(max(int(pp[-1]) if ns else 0, 1), np.int32)
    cnt = lib.rv_amalgamate(ns, _ptr(b, C.c_int32), _ptr(e, C.c_int32), _ptr(pp, C.c_int64), _ptr(pt, C.c_int32),
                            _ptr(pa, C.c_int32), nemin, _ptr(bo, C.c_int32), _ptr(eo, C.c_int32),
                            _ptr(po, C.c_int32), _ptr(rp, C.c_int64), _ptr(rl, C.c_int32), _ptr(rh, C.c_int32),
                            _ptr(ppo, C.c_int64), _ptr(pto, C.c_int32))
    if cnt < 0:
        _check(RV_INVALID_INPUT)
    return [dict(col_begin=int(bo[i]), col_end=int(eo[i]), parent=int(po[i]),
                 owned=[(int(rl[q]), int(rh[q])) for q in range(rp[i], rp[i + 1])],
                 pattern=[int(t) for t in pto[ppo[i]:ppo[i + 1]]]) for i in range(cnt)]


def dense_aptp_factor(a: np.ndarray, num_fully_summed: int, options: Optional[FactorOptions] = None) -> dict:
    """aptp_factor_in_place (factor.rs:301-404) on the GPU for a dense m x m matrix (lower
    triangle); returns the factored copy and the AptpFactorResult fields."""
    lib = load_library()
    options = options or FactorOptions()
    m = a.shape[0]
    work = np.asfortranarray(np.array(a, dtype=np.float64, copy=True))
    p = int(num_fully_summed)
    perm = np.zeros(max(m, 1), np.int32)
    dk = np.zeros(max(p, 1), np.int8)
    dg = np.zeros(max(p, 1))
    of = np.zeros(max(p, 1))
    st = np.zeros(4, np.int64)
    ml = C.c_double()
    o = options.to_c()
    _check(lib.rv_dense_aptp_factor(_ptr(work, C.c_double), m, m, p, C.byref(o), _ptr(perm, C.c_int32),
                                    _ptr(dk, C.c_int8), _ptr(dg, C.c_double), _ptr(of, C.c_double),
                                    _ptr(st, C.c_int64), C.byref(ml)))
    return dict(a=work, perm=perm[:m], dkind=dk[:p], ddiag=dg[:p], doff=of[:p], num_eliminated=int(st[0]),
                num_1x1=int(st[1]), num_2x2=int(st[2]), num_delayed=int(st[3]), max_l=float(ml.value))
