"""Learned per-front update precision (an extension: the reference's factor is FP64 only).

A small network (rv_mlp_*: 16 -> 64 -> 64 -> 3, trained and evaluated by the library's own GPU
kernels) picks each front's Schur-update arithmetic — BF16, FP32 or FP64 — from the front's
features (rv_front_features: sizes, update flops, tree position, delays, 2x2 and negative-pivot
fractions, max|L|).  Pivoting, D, the solve and the residual stay FP64; FP64 iterative
refinement brings the solution to the requested backward error, and when it does not within the
allowed steps the whole factorization falls back to FP64.

Labels come from the solver itself (`label_fronts`): on a training matrix, fronts are assigned
reduced precision below a front-order threshold, the factorization is refined, and the largest
thresholds that still reach the tolerance within `max_iter` refinement steps (with the FP64
inertia) define each front's label — the cheapest precision that keeps refinement converging.
"""
import json
import time
from dataclasses import dataclass
from typing import List, Optional, Tuple

import ctypes as C
import numpy as np

import rivrs_sparse as rs

N_FEATURES, N_HIDDEN, N_CLASSES = 16, 64, 3
N_PARAMS = N_FEATURES * N_HIDDEN + N_HIDDEN + N_HIDDEN * N_HIDDEN + N_HIDDEN + N_HIDDEN * N_CLASSES + N_CLASSES
PREC_FP64, PREC_FP32, PREC_BF16 = 0, 1, 2
CLASS_PREC = np.array([PREC_BF16, PREC_FP32, PREC_FP64], dtype=np.int8)  # class index -> RV_PREC_*
CLASS_NAMES = ("bf16", "fp32", "fp64")


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class PrecisionMLP:
    """Parameters (rv_mlp_* layout) plus the feature standardisation fitted on training data."""

    def __init__(self, params: np.ndarray, mu: np.ndarray, sd: np.ndarray):
        self.params = _f32(params)
        self.mu = _f32(mu)
        self.sd = _f32(sd)
        assert self.params.shape == (N_PARAMS,)

    @staticmethod
    def initial(seed: int = 0) -> "PrecisionMLP":
        """He-initialised weights, zero biases, identity standardisation."""
        rng = np.random.default_rng(seed)
        parts = []
        for fan_in, fan_out in ((N_FEATURES, N_HIDDEN), (N_HIDDEN, N_HIDDEN), (N_HIDDEN, N_CLASSES)):
            parts.append(rng.standard_normal(fan_in * fan_out) * np.sqrt(2.0 / fan_in))
            parts.append(np.zeros(fan_out))
        return PrecisionMLP(np.concatenate(parts), np.zeros(N_FEATURES), np.ones(N_FEATURES))

    def standardize(self, F: np.ndarray) -> np.ndarray:
        return _f32((np.asarray(F, dtype=np.float32) - self.mu) / self.sd)

    def probabilities(self, F: np.ndarray) -> np.ndarray:
        X = self.standardize(F)
        out = np.zeros((X.shape[0], N_CLASSES), dtype=np.float32)
        if X.shape[0]:
            rs._check(rs.load_library().rv_mlp_forward(_fp(self.params), X.shape[0], _fp(X), _fp(out)))
        return out

    def predict(self, F: np.ndarray) -> np.ndarray:
        return np.argmax(self.probabilities(F), axis=1).astype(np.int32)

    def loss_grad(self, F, y, class_weight=None) -> Tuple[float, np.ndarray]:
        X = self.standardize(F)
        yy = np.ascontiguousarray(y, dtype=np.int32)
        cw = None if class_weight is None else _f32(class_weight)
        loss = np.zeros(1, dtype=np.float32)
        grad = np.zeros(N_PARAMS, dtype=np.float32)
        rs._check(rs.load_library().rv_mlp_loss_grad(_fp(self.params), X.shape[0], _fp(X),
                                                     yy.ctypes.data_as(C.POINTER(C.c_int32)),
                                                     None if cw is None else _fp(cw), _fp(loss), _fp(grad)))
        return float(loss[0]), grad

    def fit(self, F, y, class_weight=None, epochs: int = 2000, lr: float = 3e-3) -> float:
        """Standardise on F, then full-batch Adam on the GPU (rv_mlp_train); returns final loss."""
        F = np.asarray(F, dtype=np.float32)
        self.mu = _f32(F.mean(axis=0))
        self.sd = _f32(np.maximum(F.std(axis=0), 1e-3))
        X = self.standardize(F)
        yy = np.ascontiguousarray(y, dtype=np.int32)
        cw = None if class_weight is None else _f32(class_weight)
        loss = np.zeros(1, dtype=np.float32)
        rs._check(rs.load_library().rv_mlp_train(_fp(self.params), X.shape[0], _fp(X),
                                                 yy.ctypes.data_as(C.POINTER(C.c_int32)),
                                                 None if cw is None else _fp(cw), int(epochs), float(lr), _fp(loss)))
        return float(loss[0])

    def to_json(self) -> dict:
        return {"layers": [N_FEATURES, N_HIDDEN, N_HIDDEN, N_CLASSES], "classes": list(CLASS_NAMES),
                "params": self.params.tolist(), "mu": self.mu.tolist(), "sd": self.sd.tolist()}

    @staticmethod
    def from_json(d: dict) -> "PrecisionMLP":
        assert d["layers"] == [N_FEATURES, N_HIDDEN, N_HIDDEN, N_CLASSES]
        return PrecisionMLP(np.array(d["params"]), np.array(d["mu"]), np.array(d["sd"]))

    def save(self, path: str, extra: Optional[dict] = None):
        d = self.to_json()
        if extra:
            d["training"] = extra
        with open(path, "w") as f:
            json.dump(d, f)

    @staticmethod
    def load(path: str) -> "PrecisionMLP":
        with open(path) as f:
            return PrecisionMLP.from_json(json.load(f))


def update_work_mask(s: rs.SparseLDLT, nb: int = 256) -> np.ndarray:
    """Fronts whose precision matters: a contribution block, or outer-block updates (k > nb)."""
    per = s.per_supernode_stats()
    m = np.array([q.front_size for q in per])
    k = np.array([q.num_fully_summed for q in per])
    root = np.array([q.parent is None or q.parent < 0 for q in per])
    return ((m - k > 0) & ~root) | (k > nb)


def front_orders(s: rs.SparseLDLT) -> np.ndarray:
    return np.array([q.front_size for q in s.per_supernode_stats()])


@dataclass
class RefineResult:
    x: np.ndarray
    corrections: int
    backward_error: float
    converged: bool
    fell_back: bool


def _refine(s, A, b, prec, tol, max_iter, inertia_ref):
    s.set_front_precision(prec)
    s.refactor(A)
    x, it, be = s.solve_refine(b, tol=tol, max_iter=max_iter)
    inert = s.inertia()
    ok = be <= tol and (inertia_ref is None or (inert.positive, inert.negative, inert.zero) == inertia_ref)
    return x, it, be, ok


def assign(s: rs.SparseLDLT, model: PrecisionMLP, mask: Optional[np.ndarray] = None) -> np.ndarray:
    """Per-front RV_PREC_* codes chosen by the network from the last factorization's features
    (fronts outside `mask`, default update_work_mask(s), stay FP64)."""
    F = s.front_features()
    prec = CLASS_PREC[model.predict(F)]
    prec[~(update_work_mask(s) if mask is None else mask)] = PREC_FP64
    return prec


def solve_learned(s: rs.SparseLDLT, A: rs.SparseColMat, b: np.ndarray, model: PrecisionMLP,
                  tol: float = 1e-12, max_iter: int = 3, inertia_ref=None,
                  prec: Optional[np.ndarray] = None) -> RefineResult:
    """Factor with the network's per-front precisions (s must hold a factorization of a matrix
    with A's pattern, whose features are used), refine; FP64 fallback when refinement does not
    reach tol within max_iter steps."""
    if prec is None:
        prec = assign(s, model)
    x, it, be, ok = _refine(s, A, b, prec, tol, max_iter, inertia_ref)
    if ok:
        return RefineResult(x, it, be, True, False)
    x, it, be, ok = _refine(s, A, b, np.zeros_like(prec), tol, max_iter, inertia_ref)
    return RefineResult(x, it, be, ok, True)


def label_fronts(s: rs.SparseLDLT, A: rs.SparseColMat, b: np.ndarray, tol: float = 1e-12, max_iter: int = 3,
                 log=None) -> Tuple[np.ndarray, np.ndarray, dict]:
    """Solver-generated labels for every front with update work of one matrix.

    After an FP64 factorization (its inertia is the reference), two bisections over the sorted
    front orders find the largest t32 such that FP32 on fronts of order <= t32 (FP64 above) still
    refines to tol within max_iter steps, then the largest t16 <= t32 such that BF16 on fronts of
    order <= t16 (FP32 up to t32) does.  Returns (features, class labels, info) of the fronts with
    update work; features are those of the FP64 factorization."""
    t0 = time.time()
    opts = getattr(s, "_options", None)
    s.set_front_precision(None)
    s.refactor(A, opts)
    s.set_front_precision(np.zeros(s.num_supernodes(), dtype=np.int8))
    s.refactor(A, opts)
    inert = s.inertia()
    iref = (inert.positive, inert.negative, inert.zero)
    F = s.front_features()
    work = update_work_mask(s)
    m = front_orders(s)
    sizes = np.unique(m[work])
    calls = [0]

    def ok(t32, t16):
        calls[0] += 1
        prec = np.zeros(len(m), dtype=np.int8)
        prec[work & (m <= t32)] = PREC_FP32
        prec[work & (m <= t16)] = PREC_BF16
        return _refine(s, A, b, prec, tol, max_iter, iref)[3]

    def largest(test, cand):
        # largest index i of cand with test(cand[i]) (assumed monotone), -1 if none
        if len(cand) == 0:
            return -1
        if test(cand[-1]):
            return len(cand) - 1
        lo, hi = -1, len(cand) - 1
        while hi - lo > 1:
            mid = (lo + hi) // 2
            if test(cand[mid]):
                lo = mid
            else:
                hi = mid
        return lo

    i32 = largest(lambda t: ok(t, -1), sizes)
    t32 = sizes[i32] if i32 >= 0 else -1
    i16 = largest(lambda t: ok(t32, t), sizes[:i32 + 1]) if i32 >= 0 else -1
    t16 = sizes[i16] if i16 >= 0 else -1
    y = np.full(len(m), 2, dtype=np.int32)
    y[m <= t32] = 1
    y[m <= t16] = 0
    s.set_front_precision(None)
    info = dict(t32=int(t32), t16=int(t16), fronts=int(work.sum()), refactorizations=calls[0] + 2,
                seconds=time.time() - t0, inertia=iref)
    if log:
        log(info)
    return F[work], y[work], info
