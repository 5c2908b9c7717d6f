"""Batched factorization of many independent symmetric indefinite systems on one GPU.

The systems are stacked into one block-diagonal matrix.  Each keeps its own MC64 scaling and
matching-based ordering (rv_match_order), so its supernodes, pivots and factors are those it would
have alone; the union of their assembly trees is a forest, and the level-set schedule then puts
the fronts of equal height of all systems into the same launches — every level is one batched
launch sequence over all systems.  Refactorizations on unchanged plans replay the whole level
sequence as one hipGraph launch (rv_set_graph_mode).  The solve takes all systems' right-hand
sides as one multi-column block.

    python rivrs-linalg_amd/batched.py --systems 2000 --nrhs 64     # throughput and latency
"""
import argparse
import json
import os
import sys
import time
from typing import List, Sequence

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import ctypes as C  # noqa: E402

import numpy as np  # noqa: E402

import rivrs_sparse as rs  # noqa: E402
import workloads  # noqa: E402


class DeviceArray:
    """An FP64 array in device memory (inputs resident in HBM for the timed refactor/solve)."""
    _hip = None

    def __init__(self, host: np.ndarray):
        if DeviceArray._hip is None:
            h = C.CDLL("libamdhip64.so")
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipFree.argtypes = [C.c_void_p]
            DeviceArray._hip = h
        a = np.asfortranarray(host, dtype=np.float64)
        self.shape, self.nbytes = a.shape, a.nbytes
        self.ptr = C.c_void_p()
        if self._hip.hipMalloc(C.byref(self.ptr), max(a.nbytes, 8)) != 0:
            raise MemoryError("device allocation of %d bytes failed" % a.nbytes)
        self.upload(a)

    def upload(self, a):
        a = np.asfortranarray(a, dtype=np.float64)
        assert a.nbytes == self.nbytes
        assert self._hip.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, 1) == 0

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float64, order="F")
        assert self._hip.hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2) == 0
        return out

    def __del__(self):
        if getattr(self, "ptr", None) and self.ptr.value and self._hip is not None:
            self._hip.hipFree(self.ptr)


class BatchedLDLT:
    """One numeric handle over the block-diagonal union of `systems`."""

    def __init__(self, systems: Sequence[rs.SparseColMat], graph: bool = True):
        self.sizes = np.array([A.n for A in systems], dtype=np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)])
        n = int(self.offsets[-1])
        nnz = np.array([A.nnz for A in systems], dtype=np.int64)
        noff = np.concatenate([[0], np.cumsum(nnz)])
        cp = np.empty(n + 1, dtype=np.int64)
        ri = np.empty(int(noff[-1]), dtype=np.int64)
        order = np.empty(n, dtype=np.int32)
        scaling = np.empty(n, dtype=np.float64)
        for i, A in enumerate(systems):
            o, q = int(self.offsets[i]), int(noff[i])
            cp[o:o + A.n] = A.col_ptr[:-1] + q
            ri[q:q + A.nnz] = A.row_idx + o
            mo = rs.match_order_metis(A)  # per-system MC64 + ordering (metis.rs:118-190)
            order[o:o + A.n] = mo.ordering + o
            scaling[o:o + A.n] = mo.scaling
        cp[n] = noff[-1]
        self._val_off = noff
        self.pattern = rs.SparseColMat(n, cp, ri, np.zeros(int(noff[-1])))
        self.solver = rs.SparseLDLT.analyze(self.pattern, rs.AnalyzeOptions(rs.OrderingStrategy.UserSupplied,
                                                                           perm=order))
        # elimination-order scaling (solver.rs:276-279)
        self.solver._scaling = np.ascontiguousarray(scaling[np.asarray(self.solver.structure().perm_fwd)])
        self.graph = graph
        self.n = n

    def values(self, systems: Sequence[rs.SparseColMat]) -> np.ndarray:
        return np.concatenate([A.values for A in systems])

    def factor(self, systems: Sequence[rs.SparseColMat]):
        M = rs.SparseColMat(self.n, self.pattern.col_ptr, self.pattern.row_idx, self.values(systems))
        if self.solver._num:
            self.solver.refactor(M)
        else:
            self.solver.factor(M)
            if self.graph:
                self.solver.set_graph_mode(True)
        return M

    def refactor_device(self, d_values: "DeviceArray"):
        """Refactorization from values already in device memory (rv_refactor_device)."""
        rs._check(rs.load_library().rv_refactor_device(self.solver._num, d_values.ptr))

    def solve_device(self, d_rhs: "DeviceArray"):
        """In-place solve of the stacked right-hand sides (n_total x nrhs) in device memory."""
        nrhs = d_rhs.shape[1] if len(d_rhs.shape) > 1 else 1
        rs._check(rs.load_library().rv_solve_device(self.solver._num, d_rhs.ptr, self.n, nrhs))

    def solve(self, rhs: Sequence[np.ndarray]) -> List[np.ndarray]:
        """rhs[i]: (n_i, nrhs) for system i; returns the solutions in the same layout."""
        B = np.asfortranarray(np.concatenate([np.asarray(b, dtype=np.float64).reshape(b.shape[0], -1)
                                              for b in rhs]))
        X = self.solver.solve(B)
        return [X[self.offsets[i]:self.offsets[i + 1]] for i in range(len(self.sizes))]


def kkt_batch(count: int, seed: int = 0, lo: int = 14, hi: int = 70):
    """`count` seeded KKT saddle-point systems, grid side uniform in [lo, hi], A with side^2/4 rows."""
    rng = np.random.default_rng(seed)
    sides = rng.integers(lo, hi + 1, count)
    out = []
    for i, s in enumerate(sides):
        A, _b, _x, info = workloads.kkt(int(s), seed=seed * 100003 + i)
        out.append((A, info))
    return out


def run_group(systems, rhs, steps):
    """Batched factor + solve of one group of systems, values and right-hand sides resident in
    device memory; returns (seconds of every timed step, solutions, solver)."""
    B = BatchedLDLT(systems)
    B.factor(systems)  # plans the levels
    d_val = DeviceArray(B.values(systems))
    Bh = np.asfortranarray(np.concatenate(rhs))
    d_rhs = DeviceArray(Bh)
    B.refactor_device(d_val)  # captures the level sequence
    times = []
    for _ in range(steps):
        d_rhs.upload(Bh)
        ta = time.perf_counter()
        B.refactor_device(d_val)
        B.solve_device(d_rhs)  # returns once its stream has finished
        times.append(time.perf_counter() - ta)
    Xall = d_rhs.download()
    X = [Xall[B.offsets[i]:B.offsets[i + 1]] for i in range(len(systems))]
    return times, X, B


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--systems", type=int, default=2000)
    ap.add_argument("--group", type=int, default=2000,
                    help="systems per batched factorization (working fronts live from their level to "
                         "their parent's, so 2000 of these KKT systems fit one factorization)")
    ap.add_argument("--nrhs", type=int, default=64)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--latency-sample", type=int, default=100)
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args()
    print(json.dumps(run(args)), flush=True)


def run(args) -> dict:
    """Throughput of `args.systems` seeded KKT systems factored and solved in groups of
    `args.group` (one block-diagonal factorization each), plus the single-system latency of a
    sample; returns the result record."""
    t0 = time.time()
    batch = kkt_batch(args.systems, args.seed)
    systems = [A for A, _ in batch]
    rng = np.random.default_rng(args.seed + 1)
    rhs = [rng.standard_normal((A.n, args.nrhs)) for A in systems]
    print("generated %d systems in %.1f s" % (len(systems), time.time() - t0), file=sys.stderr, flush=True)
    total = total_min = 0.0
    inert = np.zeros(3, dtype=np.int64)
    be = 0.0
    groups = []
    step_times = []
    replays = 0
    gbytes = 0.0
    for g0 in range(0, len(systems), args.group):
        sl = slice(g0, min(g0 + args.group, len(systems)))
        times, X, B = run_group(systems[sl], rhs[sl], args.steps)
        sec = float(np.median(times))
        step_times.append(times)
        total += sec
        total_min += min(times)
        i = B.solver.inertia()
        inert += np.array([i.positive, i.negative, i.zero])
        # backward error of the first, a middle and the last right-hand side of every system
        cols = sorted({0, args.nrhs // 2, args.nrhs - 1})
        for k, A in enumerate(systems[sl]):
            for q in cols:
                be = max(be, rs.sparse_backward_error(A, X[k][:, q], rhs[sl][k][:, q]))
        replays += B.solver.graph_launches()
        gbytes = max(gbytes, B.solver.device_bytes() / 1e9)
        groups.append({"systems": len(X), "n_total": int(B.n), "sec_median": sec, "sec_min": min(times),
                       "sec_p99": float(np.percentile(times, 99)), "steps": len(times),
                       "device_gbytes": B.solver.device_bytes() / 1e9})
        print("group %d: %d systems, n=%d, %.3f s, %.1f GB device" % (
            len(groups), len(X), B.n, sec, B.solver.device_bytes() / 1e9), file=sys.stderr, flush=True)
        del B, X
    exp = np.sum([info["expected_inertia"] for _, info in batch], axis=0)
    # per-system latency: each system of a sample alone (own handle, graph replay), factor + solve
    lat = []
    for i in range(min(args.latency_sample, len(systems))):
        A = systems[i]
        s = rs.SparseLDLT.analyze_with_matrix(A, rs.AnalyzeOptions(rs.OrderingStrategy.MatchOrderMetis))
        s.factor(A)
        s.set_graph_mode(True)
        s.refactor(A)
        best = None
        for _ in range(3):
            ta = time.perf_counter()
            s.refactor(A)
            s.solve(rhs[i])
            dt = time.perf_counter() - ta
            best = dt if best is None else min(best, dt)
        lat.append(best)
    out = {"workload": "batched KKT systems (config-0 construction, side U{14..70}, A rows side^2/4), MC64 + "
                       "matching-based ordering per system, FP64, %d RHS each" % args.nrhs,
           "systems": args.systems, "group": args.group, "groups": groups, "nrhs": args.nrhs,
           "statistic": "per group the median (and min) over %d timed steps; totals sum the groups" % args.steps,
           "batched_sec": total, "systems_per_sec": args.systems / total,
           # one group: every system of the batch completes with the step, so the step time is
           # each system's latency within the batch
           "batch_step_ms": {"p50": 1e3 * float(np.percentile(step_times[0], 50)),
                             "p99": 1e3 * float(np.percentile(step_times[0], 99)),
                             "steps": len(step_times[0])} if len(step_times) == 1 else None,
           "batched_sec_min": total_min, "systems_per_sec_best": args.systems / total_min,
           "graph_replays": replays, "max_device_gbytes": gbytes,
           "inertia": inert.tolist(), "expected_inertia": exp.tolist(),
           "max_backward_error": be,
           "backward_error_columns": "first, middle and last right-hand side of every system",
           "single_system_latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)),
                                        "p99": 1e3 * float(np.percentile(lat, 99)),
                                        "sample": len(lat)} if lat else None,
           "single_system_systems_per_sec": len(lat) / float(np.sum(lat)) if lat else None}
    return out


if __name__ == "__main__":
    main()
