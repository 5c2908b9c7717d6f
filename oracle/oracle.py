"""TEST INFRASTRUCTURE ONLY: ctypes loader for the CPU parity oracle (liboracle.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this module.  The
oracle restates rivrs-sparse's numeric phase on the CPU (see rivrs_oracle.cpp for the
file:line each function follows); it is the checker, never the product.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


class Options(C.Structure):
    _fields_ = [
        ("threshold", C.c_double),
        ("small", C.c_double),
        ("outer_block_size", C.c_int32),
        ("inner_block_size", C.c_int32),
        ("failed_pivot_tpp", C.c_int32),
        ("nemin", C.c_int32),
        ("small_leaf_threshold", C.c_int32),
    ]


def options(threshold=0.01, small=1e-20, nb=256, ib=32, tpp=True, nemin=32, small_leaf=256):
    o = Options()
    o.threshold = threshold
    o.small = small
    o.outer_block_size = nb
    o.inner_block_size = ib
    o.failed_pivot_tpp = 1 if tpp else 0
    o.nemin = nemin
    o.small_leaf_threshold = small_leaf
    return o


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P = C.c_void_p
        I64P, I32P, I8P, DP = (C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_int8),
                               C.POINTER(C.c_double))
        L.orc_aptp_factor_in_place.restype = C.c_int
        L.orc_aptp_factor_in_place.argtypes = [DP, C.c_int64, C.c_int32, C.c_int32, C.POINTER(Options), I32P,
                                               I8P, DP, DP, I64P, DP]
        L.orc_factor.restype = P
        L.orc_factor.argtypes = [C.c_int64, I64P, I64P, DP, C.c_int32, I32P, I32P, I64P, I32P, I32P, I32P, DP,
                                 C.POINTER(Options)]
        L.orc_solve.restype = C.c_int
        L.orc_solve.argtypes = [P, DP, C.c_int64, C.c_int32]
        L.orc_inertia.argtypes = [P, I64P]
        L.orc_stats.argtypes = [P, I64P]
        L.orc_num_supernodes.restype = C.c_int32
        L.orc_num_supernodes.argtypes = [P]
        L.orc_supernode_info.argtypes = [P, C.c_int32, I64P, DP]
        L.orc_supernode_structure.argtypes = [P, C.c_int32, I32P, I32P, I64P]
        L.orc_front_factors.argtypes = [P, C.c_int32, DP, DP, I8P, DP, DP, I32P, I32P, I32P]
        L.orc_free.argtypes = [P]
        L.orc_backward_error.restype = C.c_double
        L.orc_backward_error.argtypes = [C.c_int64, I64P, I64P, DP, DP, DP]
        L.orc_set_keep_contributions.argtypes = [C.c_int32]
        L.orc_contribution.restype = C.c_int64
        L.orc_contribution.argtypes = [P, C.c_int32, DP, I32P, I64P]
        L.orc_extend_add_mapped.argtypes = [DP, C.c_int64, DP, C.c_int32, C.POINTER(C.c_uint32)]
        L.orc_set_threads.argtypes = [C.c_int32]
        L.orc_max_threads.restype = C.c_int32
        L.orc_amalgamate.restype = C.c_int32
        L.orc_amalgamate.argtypes = [C.c_int32, I32P, I32P, I64P, I32P, I32P, C.c_int32, I32P, I32P, I32P,
                                     I64P, I32P, I32P, I64P, I32P]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def dense_factor(a, p, opts=None):
    """orc_aptp_factor_in_place on a copy of the dense matrix a (lower triangle)."""
    L = lib()
    m = a.shape[0]
    w = np.asfortranarray(np.array(a, dtype=np.float64, copy=True))
    perm = np.zeros(max(m, 1), np.int32)
    dk = np.zeros(max(p, 1), np.int8)
    dg = np.zeros(max(p, 1))
    of = np.zeros(max(p, 1))
    st = np.zeros(4, np.int64)
    ml = C.c_double()
    o = opts or options()
    rc = L.orc_aptp_factor_in_place(_p(w, C.c_double), m, m, p, C.byref(o), _p(perm, C.c_int32),
                                    _p(dk, C.c_int8), _p(dg, C.c_double), _p(of, C.c_double),
                                    _p(st, C.c_int64), C.byref(ml))
    if rc != 0:
        raise ValueError("invalid input")
    return dict(a=w, perm=perm[:m], dkind=dk[:p], ddiag=dg[:p], doff=of[:p], num_eliminated=int(st[0]),
                num_1x1=int(st[1]), num_2x2=int(st[2]), num_delayed=int(st[3]), max_l=float(ml.value))


def set_threads(n):
    """Thread count of later oracle calls (results do not depend on it)."""
    lib().orc_set_threads(int(n))


def max_threads():
    return int(lib().orc_max_threads())


def extend_add_mapped(parent, child, row_map):
    """orc_extend_add_mapped: parent (square, modified copy returned) += child's lower triangle
    scattered through row_map (extend_add_mapped, numeric.rs:2507-2549)."""
    P_ = np.asfortranarray(np.array(parent, dtype=np.float64, copy=True))
    Cm = np.asfortranarray(np.array(child, dtype=np.float64))
    mp = np.ascontiguousarray(row_map, np.uint32)
    lib().orc_extend_add_mapped(_p(P_, C.c_double), P_.shape[0], _p(Cm, C.c_double), Cm.shape[0],
                                mp.ctypes.data_as(C.POINTER(C.c_uint32)))
    return P_


class OracleFactor:
    """AptpNumeric::factor + SparseLDLT::solve_in_place on the CPU.  keep_contributions=True keeps
    every front's contribution block for contribution()."""

    def __init__(self, matrix, structure, scaling=None, opts=None, keep_contributions=False):
        L = lib()
        L.orc_set_keep_contributions(1 if keep_contributions else 0)
        self.L = L
        self.n = matrix.n
        self._keep = [np.ascontiguousarray(matrix.col_ptr, np.int64), np.ascontiguousarray(matrix.row_idx, np.int64),
                      np.ascontiguousarray(matrix.values, np.float64)]
        st = structure
        self._st = [np.ascontiguousarray(st.sn_begin, np.int32), np.ascontiguousarray(st.sn_end, np.int32),
                    np.ascontiguousarray(st.pat_ptr, np.int64),
                    np.ascontiguousarray(st.pat if len(st.pat) else np.zeros(1), np.int32),
                    np.ascontiguousarray(st.parent, np.int32), np.ascontiguousarray(st.perm_fwd, np.int32)]
        self._sc = None if scaling is None else np.ascontiguousarray(scaling, np.float64)
        o = opts or options()
        cp, ri, v = self._keep
        b, e, pp, pt, pa, pf = self._st
        self.h = L.orc_factor(self.n, _p(cp, C.c_int64), _p(ri, C.c_int64), _p(v, C.c_double), len(b),
                              _p(b, C.c_int32), _p(e, C.c_int32), _p(pp, C.c_int64), _p(pt, C.c_int32),
                              _p(pa, C.c_int32), _p(pf, C.c_int32),
                              None if self._sc is None else _p(self._sc, C.c_double), C.byref(o))
        L.orc_set_keep_contributions(0)
        if not self.h:
            raise ValueError("oracle factor rejected the options")

    def contribution(self, s):
        """(lower-triangle block size x size, rows, num_delayed) of supernode s, or None."""
        nd = C.c_int64()
        size = int(self.L.orc_contribution(self.h, s, None, None, C.byref(nd)))
        if size < 0:
            return None
        data = np.zeros((size, size), order="F")
        rows = np.zeros(max(size, 1), np.int32)
        self.L.orc_contribution(self.h, s, _p(data, C.c_double) if size else None, _p(rows, C.c_int32),
                                C.byref(nd))
        return data, rows[:size], int(nd.value)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_free(self.h)
            self.h = None

    def solve(self, b):
        x = np.asfortranarray(np.array(b, dtype=np.float64, copy=True).reshape(self.n, -1))
        self.L.orc_solve(self.h, _p(x, C.c_double), self.n, x.shape[1])
        return x.reshape(np.shape(b))

    def inertia(self):
        v = np.zeros(3, np.int64)
        self.L.orc_inertia(self.h, _p(v, C.c_int64))
        return tuple(int(t) for t in v)

    def stats(self):
        v = np.zeros(10, np.int64)
        self.L.orc_stats(self.h, _p(v, C.c_int64))
        keys = ["total_1x1_pivots", "total_2x2_pivots", "total_delayed", "zero_pivots", "max_front_size",
                "supernodes_before_amalgamation", "supernodes_after_amalgamation", "merges_performed",
                "small_leaf_subtrees", "small_leaf_nodes"]
        return dict(zip(keys, (int(t) for t in v)))

    def num_supernodes(self):
        return int(self.L.orc_num_supernodes(self.h))

    def supernode(self, s):
        v = np.zeros(8, np.int64)
        ml = C.c_double()
        self.L.orc_supernode_info(self.h, s, _p(v, C.c_int64), C.byref(ml))
        return dict(front_size=int(v[0]), num_fully_summed=int(v[1]), num_eliminated=int(v[2]),
                    num_delayed=int(v[3]), num_1x1=int(v[4]), num_2x2=int(v[5]), parent=int(v[6]),
                    in_small_leaf=bool(v[7]), max_l_entry=float(ml.value))

    def structure(self, s):
        out = np.zeros(2, np.int64)
        self.L.orc_supernode_structure(self.h, s, None, None, _p(out, C.c_int64))
        ow = np.zeros(max(int(out[0]), 1), np.int32)
        pt = np.zeros(max(int(out[1]), 1), np.int32)
        self.L.orc_supernode_structure(self.h, s, _p(ow, C.c_int32), _p(pt, C.c_int32), _p(out, C.c_int64))
        return ow[: int(out[0])], pt[: int(out[1])]

    def front(self, s):
        info = self.supernode(s)
        m, k, ne = info["front_size"], info["num_fully_summed"], info["num_eliminated"]
        r = m - ne
        l11 = np.zeros((ne, ne), order="F")
        l21 = np.zeros((r, ne), order="F")
        dk = np.zeros(max(ne, 1), np.int8)
        dg = np.zeros(max(ne, 1))
        of = np.zeros(max(ne, 1))
        ci = np.zeros(max(ne, 1), np.int32)
        ri = np.zeros(max(r, 1), np.int32)
        lp = np.zeros(max(k, 1), np.int32)
        self.L.orc_front_factors(self.h, s, _p(l11, C.c_double) if ne else None,
                                 _p(l21, C.c_double) if ne * r else None, _p(dk, C.c_int8), _p(dg, C.c_double),
                                 _p(of, C.c_double), _p(ci, C.c_int32), _p(ri, C.c_int32), _p(lp, C.c_int32))
        return dict(l11=l11, l21=l21, dkind=dk[:ne], ddiag=dg[:ne], doff=of[:ne], col_indices=ci[:ne],
                    row_indices=ri[:r], local_perm=lp[:k], num_eliminated=ne)


def backward_error(matrix, x, b):
    L = lib()
    cp = np.ascontiguousarray(matrix.col_ptr, np.int64)
    ri = np.ascontiguousarray(matrix.row_idx, np.int64)
    v = np.ascontiguousarray(matrix.values, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return float(L.orc_backward_error(matrix.n, _p(cp, C.c_int64), _p(ri, C.c_int64), _p(v, C.c_double),
                                      _p(x, C.c_double), _p(b, C.c_double)))


def amalgamate(begin, end, pat_ptr, pat, parent, nemin):
    """orc_amalgamate: returns list of dicts (col_begin, col_end, owned ranges, pattern, parent)."""
    L = lib()
    ns = len(begin)
    b = np.ascontiguousarray(begin, np.int32)
    e = np.ascontiguousarray(end, np.int32)
    pp = np.ascontiguousarray(pat_ptr, np.int64)
    pt = np.ascontiguousarray(pat if len(pat) else [0], np.int32)
    pa = np.ascontiguousarray(parent, np.int32)
    bo = np.zeros(max(ns, 1), np.int32)
    eo = np.zeros(max(ns, 1), np.int32)
    po = np.zeros(max(ns, 1), np.int32)
    rp = np.zeros(ns + 1, np.int64)
    rl = np.zeros(max(ns, 1), np.int32)
    rh = np.zeros(max(ns, 1), np.int32)
    ppo = np.zeros(ns + 1, np.int64)
    pto = np.zeros(max(int(pp[-1]) if ns else 0, 1), np.int32)
    cnt = L.orc_amalgamate(ns, _p(b, C.c_int32), _p(e, C.c_int32), _p(pp, C.c_int64), _p(pt, C.c_int32),
                           _p(pa, C.c_int32), nemin, _p(bo, C.c_int32), _p(eo, C.c_int32), _p(po, C.c_int32),
                           _p(rp, C.c_int64), _p(rl, C.c_int32), _p(rh, C.c_int32), _p(ppo, C.c_int64),
                           _p(pto, C.c_int32))
    out = []
    for i in range(cnt):
        out.append(dict(col_begin=int(bo[i]), col_end=int(eo[i]), parent=int(po[i]),
                        owned=[(int(rl[q]), int(rh[q])) for q in range(rp[i], rp[i + 1])],
                        pattern=[int(t) for t in pto[ppo[i]:ppo[i + 1]]]))
    return out
