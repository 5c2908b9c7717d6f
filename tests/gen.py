"""Test-matrix generators restated from the reference's tests, plus dense helpers.

deterministic_indefinite_matrix / cause_delays follow the LCG helpers in
sparse/src/symmetric/factor.rs:4886-4944 (pure u64 integer arithmetic, restated exactly).
"""
import json
import os

import numpy as np

M64 = (1 << 64) - 1
MUL = 6364136223846793005
INC = 1442695040888963407
U32MAX = 4294967295.0


def _lcg_hash(seed, a, b):
    s = (seed + a * 10007 + b * 7) & M64
    s = (s * MUL + INC) & M64
    s = (s * MUL + INC) & M64
    return (s >> 33) / (U32MAX / 2.0) - 1.0


def deterministic_indefinite_matrix(n, seed, diag_scale):
    """factor.rs:4886-4909: entries in [-1, 1], diagonal scaled by diag_scale."""
    a = np.zeros((n, n))
    for i in range(n):
        a[i, i] = _lcg_hash(seed, i, i) * diag_scale
        for j in range(i):
            v = _lcg_hash(seed, j, i + n)
            a[i, j] = v
            a[j, i] = v
    return a


def deterministic_indefinite_matrix_fast(n, seed, diag_scale):
    """deterministic_indefinite_matrix with the hash evaluated on uint64 arrays (numpy wraps
    uint64 products modulo 2**64, which is the reference's wrapping_mul); same values."""
    with np.errstate(over="ignore"):
        i = np.arange(n, dtype=np.uint64)
        mul, inc = np.uint64(MUL), np.uint64(INC)
        s = np.uint64(seed) + i[None, :] * np.uint64(10007) + (i[:, None] + np.uint64(n)) * np.uint64(7)
        s = s * mul + inc
        s = s * mul + inc
        off = (s >> np.uint64(33)).astype(np.float64) / (U32MAX / 2.0) - 1.0  # off[i, j]: a = j, b = i + n
        d = np.uint64(seed) + i * np.uint64(10007) + i * np.uint64(7)
        d = d * mul + inc
        d = d * mul + inc
        diag = ((d >> np.uint64(33)).astype(np.float64) / (U32MAX / 2.0) - 1.0) * diag_scale
    a = np.tril(off, -1)
    a = a + a.T
    a[np.diag_indices(n)] = diag
    return a


def cause_delays(a, seed, scale):
    """factor.rs:4914-4944: scale n/8 pseudo-random rows and columns by `scale`."""
    n = a.shape[0]
    n_scaled = max(n // 8, 1)
    state = seed
    rows = []
    while len(rows) < n_scaled:
        state = (state * MUL + INC) & M64
        idx = (state >> 33) % n
        if idx not in rows:
            rows.append(idx)
    a = a.copy()
    for r in rows:
        a[r, :] *= scale
        a[:, r] *= scale
    return a


def extract_l(a, dkind, ne):
    """extract_l (factor.rs:656-698): unit lower L from the in-place factor; for a 2x2 pivot the
    (col+1, col) entry is D, not L."""
    n = a.shape[0]
    L = np.eye(n)
    c = 0
    while c < ne:
        if dkind[c] == 2:
            L[c + 2:, c] = a[c + 2:, c]
            L[c + 2:, c + 1] = a[c + 2:, c + 1]
            c += 2
        else:
            L[c + 1:, c] = a[c + 1:, c]
            c += 1
    return L


def d_matrix(dkind, ddiag, doff, ne):
    D = np.zeros((ne, ne))
    c = 0
    while c < ne:
        if dkind[c] == 2:
            D[c, c] = ddiag[c]
            D[c + 1, c + 1] = ddiag[c + 1]
            D[c, c + 1] = D[c + 1, c] = doff[c]
            c += 2
        else:
            D[c, c] = ddiag[c]
            c += 1
    return D


def dense_reconstruction_error(a0, res):
    """||P^T A P - L D L^T|| / ||A|| for a full factorization (p == m, nothing delayed)."""
    ne = res["num_eliminated"]
    n = a0.shape[0]
    assert ne == n
    L = extract_l(res["a"], res["dkind"], ne)
    D = d_matrix(res["dkind"], res["ddiag"], res["doff"], ne)
    p = res["perm"]
    pap = a0[np.ix_(p, p)]
    return np.linalg.norm(pap - L @ D @ L.T) / max(np.linalg.norm(a0), 1e-300)


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def hand_constructed():
    d = os.path.join(GOLDEN, "hand-constructed")
    names = sorted(f[:-4] for f in os.listdir(d) if f.endswith(".mtx"))
    return [(n, os.path.join(d, n + ".mtx"), json.load(open(os.path.join(d, n + ".json")))) for n in names]


def suitesparse_ci():
    d = os.path.join(GOLDEN, "suitesparse-ci")
    return sorted((f.split(".")[0], os.path.join(d, f)) for f in os.listdir(d) if f.endswith(".mtx.gz"))
