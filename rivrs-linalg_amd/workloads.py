"""Seeded synthetic systems of the benchmark configurations (full-symmetric CSC, FP64).

  kkt(...)        saddle-point [[H, A^T], [A, -delta*I]] with H the h^-2-scaled 2-D 5-point
                  Laplacian on a side x side grid plus diag N(0, noise^2) and A with three
                  distinct uniform columns per row, N(0,1) values.
  helmholtz(...)  3-D 7-point Laplacian on an N^3 grid (diag 6, off-diagonals
                  -(1 + 0.01 u), u ~ U(-1,1), symmetric) minus sigma*I with sigma in the
                  spectral gap nearest the requested fraction of negative eigenvalues of the
                  unjittered operator (closed form lambda = sum_d 2 - 2 cos(pi i_d/(N+1))).
Grid unknowns are numbered x-fastest, which RV_ORDER_ND_GRID expects.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from rivrs_sparse import SparseColMat


def _lap1d(n):
    return sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1], format="csr")


def laplacian_2d(side: int):
    I = sp.identity(side, format="csr")
    T = _lap1d(side)
    return (sp.kron(I, T) + sp.kron(T, I)).tocsr()


def laplacian_3d(N: int):
    I = sp.identity(N, format="csr")
    T = _lap1d(N)
    return (sp.kron(sp.kron(I, I), T) + sp.kron(sp.kron(I, T), I) + sp.kron(sp.kron(T, I), I)).tocsr()


def helmholtz_sigma(N: int, frac: float = 0.05):
    """Shift in a spectral gap so that about frac*N^3 eigenvalues of L - sigma*I are negative.
    Returns (sigma, number of negative eigenvalues of the unjittered operator)."""
    t = 2.0 - 2.0 * np.cos(np.pi * np.arange(1, N + 1) / (N + 1))
    lam = np.sort((t[:, None, None] + t[None, :, None] + t[None, None, :]).ravel())
    k = int(round(frac * N ** 3))
    # first strict gap at or after position k
    while k < len(lam) and lam[k] - lam[k - 1] <= 1e-12 * lam[-1]:
        k += 1
    return 0.5 * (lam[k - 1] + lam[k]), k


def helmholtz(N: int, seed: int = 0, frac: float = 0.05, jitter: float = 0.01):
    """Indefinite Helmholtz-type system on an N^3 grid; returns (matrix, rhs, info)."""
    rng = np.random.default_rng(seed)
    L = sp.triu(laplacian_3d(N), k=1).tocoo()
    u = rng.uniform(-1.0, 1.0, size=L.nnz)
    off = -(1.0 + jitter * u)
    n = N ** 3
    sigma, nneg = helmholtz_sigma(N, frac)
    rows = np.concatenate([L.row, L.col, np.arange(n)])
    cols = np.concatenate([L.col, L.row, np.arange(n)])
    vals = np.concatenate([off, off, np.full(n, 6.0 - sigma)])
    K = sp.csc_matrix((vals, (rows, cols)), shape=(n, n))
    b = rng.standard_normal(n)
    return SparseColMat.from_scipy(K), b, dict(sigma=sigma, expected_negative=nneg, grid=(N, N, N))


def kkt(side: int = 64, m: int | None = None, seed: int = 0, noise: float = 0.1, delta: float = 1e-8):
    """Saddle-point system; returns (matrix, rhs, x_true, info).  With the h^-2 scaling H is
    positive definite, so the inertia is (side^2, m, 0) by Haynsworth's formula."""
    rng = np.random.default_rng(seed)
    nh = side * side
    if m is None:
        m = nh // 4
    h = 1.0 / (side + 1)
    H = laplacian_2d(side) / (h * h) + sp.diags(noise * rng.standard_normal(nh))
    cols = np.stack([rng.choice(nh, size=3, replace=False) for _ in range(m)])
    vals = rng.standard_normal((m, 3))
    A = sp.csr_matrix((vals.ravel(), (np.repeat(np.arange(m), 3), cols.ravel())), shape=(m, nh))
    K = sp.bmat([[H, A.T], [A, -delta * sp.identity(m)]], format="csc")
    x = rng.standard_normal(nh + m)
    b = K @ x
    return SparseColMat.from_scipy(K), b, x, dict(expected_inertia=(nh, m, 0), side=side, m=m)
