"""Deterministic test matrices of the crate's own tests, as column-major flat float64 arrays.

Each helper cites where the crate defines the same matrix.
"""
from __future__ import annotations

import numpy as np


def cm(a2d) -> np.ndarray:
    """Column-major flat copy of a 2-D array (numpy row/col indexing)."""
    return np.asarray(a2d, dtype=np.float64).ravel(order="F").copy()


def mat(flat, nrow: int, ncol: int) -> np.ndarray:
    """2-D view of a column-major flat array."""
    return np.asarray(flat, dtype=np.float64).reshape((ncol, nrow)).T


def sample4() -> np.ndarray:
    """src/lapack/mod.rs:463-471: [[4,1,2,1],[1,3,0,1],[2,0,5,2],[1,1,2,6]]."""
    return np.array([4.0, 1.0, 2.0, 1.0, 1.0, 3.0, 0.0, 1.0, 2.0, 0.0, 5.0, 2.0, 1.0, 1.0, 2.0, 6.0])


def singular3() -> np.ndarray:
    """src/lapack/mod.rs:497-502: second column twice the first."""
    return np.array([1.0, 2.0, 3.0, 2.0, 4.0, 6.0, 1.0, 1.0, 1.0])


def sample(n: int) -> np.ndarray:
    """src/parity.rs:57-70: dominant diagonal n+1+i/2, off-diagonal ((7i+13j) % 11) - 5."""
    a = np.zeros(n * n)
    for i in range(n):
        for j in range(n):
            a[i + j * n] = (n + 1.0 + i * 0.5) if i == j else float((i * 7 + j * 13) % 11) - 5.0
    return a


def spd(n: int) -> np.ndarray:
    """src/lapack/checkers.rs:260-279: A = L L^T with integer L."""
    L = np.zeros((n, n))
    for j in range(n):
        L[j, j] = j + 2.0
        for i in range(j + 1, n):
            L[i, j] = float((i + j) % 3) - 1.0
    return cm(L @ L.T)


def hilbert(n: int) -> np.ndarray:
    """src/lapack/checkers.rs:283-291."""
    return cm([[1.0 / (i + j + 1) for j in range(n)] for i in range(n)])


def lower3() -> np.ndarray:
    """src/blas.rs:2750-2753: [[2,0,0],[1,3,0],[4,5,6]]."""
    return np.array([2.0, 1.0, 4.0, 0.0, 3.0, 5.0, 0.0, 0.0, 6.0])


def upper3() -> np.ndarray:
    """src/blas.rs:2755-2758: [[2,1,4],[0,3,5],[0,0,6]]."""
    return np.array([2.0, 0.0, 0.0, 1.0, 3.0, 0.0, 4.0, 5.0, 6.0])


# src/blas.rs:747-794 and :2264-2319: the shapes that reach every branch of the blocked CPU
# kernel (micro-kernel remainders, MC=128 / KC=256 / NC=256 boundaries).
SWEEP_SHAPES = [
    (1, 1, 1), (1, 5, 1), (5, 1, 1), (1, 1, 5), (2, 3, 4),
    (3, 3, 3), (4, 4, 4), (5, 5, 5), (7, 9, 11), (8, 8, 8), (13, 6, 17),
    (128, 8, 8), (129, 8, 8), (8, 128, 8), (8, 129, 8), (8, 8, 256), (8, 8, 257),
    (200, 300, 260), (260, 130, 300),
]
SWEEP_ALPHA_BETA = [(1.0, 1.0), (1.0, 0.0), (2.5, -1.5), (0.0, 3.0)]


def sweep_operands(m: int, n: int, k: int):
    """The deterministic operands of src/blas.rs:775-777: A_op (m x k), B_op (k x n), C0."""
    a = np.array([float((i * 7) % 19) - 9.0 for i in range(m * k)])
    b = np.array([float((i * 11) % 23) - 11.0 for i in range(k * n)])
    c0 = np.array([float(i % 5) * 0.25 for i in range(m * n)])
    return a, b, c0


def transpose_flat(a, nrow: int, ncol: int) -> np.ndarray:
    """Column-major transpose: (nrow x ncol) -> (ncol x nrow)."""
    return cm(mat(a, nrow, ncol).T)


def uniform_pm1(n_entries: int, seed: int) -> np.ndarray:
    """Seeded uniform(-1, 1) entries for size sweeps (numpy PCG64; not the LAPACK stream)."""
    return np.random.default_rng(seed).uniform(-1.0, 1.0, n_entries)


def gemm_error_bound(alpha, a_op, b_op, beta, c0, k: int) -> np.ndarray:
    """Elementwise bound 4*k*eps*(|alpha| (|A_op||B_op|)_ij + |beta| |C0_ij|) + tiny, 2-D."""
    eps = np.finfo(np.float64).eps
    return 4.0 * max(k, 1) * eps * (abs(alpha) * (np.abs(a_op) @ np.abs(b_op)) + abs(beta) * np.abs(c0)) + 1e-300
