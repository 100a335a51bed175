"""Shared inputs for the line-solver tests: seeded random diagonally dominant
block-pentadiagonal lines, and LHS blocks assembled from the benchmark
baseflow by the oracle."""
import numpy as np


def random_lines(nsys, nrows, layout, seed, dom=3.0, scale=0.1):
    rng = np.random.default_rng(seed)
    shp = (5, 5, 5, nsys, nrows) if layout == 1 else (5, 5, 5, nrows, nsys)
    mat = (rng.standard_normal(shp) + 1j * rng.standard_normal(shp)) * scale
    for r in range(5):
        mat[2, r, r] += dom
    rshp = (5, nsys, nrows) if layout == 1 else (5, nrows, nsys)
    rhs = rng.standard_normal(rshp) + 1j * rng.standard_normal(rshp)
    return np.asfortranarray(mat), np.asfortranarray(rhs)


def relerr(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _bc_col(n, i, k):
    """column of storage slot k (0..4) in row i of a non-periodic line: the
    band i-2..i+2, with rows 0, 1, n-2, n-1 wrapping their slots onto
    columns 0..4 / n-5..n-1 (the storage contract of src/cpenta1bc.f)"""
    if i == 0:
        return [3, 4, 0, 1, 2][k]
    if i == 1:
        return [4, 0, 1, 2, 3][k]
    if i == n - 2:
        return [n - 4, n - 3, n - 2, n - 1, n - 5][k]
    if i == n - 1:
        return [n - 3, n - 2, n - 1, n - 5, n - 4][k]
    return i - 2 + k


def dense_bc(ml):
    """dense (5n, 5n) matrix of one line, blocks ml (5 slots, 5, 5, n)"""
    n = ml.shape[3]
    A = np.zeros((5 * n, 5 * n), complex)
    for i in range(n):
        for k in range(5):
            c = _bc_col(n, i, k)
            A[5 * i:5 * i + 5, 5 * c:5 * c + 5] += ml[k, :, :, i]
    return A


def dense_periodic(ml):
    """dense cyclic (5n, 5n) matrix, slot k of row i at column (i-2+k) mod n"""
    n = ml.shape[3]
    A = np.zeros((5 * n, 5 * n), complex)
    for i in range(n):
        for k in range(5):
            c = (i - 2 + k) % n
            A[5 * i:5 * i + 5, 5 * c:5 * c + 5] += ml[k, :, :, i]
    return A
