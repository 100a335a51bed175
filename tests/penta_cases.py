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
