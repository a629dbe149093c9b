"""Data initialisation for the 2-D Jacobi benchmark.

``initialize`` restates npbench/benchmarks/polybench/jacobi_2d/jacobi_2d.py:6-10
(the S/M/L/paper presets): A[i,j] = i*(j+2)/N, B[i,j] = i*(j+3)/N.
``initialize_uniform`` is the seeded configuration: A then B drawn
uniform[0,1) from numpy.random.default_rng(seed).
"""

import numpy as np


def initialize(N, datatype=np.float64):
    i, j = np.indices((N, N), dtype=datatype)
    A = i * (j + 2) / N
    B = i * (j + 3) / N
    return A, B


def initialize_uniform(N, seed=42):
    rng = np.random.default_rng(seed)
    A = rng.random((N, N))
    B = rng.random((N, N))
    return A, B
