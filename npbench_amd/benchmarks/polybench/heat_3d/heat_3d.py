"""Data initialisation for the 3-D heat benchmark.

``initialize`` restates npbench/benchmarks/polybench/heat_3d/heat_3d.py:6-11
(the S/M/L/paper presets): A[i,j,k] = (i + j + (N - k)) * 10 / N, B = A.copy().
``initialize_uniform`` is the seeded configuration: A uniform[0,1) from
numpy.random.default_rng(seed), B = A.copy().
"""

import numpy as np


def initialize(N, datatype=np.float64):
    i, j, k = np.indices((N, N, N), dtype=datatype)
    A = (i + j + (N - k)) * 10 / N
    B = A.copy()
    return A, B


def initialize_uniform(N, seed=42):
    rng = np.random.default_rng(seed)
    A = rng.random((N, N, N))
    B = A.copy()
    return A, B
