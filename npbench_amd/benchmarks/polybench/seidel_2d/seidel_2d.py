"""Data initialisation for the 2-D Gauss-Seidel benchmark.

``initialize`` restates npbench/benchmarks/polybench/seidel_2d/seidel_2d.py:6-10
(the S/M/L/paper presets): A[i, j] = (i (j + 2) + 2) / N.
"""

import numpy as np


def initialize(N, datatype=np.float64):
    i, j = np.indices((N, N), dtype=datatype)
    return (i * (j + 2) + 2) / N
