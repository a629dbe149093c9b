"""Data initialisation for the 1-D Jacobi benchmark.

``initialize`` restates npbench/benchmarks/polybench/jacobi_1d/jacobi_1d.py:6-10
(the S/M/L/paper presets): A[i] = (i + 2) / N, B[i] = (i + 3) / N.
"""

import numpy as np


def initialize(N, datatype=np.float64):
    i = np.arange(N, dtype=datatype)
    return (i + 2) / N, (i + 3) / N
