"""Data initialisation for the ADI benchmark.

``initialize`` restates npbench/benchmarks/polybench/adi/adi.py:6-9 (the
S/M/L/paper presets): u[i, j] = (i + N - j) / N.
"""

import numpy as np


def initialize(N, datatype=np.float64):
    i, j = np.indices((N, N), dtype=datatype)
    return (i + N - j) / N
