"""Data initialisation for the 2-D FDTD benchmark.

``initialize`` restates npbench/benchmarks/polybench/fdtd_2d/fdtd_2d.py:6-16
(the S/M/L/paper presets): ex = i (j + 1) / NX, ey = i (j + 2) / NY,
hz = i (j + 3) / NX, fict[t] = t.
"""

import numpy as np


def initialize(TMAX, NX, NY, datatype=np.float64):
    i, j = np.indices((NX, NY), dtype=datatype)
    ex = (i * (j + 1)) / NX
    ey = (i * (j + 2)) / NY
    hz = (i * (j + 3)) / NX
    fict = np.arange(TMAX, dtype=datatype)
    return ex, ey, hz, fict
