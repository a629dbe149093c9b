"""Data initialisation for the channel flow micro-app.

``initialize`` restates npbench/benchmarks/channel_flow/channel_flow.py:6-13:
u and v start at zero and p at one on an (ny, nx) grid over [0, 2]^2, with
dt = 0.1 / ((nx - 1)(ny - 1)).
"""

import numpy as np


def initialize(ny, nx):
    u = np.zeros((ny, nx), dtype=np.float64)
    v = np.zeros((ny, nx), dtype=np.float64)
    p = np.ones((ny, nx), dtype=np.float64)
    dx = 2 / (nx - 1)
    dy = 2 / (ny - 1)
    dt = .1 / ((nx - 1) * (ny - 1))
    return u, v, p, dx, dy, dt
