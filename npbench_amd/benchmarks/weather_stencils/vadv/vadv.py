"""Data initialisation for implicit vertical advection.

``initialize`` restates npbench/benchmarks/weather_stencils/vadv/vadv.py:6-19:
dtr_stage = 3/20 and default_rng(42) draws utens_stage, u_stage,
wcon (I+1, J, K), u_pos, utens uniform[0,1), in that order.
``initialize_config`` keeps the draw order with wcon = 2*U - 1, i.e.
uniform[-1, 1).
"""

import numpy as np


def initialize(I, J, K):
    rng = np.random.default_rng(42)
    dtr_stage = 3. / 20.
    utens_stage = rng.random((I, J, K))
    u_stage = rng.random((I, J, K))
    wcon = rng.random((I + 1, J, K))
    u_pos = rng.random((I, J, K))
    utens = rng.random((I, J, K))
    return dtr_stage, utens_stage, u_stage, wcon, u_pos, utens


def initialize_config(I, J, K, seed=42):
    rng = np.random.default_rng(seed)
    dtr_stage = 3. / 20.
    utens_stage = rng.random((I, J, K))
    u_stage = rng.random((I, J, K))
    wcon = 2.0 * rng.random((I + 1, J, K)) - 1.0
    u_pos = rng.random((I, J, K))
    utens = rng.random((I, J, K))
    return dtr_stage, utens_stage, u_stage, wcon, u_pos, utens
