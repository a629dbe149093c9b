"""Data initialisation for horizontal diffusion.

``initialize`` restates npbench/benchmarks/weather_stencils/hdiff/hdiff.py:6-15:
default_rng(42) draws in_field (I+4, J+4, K), out_field and coeff (I, J, K),
uniform[0,1), in that order.  ``initialize_config`` keeps the draw order and
scales coeff to uniform[0, 0.25).
"""

import numpy as np


def initialize(I, J, K):
    rng = np.random.default_rng(42)
    in_field = rng.random((I + 4, J + 4, K))
    out_field = rng.random((I, J, K))
    coeff = rng.random((I, J, K))
    return in_field, out_field, coeff


def initialize_config(I, J, K, seed=42):
    rng = np.random.default_rng(seed)
    in_field = rng.random((I + 4, J + 4, K))
    out_field = rng.random((I, J, K))
    coeff = 0.25 * rng.random((I, J, K))
    return in_field, out_field, coeff
