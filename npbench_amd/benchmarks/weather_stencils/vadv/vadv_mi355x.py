"""MI355X implementation of vadv_numpy.vadv (same signature).

``vadv`` is the FP64 drop-in (bitwise the NumPy result); the "fp32" variant
stores the fields in FP32 and still eliminates in FP64.
"""

import functools

from npbench_amd import stencils


def vadv(utens_stage, u_stage, wcon, u_pos, utens, dtr_stage):
    stencils.vadv_dropin(utens_stage, u_stage, wcon, u_pos, utens, dtr_stage, precision="fp64")


VARIANTS = {
    "fp32": functools.partial(stencils.vadv_dropin, precision="fp32"),
}
