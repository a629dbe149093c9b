"""MI355X implementation of heat_3d_numpy.kernel (same signature).

``kernel`` is the FP64 drop-in (bitwise the NumPy result); VARIANTS are the
FP32/BF16 storage forms exposed as extra implementations.
"""

import functools

from npbench_amd import stencils


def kernel(TSTEPS, A, B):
    stencils.heat_3d(TSTEPS, A, B, precision="fp64")


VARIANTS = {
    "fp32": functools.partial(stencils.heat_3d, precision="fp32"),
    "bf16": functools.partial(stencils.heat_3d, precision="bf16"),
}
