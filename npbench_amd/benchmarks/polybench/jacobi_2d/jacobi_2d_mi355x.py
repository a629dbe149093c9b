"""MI355X implementation of jacobi_2d_numpy.kernel (same signature).

``kernel`` is the FP64 drop-in (bitwise the NumPy result).  VARIANTS lists
the storage forms the framework adapter exposes as extra implementations:
FP32 and BF16 storage with FP32 arithmetic, and temporal blocking.
"""

import functools

from npbench_amd import stencils


def kernel(TSTEPS, A, B):
    stencils.jacobi_2d(TSTEPS, A, B, precision="fp64")


VARIANTS = {
    "fp32": functools.partial(stencils.jacobi_2d, precision="fp32"),
    "bf16": functools.partial(stencils.jacobi_2d, precision="bf16"),
}
