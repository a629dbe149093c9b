"""MI355X implementation of jacobi_1d_numpy.kernel (same signature): the FP64
drop-in, bitwise the NumPy result; A and B are updated in place."""

from npbench_amd import stencils


def kernel(TSTEPS, A, B):
    stencils.jacobi_1d(TSTEPS, A, B)
