"""MI355X implementation of seidel_2d_numpy.kernel (same signature): the FP64
drop-in, bitwise the NumPy result; A is updated in place."""

from npbench_amd import stencils


def kernel(TSTEPS, N, A):
    stencils.seidel_2d(TSTEPS, N, A)
