"""MI355X implementation of adi_numpy.kernel (same signature): the FP64
drop-in, bitwise the NumPy result; u is updated in place and returned."""

from npbench_amd import stencils


def kernel(TSTEPS, N, u):
    return stencils.adi_dropin(TSTEPS, N, u)
