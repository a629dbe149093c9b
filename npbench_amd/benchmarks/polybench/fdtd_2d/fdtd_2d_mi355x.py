"""MI355X implementation of fdtd_2d_numpy.kernel (same signature): the FP64
drop-in, bitwise the NumPy result; ex, ey and hz are updated in place."""

from npbench_amd import stencils


def kernel(TMAX, ex, ey, hz, _fict_):
    stencils.fdtd_2d(TMAX, ex, ey, hz, _fict_)
