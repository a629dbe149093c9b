"""MI355X implementation of cavity_flow_numpy.cavity_flow (same signature):
the FP64 drop-in, bitwise the NumPy result; u, v and p are updated in place."""

from npbench_amd import stencils


def cavity_flow(nx, ny, nt, nit, u, v, dt, dx, dy, p, rho, nu):
    stencils.cavity_flow_dropin(nx, ny, nt, nit, u, v, dt, dx, dy, p, rho, nu)
