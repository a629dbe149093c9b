"""MI355X implementation of channel_flow_numpy.channel_flow (same signature):
the FP64 drop-in; u, v and p are updated in place and the step count is
returned."""

from npbench_amd import stencils


def channel_flow(nit, u, v, dt, dx, dy, p, rho, nu, F):
    return stencils.channel_flow_dropin(nit, u, v, dt, dx, dy, p, rho, nu, F)
