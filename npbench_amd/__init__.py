"""MI355X-native structured-grid stencil kernels for npbench.

Hot path: 2-D 5-point Jacobi, 3-D 7-point heat, horizontal diffusion with
flux limiter and implicit vertical advection, as hand-written HIP kernels for
gfx950 in libnpbx.so (C ABI in include/npbx.h), driven from Python through
ctypes.  ``npbench_amd.infrastructure`` mirrors npbench's harness and holds
the "mi355x" framework adapter; ``npbench_amd.benchmarks`` holds the
per-kernel drop-in modules with the reference signatures.
"""

__version__ = "0.1.0"
