"""CPU FP64 oracle for the structured-grid stencil path (test infrastructure).

This package is the correctness checker, not the product.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it; the MI355X path in ``npbench_amd`` never calls into it and fails
loudly when its HIP library is missing.

Contents
--------
``stencils``  vectorised NumPy FP64 restatement of the four reference kernels,
              with the reference's element-wise operation order kept, so that
              results are bitwise those of the reference NumPy code on the
              same inputs (IEEE-754 ``+ - * /`` are correctly rounded, so the
              NumPy version does not matter once inputs are identical).
``scalar``    an independent pure-Python scalar-loop restatement for tiny
              shapes (brute force), used to pin ``stencils`` bit for bit.
``numpy_impl`` the reference-named NumPy framework modules used by the
              harness counterpart for validation (``kernel``/``hdiff``/``vadv``
              with the reference signatures).

Pinning (see ``tests/test_oracle.py``): every function here is checked against
closed forms or invariants that the mathematics fixes (affine fixed points,
discrete Fourier eigenmodes, the quadratic cone of heat, hdiff's polynomial
inputs, vadv's exact-solution construction and a banded-solver cross-check)
and against the scalar brute force.  The reference holds no golden vectors
for this path.  ``hdiff_iterated`` is a build-defined extension (the reference
hdiff has no time loop); its ``nsteps=1`` case is the reference hdiff.
"""

from .stencils import (  # noqa: F401
    jacobi_2d_kernel,
    heat_3d_kernel,
    hdiff,
    hdiff_iterated,
    vadv,
    jacobi_sweeps,
    heat_sweeps,
    adi_kernel,
    jacobi_1d_kernel,
    fdtd_2d_kernel,
)
