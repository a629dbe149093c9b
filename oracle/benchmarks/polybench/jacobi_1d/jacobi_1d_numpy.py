from oracle.stencils import jacobi_1d_kernel as kernel  # noqa: F401  (jacobi_1d_numpy.py:4)
