from oracle.stencils import jacobi_2d_kernel as kernel  # noqa: F401  (jacobi_2d_numpy.py:4)
