from oracle.stencils import seidel_2d_kernel as kernel  # noqa: F401  (seidel_2d_numpy.py:4)
