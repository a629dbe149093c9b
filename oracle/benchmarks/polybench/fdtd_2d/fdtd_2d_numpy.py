from oracle.stencils import fdtd_2d_kernel as kernel  # noqa: F401  (fdtd_2d_numpy.py:4)
