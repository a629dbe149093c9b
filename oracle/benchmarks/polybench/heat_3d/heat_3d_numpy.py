from oracle.stencils import heat_3d_kernel as kernel  # noqa: F401  (heat_3d_numpy.py:4)
