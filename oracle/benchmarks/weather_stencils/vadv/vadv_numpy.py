from oracle.stencils import vadv  # noqa: F401  (vadv_numpy.py:9)
