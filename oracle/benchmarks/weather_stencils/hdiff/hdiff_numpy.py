from oracle.stencils import hdiff  # noqa: F401  (hdiff_numpy.py:5)
