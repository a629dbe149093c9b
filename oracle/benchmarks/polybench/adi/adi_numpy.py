from oracle.stencils import adi_kernel as kernel  # noqa: F401  (adi_numpy.py:6)
