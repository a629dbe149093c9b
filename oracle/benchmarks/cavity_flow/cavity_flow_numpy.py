from oracle.stencils import cavity_flow  # noqa: F401  (cavity_flow_numpy.py:41)
