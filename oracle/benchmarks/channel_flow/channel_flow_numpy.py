from oracle.stencils import channel_flow  # noqa: F401  (channel_flow_numpy.py:69)
