"""The NumPy framework's per-benchmark modules for the harness counterpart.

Test infrastructure: the harness's "numpy" framework (framework_info/numpy.json)
resolves ``<module>_numpy`` here, exactly as the reference resolves
npbench/benchmarks/<path>/<module>_numpy.py, and validates other frameworks
against it.  Each module re-exports the oracle function under the
reference's function name.
"""
