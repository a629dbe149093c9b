"""ctypes binding of the oracle's C restatements (oracle/c/, test
infrastructure only; see ``oracle/__init__.py``).

``seidel_2d_kernel(TSTEPS, N, A)`` is the C statement of
seidel_2d_numpy.kernel (npbench/benchmarks/polybench/seidel_2d/
seidel_2d_numpy.py:4-13) for grids where the NumPy oracle's per-element
Python loop is too slow; it is pinned bit for bit against
``oracle.stencils.seidel_2d_kernel``.  ``build()`` runs oracle/c/Makefile.
"""

import ctypes
import pathlib
import subprocess

import numpy as np

DIR = pathlib.Path(__file__).resolve().parent / "c"
LIB = DIR / "liboracle_c.so"
_lib = None


def build():
    subprocess.run(["make", "-s", "-B", "-C", str(DIR)], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        l = ctypes.CDLL(str(LIB))
        l.oracle_seidel_2d.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p]
        l.oracle_seidel_2d.restype = ctypes.c_int
        _lib = l
    return _lib


def seidel_2d_kernel(TSTEPS, N, A):
    if not (isinstance(A, np.ndarray) and A.dtype == np.float64 and A.flags.c_contiguous
            and A.shape == (N, N)):
        raise TypeError("seidel_2d_kernel needs a C-contiguous float64 (N, N) array")
    if lib().oracle_seidel_2d(int(TSTEPS), int(N), A.ctypes.data) != 0:
        raise MemoryError("oracle_seidel_2d: allocation failed")
