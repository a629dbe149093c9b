"""Framework adapters: the plugin boundary of the harness.

Follows npbench/infrastructure/framework.py:12-189.  A framework is described
by ``framework_info/<name>.json`` ({"framework": {simple_name, full_name,
prefix, postfix, class, arch}}); ``generate_framework`` instantiates the class
named there from ``npbench_amd.infrastructure``.  The strings built here are
what the harness hands to timeit:

  setup_str  "__npb_<prefix>_A, ... = __npb_copy(A), ..."  (untimed copies)
  exec_str   "__npb_result = __npb_impl(<input_args>)"     (timed call)

with array arguments renamed to their per-repetition copies.

Extension: an optional ``module_root`` key names the package holding the
framework's ``<module>_<postfix>`` modules (default ``npbench_amd.benchmarks``).
"""

import importlib
import json
import pathlib

import numpy as np

from .benchmark import BENCH_ROOT

PKG = pathlib.Path(__file__).resolve().parent.parent
FRAMEWORK_INFO = PKG / "framework_info"


def _load_info(fname):
    path = FRAMEWORK_INFO / f"{fname}.json"
    try:
        with open(path) as f:
            return json.load(f)["framework"]
    except Exception:
        print(f"Framework JSON file {path.name} could not be opened.")
        raise


class Framework:
    """Base adapter: host NumPy copies, no synchronisation, one implementation."""

    def __init__(self, fname):
        self.fname = fname
        self.info = _load_info(fname)

    # -- identity ----------------------------------------------------------
    def version(self):
        return np.__version__ if self.fname == "numpy" else "unknown"

    def imports(self):
        return {}

    # -- data movement -----------------------------------------------------
    def copy_func(self):
        return np.copy

    def copy_back_func(self):
        return lambda x: x

    # -- implementations ---------------------------------------------------
    def module_name(self, bench):
        root = self.info.get("module_root", BENCH_ROOT)
        postfix = self.info.get("postfix", self.fname)
        return f"{root}.{bench.module_path}_{postfix}"

    def impl_files(self, bench):
        mod = self.module_name(bench)
        return [(mod.replace(".", "/") + ".py", "default")]

    def implementations(self, bench):
        func = bench.info["func_name"]
        try:
            module = importlib.import_module(self.module_name(bench))
            return [(getattr(module, func), "default")]
        except Exception:
            print(f"Failed to load the {self.info['full_name']} {func} implementation.")
            raise

    # -- argument and statement strings -----------------------------------
    def _copy_name(self, arg):
        return f"__npb_{self.info['prefix']}_{arg}"

    def args(self, bench, impl=None):
        arrays = set(bench.info["array_args"])
        return [self._copy_name(a) if a in arrays else a for a in bench.info["input_args"]]

    def mutable_args(self, bench, impl=None):
        return [self._copy_name(a) for a in bench.info["array_args"]]

    def inout_args(self, bench, impl=None):
        return [self._copy_name(a) for a in bench.info["output_args"]]

    def arg_str(self, bench, impl=None):
        return ", ".join(self.args(bench, impl))

    def out_arg_str(self, bench, impl=None):
        return ", ".join(self.mutable_args(bench, impl))

    def setup_str(self, bench, impl=None):
        arrays = bench.info["array_args"]
        if not arrays:
            return "pass"
        copies = ", ".join(f"__npb_copy({a})" for a in arrays)
        return f"{self.out_arg_str(bench, impl)} = {copies}"

    def exec_str(self, bench, impl=None):
        return f"__npb_result = __npb_impl({self.arg_str(bench, impl)})"


def generate_framework(fname, save_strict=False, load_strict=False):
    """Instantiates the adapter class named in ``framework_info/<fname>.json``."""
    info = _load_info(fname)
    module = importlib.import_module("npbench_amd.infrastructure")
    cls = getattr(module, info["class"])
    return cls(fname)
