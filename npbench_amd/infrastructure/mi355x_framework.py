"""The MI355X framework adapter (framework_info/mi355x.json).

What a maintainer adds to npbench to run the HIP kernels under the existing
harness, modelled on the GPU adapters there
(npbench/infrastructure/cupy_framework.py:8-58, appy_framework.py:28-52,
numba_framework.py:8-72):

  copy_func        host float64 array -> device tensor (HBM), untimed setup
  copy_back_func   device tensor -> host NumPy array, for validation
  setup_str        the copies followed by a device synchronisation
  exec_str         the call followed by ``__npb_sync()``, a polling wait that
                   a harness timeout can interrupt
  implementations  "default" (the FP64 drop-in) plus every storage variant the
                   per-benchmark module lists in VARIANTS (fp32, bf16, ...),
                   recorded in the ``details`` column like Numba's variants
"""

import importlib

from .framework import Framework


class MI355XFramework(Framework):
    def version(self):
        from npbench_amd import _native
        return _native.version()

    def imports(self):
        from npbench_amd import device
        device.require_device()
        return {"__npb_sync": device.sync}

    def copy_func(self):
        from npbench_amd import device
        device.require_device()
        return device.to_device

    def copy_back_func(self):
        from npbench_amd import device
        return device.to_host

    def implementations(self, bench):
        func = bench.info["func_name"]
        try:
            module = importlib.import_module(self.module_name(bench))
        except Exception:
            print(f"Failed to load the {self.info['full_name']} {func} implementation.")
            raise
        impls = [(getattr(module, func), "default")]
        for name, fn in getattr(module, "VARIANTS", {}).items():
            impls.append((fn, name))
        return impls

    def setup_str(self, bench, impl=None):
        base = super().setup_str(bench, impl)
        return "__npb_sync()" if base == "pass" else base + "; __npb_sync()"

    def exec_str(self, bench, impl=None):
        return super().exec_str(bench, impl) + "; __npb_sync()"
