"""Benchmark specifications and data construction.

Follows npbench/infrastructure/benchmark.py:12-66: the spec is
``bench_info/<name>.json`` ({"benchmark": {...}}); ``get_data(preset)``
puts the preset's parameters in a dict, imports the init function from
``<relative_path>.<module_name>`` and binds its outputs by name; the result
is cached per preset and must not be mutated by callers.

Extension: a spec may carry ``preset_init`` -- a per-preset init function
(same {func_name, input_args, output_args} schema) -- used by the seeded
configuration presets C1..C4, whose distributions differ from the
reference's presets.
"""

import importlib
import json
import pathlib

PKG = pathlib.Path(__file__).resolve().parent.parent
BENCH_INFO = PKG / "bench_info"
BENCH_ROOT = "npbench_amd.benchmarks"


class Benchmark:
    def __init__(self, bname):
        self.bname = bname
        self.bdata = {}
        path = BENCH_INFO / f"{bname}.json"
        try:
            with open(path) as f:
                self.info = json.load(f)["benchmark"]
        except Exception:
            print(f"Benchmark JSON file {path.name} could not be opened.")
            raise

    @property
    def module_path(self):
        rel = self.info["relative_path"].replace("/", ".")
        return f"{rel}.{self.info['module_name']}"

    def init_spec(self, preset):
        return self.info.get("preset_init", {}).get(preset) or self.info.get("init")

    def get_data(self, preset="L"):
        if preset in self.bdata:
            return self.bdata[preset]
        params = self.info["parameters"]
        if preset not in params:
            raise NotImplementedError(f"{self.bname} doesn't have a {preset} preset.")
        data = dict(params[preset])
        spec = self.init_spec(preset)
        if spec:
            module = importlib.import_module(f"{BENCH_ROOT}.{self.module_path}")
            init = getattr(module, spec["func_name"])
            outs = init(*[data[a] for a in spec["input_args"]])
            names = spec["output_args"]
            if len(names) == 1:
                outs = (outs,)
            if len(outs) != len(names):
                raise ValueError(f"{self.bname}: init returned {len(outs)} values for "
                                 f"{len(names)} names")
            data.update(zip(names, outs))
        self.bdata[preset] = data
        return data

    def lups(self, preset):
        """Grid-point updates of one call at ``preset`` (None if unspecified)."""
        expr = self.info.get("lups")
        if not expr:
            return None
        return int(eval(expr, {"__builtins__": {}}, dict(self.info["parameters"][preset])))
