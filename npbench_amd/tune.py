"""Autotuner for the explicit stencils' launch shape.

For a workload class -- kernel kind, storage dtype, side length and number
of grids advanced together -- it times candidate shapes on scratch copies of
seeded grids (HIP events on the launch stream) and keeps the fastest:

* T = 1: the plain vector sweep, one launch per sweep and grid, with its
  workgroup shape `shape` (planes per workgroup / planes loaded ahead);
* T >= 2: the temporal-blocked tile kernel with rows per wave R (heat) and
  planes per tile `seg`, through a batch plan of all the grids.

Every candidate computes bitwise the same result, so the choice only moves
performance.  Choices are cached in memory and in a JSON file
(``$NPBX_TUNE_CACHE``, default ``~/.cache/npbench_amd/tune.json``) keyed by
the class and the device name.
"""

import json
import os
import pathlib

import torch

from . import _native as nat
from . import batch
from . import device as dev
from . import stencils as st

DEFAULT_CACHE = pathlib.Path(os.environ.get(
    "NPBX_TUNE_CACHE", pathlib.Path.home() / ".cache" / "npbench_amd" / "tune.json"))

# (T, rows) pairs instantiated in libnpbx (rows only matters for heat; FP32
# rows 10 .. 14 select the wave-private heat tile)
SHAPES = {
    torch.float32: [(2, 2), (2, 4), (2, 6), (2, 8), (3, 4), (3, 6), (4, 4), (2, 10), (2, 12),
                    (3, 12), (3, 14)],
    torch.float64: [(2, 2), (2, 4), (2, 6)],
    torch.bfloat16: [(2, 2), (2, 4)],
}
SEGS = (32, 64, 128)
SWEEP_SHAPES = (0, 1, 2, 3)  # libnpbx "sweep.shape": 16/4, 16/8, 32/8, 8/4
VEC = {torch.float64: 2, torch.float32: 4, torch.bfloat16: 8}

_memo = {}


def class_key(kind, dtype, n, count):
    return f"{kind}:{str(dtype).replace('torch.', '')}:{n}:{count}"


def _device_name():
    return torch.cuda.get_device_name(0) if torch.cuda.is_available() else "none"


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save(path, table):
    path = pathlib.Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    tmp = path.with_suffix(".tmp")
    with open(tmp, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def apply(cfg):
    """Sets the library knobs of a tuned shape (a dict from ``best``)."""
    if cfg.get("T", 1) >= 2:
        nat.call("npbx_set_tuning", b"heat.rows", int(cfg["rows"]))
        nat.call("npbx_set_tuning", b"heat.seg", int(cfg["seg"]))
    else:
        nat.call("npbx_set_tuning", b"sweep.shape", int(cfg.get("shape", 0)))


def candidates(kind, dtype):
    shapes = SWEEP_SHAPES if kind == "heat_3d" else (0,)  # Jacobi's sweep has one shape
    out = [{"T": 1, "rows": 0, "seg": 0, "shape": sh} for sh in shapes]
    for T, R in SHAPES[dtype]:
        if kind == "jacobi_2d":
            if R == 4:  # rows / seg do not apply to Jacobi tiles
                out.append({"T": T, "rows": 4, "seg": 32})
        else:
            out += [{"T": T, "rows": R, "seg": s} for s in SEGS]
    return out


def _time(fn, reps):
    fn()  # warm-up (and any first-call work)
    torch.cuda.synchronize()
    t = dev.Timer()
    t.start()
    for _ in range(reps):
        fn()
    t.stop()
    return t.elapsed_ms() / reps / 1e3


def measure(kind, dtype, n, count, cfg, sweeps=16, reps=2, seed=0):
    """GLUP/s of ``cfg`` on ``count`` seeded grids of side ``n``."""
    dev.require_device()
    ndim = 2 if kind == "jacobi_2d" else 3
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    base = torch.rand((n,) * ndim, generator=g, dtype=torch.float64, device="cuda").to(dtype)
    grids = [(base.clone(), base.clone()) for _ in range(count)]
    lups = count * sweeps * (n - 2) ** ndim
    if cfg["T"] == 1:
        apply(cfg)
        sweep = st.jacobi2d_sweeps if kind == "jacobi_2d" else st.heat3d_sweeps

        def fn():
            for a, b in grids:
                sweep(a, b, sweeps)
        secs = _time(fn, reps)
    else:
        apply(cfg)
        plan = batch.Plan([batch.Grid(kind, a, b, sweeps) for a, b in grids], tblock=cfg["T"],
                          seg=cfg["seg"])
        try:
            secs = _time(plan.launch, reps)
        finally:
            plan.close()
    return lups / secs / 1e9


def best(kind, dtype, n, count=1, cache=DEFAULT_CACHE, retune=False, sweeps=16, verbose=False):
    """The fastest shape for the class, tuned on first use and cached."""
    key = class_key(kind, dtype, n, count)
    dkey = _device_name()
    if not retune and key in _memo:
        return _memo[key]
    table = _load(cache) if cache else {}
    if not retune and key in table.get(dkey, {}):
        _memo[key] = table[dkey][key]
        return _memo[key]
    results = []
    usable = n % VEC[dtype] == 0
    for cfg in candidates(kind, dtype):
        if cfg["T"] >= 2 and not usable:
            continue
        try:
            rate = measure(kind, dtype, n, count, cfg, sweeps=sweeps)
        except nat.NpbxError as e:  # a shape the library rejects for this size
            if verbose:
                print(f"tune {key} {cfg}: {e}")
            continue
        results.append((rate, cfg))
        if verbose:
            print(f"tune {key} {cfg}: {rate:.1f} GLUP/s", flush=True)
    rate, cfg = max(results, key=lambda x: x[0])
    out = {**cfg, "glups": round(rate, 2)}
    _memo[key] = out
    if cache:
        table = _load(cache)
        table.setdefault(dkey, {})[key] = out
        _save(cache, table)
    return out
