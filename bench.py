"""Benchmark of the MI355X stencil path (one GPU).

Headline workload: 3-D 7-point heat on a 256^3 grid, TSTEPS=101 (100
ping-pong iterations = 200 sweeps), seeded uniform[0,1) FP64 inputs with
B = A.copy(), stored on the device in FP32 (the reduced-precision storage
form).  One "step" is one full kernel call (200 sweeps) on inputs already
resident in HBM.  The plain FP64 drop-in (one launch per sweep) is timed on
the same workload beside it, and the FP32 result is compared with the FP64
result (bitwise the NumPy oracle, as the parity tests show).

Also reported (same JSON line, "kernels"): Jacobi 1000^2 FP64 (TSTEPS=101),
hdiff 256x256x160 (FP64/FP32/BF16), vadv 256x256x160 (FP64, FP32 fields), and
a measured device copy rate.  Device times come from HIP events recorded on
the stream the kernels are launched on.

Prints ONE JSON line: metric/value/unit, roofline of the dominant kernel and
the CPU baseline (the NumPy FP64 oracle timed on this host, bounded sample).
"""

import argparse
import json
import os
import platform
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("GLUP/s + % HBM roofline, batched FP32 vs plain FP64 GPU; "
          "p50/p99 serve latency")
HBM_PEAK_GBS = 8000.0
HEAT_N = 256
TSTEPS = 101
SWEEPS = 2 * (TSTEPS - 1)
SIZE = {"f64": 8, "f32": 4, "bf16": 2}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


class Clock:
    """Wall clock plus HIP events on the launch stream around a region."""

    def __init__(self, dev, stream):
        self.dev = dev
        self.stream = stream
        self.timer = dev.Timer(stream)

    def run(self, fn, reps):
        import torch
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        self.timer.start()
        for _ in range(reps):
            fn()
        self.timer.stop()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, self.timer.elapsed_ms() / 1e3


def roofline(bytes_per_launch, launch_s, kernel):
    achieved = bytes_per_launch / launch_s / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": kernel, "avg_launch_us": round(launch_s * 1e6, 3)}


# ----------------------------------------------------------------- headline
def heat_headline(st, dev, torch, clock, args):
    rng = np.random.default_rng(42)
    A64 = rng.random((HEAT_N,) * 3)
    B64 = A64.copy()
    lup_sweep = (HEAT_N - 2) ** 3
    lup_step = lup_sweep * SWEEPS
    out = {}
    for tag, dtype, tblock in (("f32", torch.float32, args.tblock), ("f64", torch.float64, 1)):
        A = dev.to_device(A64, dtype)
        B = dev.to_device(B64, dtype)

        def step():
            st.heat3d_sweeps(A, B, SWEEPS, tblock=tblock, stream=clock.stream)

        for _ in range(args.warmup):
            step()
        wall, ev = clock.run(step, args.steps)
        launches = args.steps * SWEEPS if tblock <= 1 else None
        out[tag] = {"wall": wall, "event": ev, "launches": launches,
                    "glups": lup_step * args.steps / wall / 1e9,
                    "glups_device": lup_step * args.steps / ev / 1e9}
        del A, B
    # accuracy: one step from the same initial state
    a32, b32 = dev.to_device(A64, torch.float32), dev.to_device(B64, torch.float32)
    a64, b64 = dev.to_device(A64), dev.to_device(B64)
    st.heat3d_sweeps(a32, b32, SWEEPS, tblock=args.tblock)
    st.heat3d_sweeps(a64, b64, SWEEPS)
    torch.cuda.synchronize()
    err = (a32.double() - a64).abs().max().item() / a64.abs().max().item()
    return out, err, lup_sweep, lup_step


def cpu_baseline(n, sweeps):
    """NumPy FP64 oracle, median time per heat sweep on this host (1 core)."""
    from oracle import stencils as ref
    rng = np.random.default_rng(42)
    A = rng.random((n, n, n))
    B = A.copy()
    times = []
    for s in range(sweeps):
        t0 = time.perf_counter()
        if s % 2 == 0:
            ref.heat_sweeps(A, B, 1)
        else:
            ref.heat_sweeps(B, A, 1)
        times.append(time.perf_counter() - t0)
    med = statistics.median(times)
    return {
        "value": round((n - 2) ** 3 / med / 1e9, 5),
        "unit": "GLUP/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"{sweeps} sweeps of heat_3d {n}^3 FP64 with the NumPy oracle (single-"
                   f"threaded ufuncs), median sweep {med * 1e3:.1f} ms; host {cpu_model()}, "
                   f"os.cpu_count()={os.cpu_count()}, affinity={len(os.sched_getaffinity(0))}"),
    }


# ----------------------------------------------------------------- kernels
def kernel_suite(st, dev, torch, clock, reps):
    res = {}
    rng = np.random.default_rng(42)

    # copy rate (2 GiB) for reference against the datasheet peak
    nbytes = 1 << 31
    src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    from npbench_amd import _native as nat
    cp = lambda: nat.call("npbx_stream_copy", dst.data_ptr(), src.data_ptr(), nbytes,  # noqa
                          dev.stream_handle(clock.stream))
    cp()
    _, ev = clock.run(cp, reps)
    res["stream_copy_2GiB"] = {"GBps": round(2 * nbytes / (ev / reps) / 1e9, 1)}
    del src, dst

    # Jacobi 1000^2 FP64, TSTEPS=101 (plain sweeps)
    n = 1000
    A = dev.to_device(rng.random((n, n)))
    B = dev.to_device(rng.random((n, n)))
    f = lambda: st.jacobi2d_sweeps(A, B, SWEEPS, stream=clock.stream)  # noqa
    f()
    wall, ev = clock.run(f, reps)
    lups = (n - 2) ** 2 * SWEEPS
    res["jacobi2d_1000_f64"] = {"glups": round(lups * reps / wall / 1e9, 2),
                                "avg_sweep_us": round(ev / reps / SWEEPS * 1e6, 3),
                                "hbm_frac": round(16 * lups * reps / ev / 1e9 / HBM_PEAK_GBS, 4)}
    del A, B

    # hdiff 256x256x160
    I, J, K = 256, 256, 160
    inf = rng.random((I + 4, J + 4, K))
    coeff = 0.25 * rng.random((I, J, K))
    for tag, dt in (("f64", torch.float64), ("f32", torch.float32), ("bf16", torch.bfloat16)):
        di, dc = dev.to_device(inf, dt), dev.to_device(coeff, dt)
        do = torch.empty((I, J, K), dtype=dt, device="cuda")
        f = lambda: st.hdiff(di, do, dc, stream=clock.stream)  # noqa
        f()
        wall, ev = clock.run(f, reps)
        byts = SIZE[tag] * ((I + 4) * (J + 4) * K + 2 * I * J * K)
        res[f"hdiff_256x256x160_{tag}"] = {
            "glups": round(I * J * K * reps / ev / 1e9, 2),
            "avg_us": round(ev / reps * 1e6, 3),
            "hbm_frac": round(byts * reps / ev / 1e9 / HBM_PEAK_GBS, 4)}
        del di, dc, do

    # vadv 256x256x160
    fields = [rng.random((I, J, K)), rng.random((I, J, K)), 2 * rng.random((I + 1, J, K)) - 1,
              rng.random((I, J, K)), rng.random((I, J, K))]
    for tag, dt in (("f64", torch.float64), ("f32", torch.float32)):
        d = [dev.to_device(a, dt) for a in fields]
        f = lambda: st.vadv(*d, 0.15, stream=clock.stream)  # noqa
        f()
        wall, ev = clock.run(f, reps)
        byts = SIZE[tag] * (5 * I * J * K + J * K + I * J * K)
        res[f"vadv_256x256x160_{tag}"] = {
            "glups": round(I * J * K * reps / ev / 1e9, 2),
            "avg_us": round(ev / reps * 1e6, 3),
            "hbm_frac": round(byts * reps / ev / 1e9 / HBM_PEAK_GBS, 4)}
        del d
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--tblock", type=int, default=1)
    ap.add_argument("--cpu-sweeps", type=int, default=64,
                    help="sweeps in the CPU baseline sample (0 = skip)")
    ap.add_argument("--no-kernels", action="store_true", help="skip the per-kernel lines")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)

    import torch
    from npbench_amd import device as dev, stencils as st
    dev.require_device(0)
    clock = Clock(dev, torch.cuda.current_stream())

    heat, err, lup_sweep, lup_step = heat_headline(st, dev, torch, clock, args)
    h32, h64 = heat["f32"], heat["f64"]
    bpl32 = 8.0 / max(args.tblock, 1)
    line = {
        "metric": METRIC,
        "value": round(h32["glups"], 3),
        "unit": "GLUP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(h32["wall"] / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded uniform[0,1), numpy default_rng(42))",
        "config": {"workload": "heat_3d 256^3, TSTEPS=101 (200 sweeps), FP32 storage",
                   "N": HEAT_N, "tsteps": TSTEPS, "sweeps_per_step": SWEEPS,
                   "storage": "f32", "tblock": args.tblock},
        "roofline": roofline(bpl32 * lup_sweep, h32["event"] / (args.steps * SWEEPS),
                             "heat3d_sweep_vec<float>"),
        "fp64_plain": {
            "value": round(h64["glups"], 3), "unit": "GLUP/s",
            "ms_per_step": round(h64["wall"] / args.steps * 1e3, 4),
            "roofline": roofline(16.0 * lup_sweep, h64["event"] / (args.steps * SWEEPS),
                                 "heat3d_sweep_vec<double>")},
        "speedup_vs_fp64_plain": round(h32["glups"] / h64["glups"], 3),
        "max_rel_err_vs_fp64": err,
    }
    if not args.no_kernels:
        line["kernels"] = kernel_suite(st, dev, torch, clock, max(3, args.steps // 2))
    if args.cpu_sweeps > 0:
        line["cpu_baseline"] = cpu_baseline(HEAT_N, args.cpu_sweeps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
