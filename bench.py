"""Benchmark of the MI355X stencil path (one GPU).

Default workload: 3-D 7-point heat on a 256^3 grid, TSTEPS=101 (100 ping-pong
iterations = 200 sweeps), seeded uniform[0,1) FP64 inputs with B = A.copy(),
stored on the device in FP32 (the reduced-precision storage form).  One
"step" is one full kernel call (200 sweeps) on inputs already resident in
HBM.  The plain FP64 drop-in is timed on the same workload beside it, and
the FP32 result is compared with the FP64 result (bitwise the NumPy oracle,
as the parity tests show).

Prints ONE JSON line (rank 0): metric/value/unit, roofline of the dominant
kernel (HIP-event timed on the stream the kernels run on) and the CPU
baseline (the NumPy FP64 oracle timed on this host, bounded sample).
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
N = 256
TSTEPS = 101
SWEEPS = 2 * (TSTEPS - 1)


def lups_per_sweep(n):
    return (n - 2) ** 3


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def time_gpu(st, dev, A, B, steps, warmup, tblock=1):
    """Runs warmup + timed steps; returns (wall_s, event_ms)."""
    import torch
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        st.heat3d_sweeps(A, B, SWEEPS, tblock=tblock, stream=stream)
    timer = dev.Timer(stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    timer.start()
    for _ in range(steps):
        st.heat3d_sweeps(A, B, SWEEPS, tblock=tblock, stream=stream)
    timer.stop()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, timer.elapsed_ms()


def cpu_baseline(n, sweeps):
    """NumPy FP64 oracle, median time per sweep on this host (1 core)."""
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
        "value": lups_per_sweep(n) / med / 1e9,
        "unit": "GLUP/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"{sweeps} sweeps of heat_3d {n}^3 FP64 with the NumPy oracle, median "
                   f"sweep {med * 1e3:.1f} ms; host {cpu_model()}, os.cpu_count()="
                   f"{os.cpu_count()}, affinity={len(os.sched_getaffinity(0))}"),
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--tblock", type=int, default=1)
    ap.add_argument("--cpu-sweeps", type=int, default=16,
                    help="sweeps in the CPU baseline sample (0 = skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)

    import torch
    from npbench_amd import device as dev, stencils as st
    dev.require_device(0)

    rng = np.random.default_rng(42)
    A64 = rng.random((N, N, N))
    B64 = A64.copy()
    lup_step = lups_per_sweep(N) * SWEEPS

    # FP32 storage form (headline)
    A = dev.to_device(A64, torch.float32)
    B = dev.to_device(B64, torch.float32)
    wall32, ev32 = time_gpu(st, dev, A, B, args.steps, args.warmup, args.tblock)

    # plain FP64 drop-in on the same workload
    A = dev.to_device(A64)
    B = dev.to_device(B64)
    wall64, ev64 = time_gpu(st, dev, A, B, args.steps, args.warmup, 1)
    del A, B

    # accuracy: one step from the same initial state, FP32 vs FP64 (the FP64
    # GPU path is bitwise the NumPy oracle)
    a32, b32 = dev.to_device(A64, torch.float32), dev.to_device(B64, torch.float32)
    a64, b64 = dev.to_device(A64), dev.to_device(B64)
    st.heat3d_sweeps(a32, b32, SWEEPS, tblock=args.tblock)
    st.heat3d_sweeps(a64, b64, SWEEPS)
    torch.cuda.synchronize()
    err = (a32.double() - a64).abs().max().item() / a64.abs().max().item()
    del a32, b32, a64, b64

    value = lup_step * args.steps / wall32 / 1e9
    value64 = lup_step * args.steps / wall64 / 1e9
    launches = args.steps * SWEEPS
    avg_launch_s = ev32 / 1e3 / launches
    bytes_per_lup = 2 * 4 / max(args.tblock, 1)
    achieved = bytes_per_lup * lups_per_sweep(N) / avg_launch_s / 1e9
    avg64 = ev64 / 1e3 / launches
    achieved64 = 16 * lups_per_sweep(N) / avg64 / 1e9

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "GLUP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall32 / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded uniform[0,1), numpy default_rng(42))",
        "config": {"workload": "heat_3d 256^3, TSTEPS=101 (200 sweeps), FP32 storage",
                   "N": N, "tsteps": TSTEPS, "sweeps_per_step": SWEEPS, "storage": "f32",
                   "tblock": args.tblock},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "kernel": "heat3d_sweep<float>",
                     "avg_launch_us": round(avg_launch_s * 1e6, 3)},
        "fp64_plain": {"value": round(value64, 3), "unit": "GLUP/s",
                       "ms_per_step": round(wall64 / args.steps * 1e3, 4),
                       "roofline_achieved_gbs": round(achieved64, 1),
                       "roofline_frac": round(achieved64 / HBM_PEAK_GBS, 4)},
        "speedup_vs_fp64_plain": round(value / value64, 3),
        "max_rel_err_vs_fp64": err,
    }
    if args.cpu_sweeps > 0:
        line["cpu_baseline"] = cpu_baseline(N, args.cpu_sweeps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
