"""Benchmark of the MI355X stencil path on one GPU.

Headline ("value"): the batched FP32 form of the 3-D heat configuration --
BATCH independent 256^3 grids, each exactly the 3-D heat configuration
(seeded uniform[0,1) A, B = A.copy(), TSTEPS=101 = 100 ping-pong iterations
= 200 sweeps, fixed boundaries), stored in FP32, advanced together with the
launch shape the autotuner picks for this class (plain sweep or
temporal-blocked tile plan).  One "step" is one full pass: every grid of the
batch through its 200 sweeps, inputs resident in HBM.  The batch's working
set (1 GiB) is far beyond the 256 MB Infinity Cache.

Beside it, on the same 256^3 configuration: the plain FP64 drop-in (one
launch per sweep; the comparison base of the metric), a single FP32 grid,
and the FP32-vs-FP64 error (the FP64 GPU path is bitwise the NumPy oracle).
Then a serving run (Poisson mix at fixed QPS: p50/p99 latency and GLUP/s),
per-kernel lines (Jacobi 1000^2, hdiff and vadv 256x256x160, a device copy),
and the CPU baseline (NumPy FP64 oracle on this host, bounded sample).

Device times come from HIP events recorded on the stream the kernels are
launched on.  Prints ONE JSON line.
"""

import argparse
import hashlib
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
BATCH = 8
SIZE = {"f64": 8, "f32": 4, "bf16": 2}
SERVE_HEAT_ROWS = 10  # serving's heat tiles: the wave-private FP32 tile (10 rows
                      # measured lowest p50/p99 on the serving mix)
SERVE_TBLOCK = 2
WARM_S = 0.1  # untimed calls before each per-kernel timing
SERVE_SEG = 32  # heat planes per tile for serving (set explicitly: the process-wide
                # knob holds whatever the headline's tuner chose)


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
        self.stream = stream
        self.timer = dev.Timer(stream)

    def run(self, fn, reps, warm_s=0.0):
        """Times reps calls of fn; with warm_s > 0, first keeps calling fn
        (untimed) for about warm_s seconds so that the clocks of a GPU that
        sat idle (while the host ran the oracle) are back up."""
        import torch
        if warm_s > 0:
            t_end = time.perf_counter() + warm_s
            while True:
                fn()
                torch.cuda.synchronize()
                if time.perf_counter() >= t_end:
                    break
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        self.timer.start()
        for _ in range(reps):
            fn()
        self.timer.stop()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, self.timer.elapsed_ms() / 1e3


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.json")


def measured_traffic(kernel):
    """HBM bytes per launch of ``kernel`` on this workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (None when there is none)."""
    try:
        with open(TRAFFIC_FILE) as f:
            rec = json.load(f).get(kernel)
    except (OSError, ValueError):
        return None
    return None if rec is None else int(rec["read_bytes"] + rec["write_bytes"])


def roofline(bytes_per_launch, launch_s, kernel):
    achieved = bytes_per_launch / launch_s / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(kernel), "kernel": kernel,
            "avg_launch_us": round(launch_s * 1e6, 3)}


def seeded_grid(n):
    """Seeded uniform[0,1) FP64 A (numpy default_rng(42)); B = A.copy()."""
    A = np.random.default_rng(42).random((n, n, n))
    return A, A.copy()


# ----------------------------------------------------------------- headline
def batched_fp32(st, dev, batch, tune, torch, clock, args, A64):
    if args.shape:  # a fixed launch shape (profiler runs): T, rows, seg
        T, rows, seg = (int(x) for x in args.shape.split(","))
        cfg = {"T": T, "rows": rows, "seg": seg, "shape": 0}
    else:
        cfg = tune.best("heat_3d", torch.float32, HEAT_N, args.batch, cache=None,
                        sweeps=16, verbose=False)
    tune.apply(cfg)
    base = dev.to_device(A64, torch.float32)
    grids = [(base.clone(), base.clone()) for _ in range(args.batch)]
    del base
    lup_sweep = (HEAT_N - 2) ** 3
    lup_step = args.batch * lup_sweep * SWEEPS
    if cfg["T"] == 1:
        def step():
            for a, b in grids:
                st.heat3d_sweeps(a, b, SWEEPS, stream=clock.stream)
        launches = args.batch * SWEEPS
        kernel = "heat3d_sweep_vec<float>"
        plan = None
    else:
        plan = batch.Plan([batch.Grid("heat_3d", a, b, SWEEPS) for a, b in grids],
                          tblock=cfg["T"], seg=cfg["seg"])
        if args.graph:
            side = torch.cuda.Stream()
            plan.capture(side)
            clock = type(clock)(dev, side)
        launches = plan.info()["launches"]
        kind = "wave-private" if cfg["rows"] >= 10 else "shared"
        kernel = f"stencil_tiles<float,T={cfg['T']},R={cfg['rows']}> ({kind} heat tile)"

        def step():
            plan.launch(clock.stream)
    for _ in range(args.warmup):
        step()
    wall, ev = clock.run(step, args.steps)
    # algorithmic bytes: one read and one write of every grid per round trip
    bytes_per_lup = 2 * SIZE["f32"] / cfg["T"]
    glups = lup_step * args.steps / wall / 1e9
    rl = roofline(bytes_per_lup * lup_step / launches, ev / (args.steps * launches), kernel)
    # T-independent view: the bandwidth a plain sweep would need for this rate
    rl["effective_GBps"] = round(glups * 2 * SIZE["f32"], 1)
    rl["bytes_per_lup"] = bytes_per_lup
    out = {"cfg": cfg, "wall": wall, "event": ev, "glups": glups, "roofline": rl,
           "launches_per_step": launches}
    if plan is not None:
        plan.close()
    del grids
    tune.apply({"T": 1, "shape": 0})
    return out, lup_step


def single_grid(st, dev, torch, clock, args, A64, dtype, shape):
    from npbench_amd import _native as nat
    nat.call("npbx_set_tuning", b"sweep.shape", shape)
    A = dev.to_device(A64, dtype)
    B = A.clone()

    def step():
        st.heat3d_sweeps(A, B, SWEEPS, stream=clock.stream)
    for _ in range(args.warmup):
        step()
    wall, ev = clock.run(step, args.steps)
    nat.call("npbx_set_tuning", b"sweep.shape", 0)
    lup_step = (HEAT_N - 2) ** 3 * SWEEPS
    return {"wall": wall, "event": ev, "glups": lup_step * args.steps / wall / 1e9,
            "launch_s": ev / (args.steps * SWEEPS)}


def fp32_error(st, dev, torch, A64, B64):
    """FP32 storage against the FP64 GPU path (bitwise the NumPy oracle) on
    the headline configuration: the inf-norm relative error, the elementwise
    maximum of |x - x64| / max(|x64|, 1e-3 max|x64|), and the reference
    harness's 2-norm ratio (utilities.py:27-28)."""
    a32, b32 = dev.to_device(A64, torch.float32), dev.to_device(B64, torch.float32)
    a64, b64 = dev.to_device(A64), dev.to_device(B64)
    st.heat3d_sweeps(a32, b32, SWEEPS)
    st.heat3d_sweeps(a64, b64, SWEEPS)
    torch.cuda.synchronize()
    d = (a32.double() - a64).abs()
    scale = a64.abs().max()
    return {"inf_norm": d.max().item() / scale.item(),
            "elementwise": (d / torch.clamp(a64.abs(), min=1e-3 * scale)).max().item(),
            "norm2": (torch.linalg.vector_norm(d) / torch.linalg.vector_norm(a64)).item()}


# ----------------------------------------------------------------- serving
def serving(torch, args, heat_rows=SERVE_HEAT_ROWS, cpu_rates=None, tblock=SERVE_TBLOCK):
    from npbench_amd import _native as nat, serve
    nat.set_tuning("heat.rows", heat_rows)
    reqs = serve.poisson_mix(args.serve_requests, args.serve_qps, seed=42)
    srv = serve.Server(dtype=torch.float32, tblock=tblock, hbm_budget=32 << 30, max_active=48,
                       seg=SERVE_SEG)
    # warm-up: kernels' first launches and the allocator's pool for a full mix
    srv.run(serve.poisson_mix(64, 1e6, seed=7), realtime=False)
    srv.peak_active = srv.peak_used = srv.ticks = 0
    t0 = time.perf_counter()
    done = srv.run(reqs, realtime=True)
    wall = time.perf_counter() - t0
    out = serve.summarize(done, wall)
    out.update({"qps": args.serve_qps, "mix": "50% jacobi_2d side log-U[128,4096], 50% heat_3d "
                "side U[32,256], steps U[10,200], FP32, seed 42", "tblock": tblock,
                "heat_rows": heat_rows, "heat_seg": SERVE_SEG, "peak_active": srv.peak_active,
                "ticks": srv.ticks})
    # The rate above is the arrivals' (the offered load): the server keeps
    # up.  Its capacity: the same requests all offered at once, admitted as
    # memory and slots allow.
    lups = sum(r.lups for r in reqs)
    span = max(r.arrival for r in reqs) if reqs else 0.0
    out["offered_glups"] = round(lups / span / 1e9, 2) if span > 0 else None
    t0 = time.perf_counter()
    srv.run(reqs, realtime=False)
    cap_wall = time.perf_counter() - t0
    out["capacity"] = {"glups": round(lups / cap_wall / 1e9, 2), "wall_s": round(cap_wall, 4),
                       "note": "the same requests offered at t = 0 (no arrival waits): "
                               "the server's rate on this mix"}
    srv.close()
    if cpu_rates:
        out["cpu_reference"] = cpu_serving_reference(reqs, cpu_rates)
        out["cpu_reference"]["measured_sample"] = cpu_serving_sample(reqs, cpu_rates, torch)
    return out


def cpu_serving_sample(reqs, rates, torch, budget_s=12.0, max_one_s=2.5):
    """Measured, not modelled: requests of the mix drawn at random (seed 5)
    whose modelled service time is at most max_one_s, each served by the
    NumPy FP64 oracle on one host core from its own seeded inputs, until
    about budget_s of CPU time is spent.  Reports every sampled request's
    measured and modelled service time and their median ratio, which checks
    the model the full-mix figures above rest on."""
    from npbench_amd import serve
    from oracle import stencils as ref
    order = np.random.default_rng(5).permutation(len(reqs))
    spent, rows = 0.0, []
    for i in order:
        r = reqs[int(i)]
        model_s = r.lups / (rates[r.kind] * 1e9)
        if model_s > max_one_s or spent + model_s > budget_s:
            continue
        A, B = serve.make_inputs(r, torch.float64)
        a, b = A.cpu().numpy(), B.cpu().numpy()
        f = ref.jacobi_sweeps if r.kind == "jacobi_2d" else ref.heat_sweeps
        t0 = time.perf_counter()
        f(a, b, r.sweeps)
        t = time.perf_counter() - t0
        spent += t
        rows.append({"rid": r.rid, "kind": r.kind, "n": r.n, "steps": r.steps,
                     "measured_ms": round(t * 1e3, 2), "modelled_ms": round(model_s * 1e3, 2)})
    ratios = [x["measured_ms"] / x["modelled_ms"] for x in rows if x["modelled_ms"] > 0]
    return {"requests": rows, "cores": 1,
            "median_measured_over_modelled": round(float(np.median(ratios)), 3) if ratios else None,
            "note": f"random requests with a modelled service time <= {max_one_s} s, about "
                    f"{budget_s} s of oracle time in all"}


def cpu_serving_reference(reqs, rates):
    """The same arrivals served one at a time by the NumPy oracle on one host
    core: each request's service time is its LUPs at the oracle rate measured
    for its kind (GLUP/s, bounded samples), queued first come first served."""
    free_at = 0.0
    lat, service = [], []
    for r in sorted(reqs, key=lambda q: q.arrival):
        svc = r.lups / (rates[r.kind] * 1e9)
        start = max(free_at, r.arrival)
        free_at = start + svc
        lat.append(free_at - r.arrival)
        service.append(svc)
    lat = np.array(lat) * 1e3
    svc = np.array(service) * 1e3
    return {"service_p50_ms": round(float(np.percentile(svc, 50)), 1),
            "service_p99_ms": round(float(np.percentile(svc, 99)), 1),
            "queued_p50_ms": round(float(np.percentile(lat, 50)), 1),
            "queued_p99_ms": round(float(np.percentile(lat, 99)), 1),
            "rates_glups": {k: round(v, 4) for k, v in rates.items()}, "cores": 1,
            "note": "service time = request LUPs / oracle GLUP/s of its kind (sampled on this "
                    "host); latencies from a single-server FCFS queue at the same arrivals"}


# ----------------------------------------------------------------- configurations
def input_hash(*arrays):
    """sha256 (first 16 hex digits) of the FP64 inputs' bytes: identifies the
    seeded inputs a figure was measured on (numpy's Generator streams are not
    guaranteed across versions)."""
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    return h.hexdigest()[:16]


def error_vs(x, ref):
    """inf-norm relative error max|x - ref| / max|ref|, the elementwise
    maximum of |x - ref| / max(|ref|, 1e-3 max|ref|), and the reference
    harness's 2-norm ratio (utilities.py:27-28)."""
    d = np.abs(x - ref)
    scale = float(np.max(np.abs(ref)))
    return {"inf_norm": float(d.max()) / scale,
            "elementwise": float(np.max(d / np.maximum(np.abs(ref), 1e-3 * scale))),
            "norm2": float(np.linalg.norm(d) / np.linalg.norm(ref))}


def oracle_timed(fn, inputs, repeat, outputs):
    """The NumPy oracle with the harness's timing (utilities.py:135-151):
    timeit.repeat(number=1) with fresh copies of the inputs made in the
    untimed setup, median reported.  Returns (median seconds, the outputs of
    the last repetition)."""
    from npbench_amd.infrastructure import utilities as util
    box = [None]
    ctx = {"__f": fn, "__in": inputs, "__box": box, "np": np}
    _, times = util.benchmark(
        "__f(*__box[0])",
        setup="__box[0] = [x.copy() if isinstance(x, np.ndarray) else x for x in __in]",
        repeat=repeat, context=ctx, verbose=False)
    return statistics.median(times), [box[0][i] for i in outputs]


def quantized(a, torch, dt):
    return torch.from_numpy(a).to(dt).double().numpy()


def config_suite(st, dev, torch, clock, reps, cpu):
    """The measured configurations on their own seeded inputs (numpy
    default_rng(42)): 2-D Jacobi 1000^2 with 200 sweeps, hdiff and vadv
    256 x 256 x 160.  Per storage form: GPU time (HIP events on the launch
    stream), the fraction of 8 TB/s of the algorithmic bytes, and, with
    ``cpu``, the error against the FP64 oracle (FP64: bitwise or not) and the
    oracle's own time on this host with the harness's semantics (1 core)."""
    from npbench_amd import _native as nat
    from oracle import stencils as ref
    res = {}
    dts = {"f64": torch.float64, "f32": torch.float32, "bf16": torch.bfloat16}
    host = cpu_model()

    # --- 2-D Jacobi, 1000 x 1000, TSTEPS=101 (200 sweeps)
    from npbench_amd.benchmarks.polybench.jacobi_2d.jacobi_2d import initialize_uniform
    n = 1000
    A, B = initialize_uniform(n, seed=42)
    lups = (n - 2) ** 2 * SWEEPS
    want = None
    if cpu:
        t, want = oracle_timed(ref.jacobi_2d_kernel, [TSTEPS, A, B], 3, (1, 2))
        res["jacobi_2d_1000_cpu_oracle"] = {
            "ms_per_call": round(t * 1e3, 1), "glups": round(lups / t / 1e9, 4), "repeats": 3,
            "cores": 1, "host": host, "timing": "median of timeit.repeat(number=1), copies "
                                                 "untimed (the harness's)"}
    for tag in ("f64", "f32", "bf16"):
        a, b = dev.to_device(A, dts[tag]), dev.to_device(B, dts[tag])

        def jac():
            st.jacobi2d_sweeps(a, b, SWEEPS, stream=clock.stream)
        jac()  # from the seeded inputs: the result checked below
        line = {}
        if want is not None:
            torch.cuda.synchronize()
            got = [a.double().cpu().numpy(), b.double().cpu().numpy()]
            if tag == "f64":
                line["bitwise_vs_oracle"] = bool(all(np.array_equal(g, w)
                                                     for g, w in zip(got, want)))
            else:
                e = [error_vs(g, w) for g, w in zip(got, want)]
                line["max_rel_err_vs_fp64_oracle"] = max(x["inf_norm"] for x in e)
                line["error_vs_fp64_oracle"] = {k: max(x[k] for x in e) for k in e[0]}
        wall, ev = clock.run(jac, reps, WARM_S)
        # the kernel stages 8 sweeps per launch in LDS, so a sweep's own
        # read + write (2 x element x points) over 8 TB/s is a
        # sweep-equivalent rate, not HBM traffic
        line.update({"glups": round(lups * reps / ev / 1e9, 2),
                     "avg_sweep_us": round(ev / reps / SWEEPS * 1e6, 3),
                     "sweeps_per_launch": int(nat.get_tuning("jacobi.deep")),
                     "sweep_equiv_bw_frac": round(2 * SIZE[tag] * lups * reps / ev / 1e9
                                                  / HBM_PEAK_GBS, 4),
                     "input_sha256_16": input_hash(A, B)})
        if cpu:
            line["speedup_vs_cpu_oracle"] = round(
                res["jacobi_2d_1000_cpu_oracle"]["ms_per_call"] / (ev / reps * 1e3), 1)
        res[f"jacobi2d_1000_{tag}"] = line
        del a, b

    # --- hdiff, 256 x 256 x 160, coeff = 0.25 U
    from npbench_amd.benchmarks.weather_stencils.hdiff.hdiff import initialize_config as hd_init
    I, J, K = 256, 256, 160
    inf, out0, coeff = hd_init(I, J, K, seed=42)
    want = None
    if cpu:
        t, (want,) = oracle_timed(ref.hdiff, [inf, out0, coeff], 5, (1,))
        res["hdiff_256x256x160_cpu_oracle"] = {
            "ms_per_call": round(t * 1e3, 1), "glups": round(I * J * K / t / 1e9, 4),
            "repeats": 5, "cores": 1, "host": host}
    for tag in ("f64", "f32", "bf16"):
        dt = dts[tag]
        di, dc = dev.to_device(inf, dt), dev.to_device(coeff, dt)
        do = torch.empty((I, J, K), dtype=dt, device="cuda")

        def hd():
            st.hdiff(di, do, dc, stream=clock.stream)
        hd()
        line = {}
        if want is not None:
            torch.cuda.synchronize()
            got = do.double().cpu().numpy()
            if tag == "f64":
                line["bitwise_vs_oracle"] = bool(np.array_equal(got, want))
            else:
                e = error_vs(got, want)
                line["max_rel_err_vs_fp64_oracle"] = e["inf_norm"]
                line["error_vs_fp64_oracle"] = e
                line["share_beyond_1e-5"] = float(
                    np.mean(np.abs(got - want) > 1e-5 * np.max(np.abs(want))))
                rq = out0.copy()
                ref.hdiff(quantized(inf, torch, dt), rq, quantized(coeff, torch, dt))
                line["max_rel_err_vs_same_input_oracle"] = error_vs(got, rq)["inf_norm"]
        _, ev = clock.run(hd, reps, WARM_S)
        byts = SIZE[tag] * ((I + 4) * (J + 4) * K + 2 * I * J * K)
        line.update({"glups": round(I * J * K * reps / ev / 1e9, 2),
                     "avg_us": round(ev / reps * 1e6, 3),
                     "hbm_frac": round(byts * reps / ev / 1e9 / HBM_PEAK_GBS, 4),
                     "input_sha256_16": input_hash(inf, coeff)})
        if cpu:
            line["speedup_vs_cpu_oracle"] = round(
                res["hdiff_256x256x160_cpu_oracle"]["ms_per_call"] / (ev / reps * 1e3), 1)
        res[f"hdiff_256x256x160_{tag}"] = line
        if tag in ("f32", "bf16"):
            # iterated form: one launch per step (tblock 1) or tblock steps
            # fused per HBM round trip; effective_frac counts one step's
            # bytes per step, so it exceeds the roofline's meaning once steps
            # are fused
            tbs = (1, 2, 4) if tag == "f32" else (1, 2)
            for steps in (4, 8):
                for tb in tbs:
                    nat.set_tuning("hdiff.tblock", tb)

                    def hdn(steps=steps):
                        st.hdiff(di, do, dc, nsteps=steps, stream=clock.stream)
                    hdn()
                    _, ev = clock.run(hdn, reps, WARM_S)
                    res[f"hdiff_256x256x160_{tag}_steps{steps}_tblock{tb}"] = {
                        "glups": round(I * J * K * steps * reps / ev / 1e9, 2),
                        "avg_step_us": round(ev / reps / steps * 1e6, 3),
                        "effective_frac": round(byts * steps * reps / ev / 1e9 / HBM_PEAK_GBS, 4)}
            nat.set_tuning("hdiff.tblock", 1)
        del di, dc, do

    # --- vadv, 256 x 256 x 160, wcon = 2U - 1, dtr_stage = 0.15
    from npbench_amd.benchmarks.weather_stencils.vadv.vadv import initialize_config as va_init
    dtr, *fields = va_init(I, J, K, seed=42)
    want = None
    if cpu:
        t, (want,) = oracle_timed(ref.vadv, [*fields, dtr], 3, (0,))
        res["vadv_256x256x160_cpu_oracle"] = {
            "ms_per_call": round(t * 1e3, 1), "glups": round(I * J * K / t / 1e9, 4),
            "repeats": 3, "cores": 1, "host": host}
    for tag in ("f64", "f32"):
        d0 = [dev.to_device(a, dts[tag]) for a in fields]
        d = [x.clone() for x in d0]

        def va():
            st.vadv(*d, dtr, stream=clock.stream)
        va()
        line = {}
        if want is not None:
            torch.cuda.synchronize()
            got = d[0].double().cpu().numpy()
            if tag == "f64":
                line["bitwise_vs_oracle"] = bool(np.array_equal(got, want))
            else:
                e = error_vs(got, want)
                line["max_rel_err_vs_fp64_oracle"] = e["inf_norm"]
                line["error_vs_fp64_oracle"] = e
                q = [quantized(a, torch, dts[tag]) for a in fields]
                ref.vadv(q[0], *q[1:], dtr)
                line["max_rel_err_vs_same_input_oracle"] = error_vs(got, q[0])["inf_norm"]
                line["note"] = ("wcon in [-1, 1) has near-zero pivots that amplify the FP32 "
                                "rounding of the inputs; the oracle fed the same FP32 inputs "
                                "differs from the FP64 oracle by the same amount")
        _, ev = clock.run(va, reps, WARM_S)
        byts = SIZE[tag] * (5 * I * J * K + J * K + I * J * K)
        line.update({"glups": round(I * J * K * reps / ev / 1e9, 2),
                     "avg_us": round(ev / reps * 1e6, 3),
                     "hbm_frac": round(byts * reps / ev / 1e9 / HBM_PEAK_GBS, 4),
                     "input_sha256_16": input_hash(*fields)})
        if cpu:
            line["speedup_vs_cpu_oracle"] = round(
                res["vadv_256x256x160_cpu_oracle"]["ms_per_call"] / (ev / reps * 1e3), 1)
        res[f"vadv_256x256x160_{tag}"] = line
        del d0, d
    return res


# ----------------------------------------------------------------- kernels
BW_VARIANTS = {0: "U4", 1: "U8", 2: "U4 nontemporal", 3: "U8 nontemporal",
               4: "U16 nontemporal"}


def bandwidth(dev, torch, clock, reps, nbytes=1 << 31):
    """The device's practical HBM ceiling: copy (2 GiB read + 2 GiB written),
    read-only and write-only probes (2 GiB each) in every launch shape of
    npbx_bandwidth plus the grid-stride npbx_stream_copy; GB/s from HIP
    events, bytes moved = bytes read + bytes written.  The best of each kind
    is the ceiling the stencil figures are held to."""
    from npbench_amd import _native as nat
    src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    src.uniform_()
    dst = torch.empty_like(src)
    h = dev.stream_handle(clock.stream)
    out = {}

    def cp():
        nat.call("npbx_stream_copy", dst.data_ptr(), src.data_ptr(), nbytes, h)
    _, ev = clock.run(cp, reps, WARM_S)
    out["copy grid-stride"] = round(2 * nbytes / (ev / reps) / 1e9, 1)
    for mode, name, moved in ((0, "copy", 2), (1, "read", 1), (2, "write", 1)):
        for v, vname in BW_VARIANTS.items():
            def fn(mode=mode, v=v):
                nat.call("npbx_bandwidth", mode, v, dst.data_ptr(), src.data_ptr(), nbytes, h)
            _, ev = clock.run(fn, reps, WARM_S)
            out[f"{name} {vname}"] = round(moved * nbytes / (ev / reps) / 1e9, 1)
    del src, dst
    best = {k: max(v for n, v in out.items() if n.startswith(k)) for k in ("copy", "read",
                                                                          "write")}
    return {"GBps": out, "best_GBps": best, "bytes": nbytes,
            "best_frac_of_peak": {k: round(v / HBM_PEAK_GBS, 4) for k, v in best.items()}}


def kernel_suite(st, dev, torch, clock, reps, copy_only=False, cpu=False):
    res = {}
    from npbench_amd import _native as nat

    res["bandwidth"] = bandwidth(dev, torch, clock, reps)
    if copy_only:
        return res

    res.update(config_suite(st, dev, torch, clock, reps, cpu))
    # ADI (polybench adi, FP64) at the reference's paper and L presets, on the
    # reference initialisation restored before each call (its values grow
    # from call to call).  Bound: the serial division chain of each line, not
    # HBM (a call moves a few MB); a "lup" is one point of one half step.
    from npbench_amd.benchmarks.polybench.adi import adi as adi_init
    for preset, T, N in (("paper", 100, 200), ("L", 50, 500)):
        u0 = dev.to_device(adi_init.initialize(N), torch.float64)
        u = u0.clone()

        def ad():
            u.copy_(u0)
            st.adi(u, T, stream=clock.stream)
        ad()
        _, ev = clock.run(ad, reps, WARM_S)
        res[f"adi_{preset}_N{N}_T{T}_f64"] = {
            "ms_per_call": round(ev / reps * 1e3, 4),
            "glups": round(2 * T * (N - 2) ** 2 * reps / ev / 1e9, 4),
            "bound": "latency (one serial recurrence per line)"}
        del u0, u

    # 1-D Jacobi (polybench jacobi_1d, FP64) at the paper preset: 7998 sweeps
    # of 32000 points in rounds of 64 sweeps per launch
    from npbench_amd.benchmarks.polybench.jacobi_1d import jacobi_1d as j1_init
    T1, N1 = 4000, 32000
    a0, b0 = (dev.to_device(x, torch.float64) for x in j1_init.initialize(N1))
    a1, b1 = a0.clone(), b0.clone()

    def j1():
        a1.copy_(a0)
        b1.copy_(b0)
        st.jacobi1d(a1, b1, T1, stream=clock.stream)
    j1()
    _, ev = clock.run(j1, reps, WARM_S)
    res[f"jacobi_1d_paper_N{N1}_T{T1}_f64"] = {
        "ms_per_call": round(ev / reps * 1e3, 4),
        "glups": round(2 * (T1 - 1) * (N1 - 2) * reps / ev / 1e9, 2),
        "bound": "latency (rounds of 64 dependent sweeps, a launch each)"}
    del a0, b0, a1, b1

    # 2-D FDTD (polybench fdtd_2d, FP64) at the L and paper presets: two
    # launches per time step; 9 field passes per step (ex, ey, hz read and
    # ex, ey written by the first launch; ex, ey, hz read and hz written by
    # the second), the fields resident in the 256 MB cache
    from npbench_amd.benchmarks.polybench.fdtd_2d import fdtd_2d as fd_init
    for preset, TF, NX, NY in (("L", 150, 800, 900), ("paper", 500, 1000, 1200)):
        f0 = [dev.to_device(x, torch.float64) for x in fd_init.initialize(TF, NX, NY)]
        f1 = [x.clone() for x in f0]

        def fd():
            for x, y in zip(f1, f0):
                x.copy_(y)
            st.fdtd2d(*f1, TF, stream=clock.stream)
        fd()
        _, ev = clock.run(fd, reps, WARM_S)
        res[f"fdtd_2d_{preset}_{NX}x{NY}_T{TF}_f64"] = {
            "ms_per_call": round(ev / reps * 1e3, 4),
            "glups": round(TF * NX * NY * reps / ev / 1e9, 2),
            "field_pass_GBps": round(9 * 8 * TF * NX * NY * reps / ev / 1e9, 1)}
        del f0, f1

    # lid-driven cavity flow (FP64) at the paper and L presets, from the
    # reference initialisation (fields at rest) on every call: one workgroup
    # runs nt * (nit + 3) barrier-separated phases; a "lup" is one interior
    # point of one source, pressure or velocity phase
    from npbench_amd.benchmarks.cavity_flow import cavity_flow as cf_init
    for preset, NYC, NXC, NT, NIT in (("paper", 101, 101, 700, 50), ("L", 201, 201, 100, 20)):
        u, v, p, dx, dy, dt = cf_init.initialize(NYC, NXC)
        c0 = [dev.to_device(x, torch.float64) for x in (u, v, p)]
        c1 = [x.clone() for x in c0]

        def cf():
            for x, y in zip(c1, c0):
                x.copy_(y)
            st.cavity_flow(NXC, NYC, NT, NIT, *c1[:2], dt, dx, dy, c1[2], 1.0, 0.1,
                           stream=clock.stream)
        cf()
        _, ev = clock.run(cf, reps, WARM_S)
        phases = NT * (NIT + 3)
        res[f"cavity_flow_{preset}_{NYC}x{NXC}_nt{NT}_nit{NIT}_f64"] = {
            "ms_per_call": round(ev / reps * 1e3, 4),
            "us_per_phase": round(ev / reps / phases * 1e6, 3),
            "glups": round(NT * (NIT + 2) * (NXC - 2) * (NYC - 2) * reps / ev / 1e9, 3),
            "bound": "latency (one workgroup, a barrier per phase)"}
        del c0, c1

    # channel flow (FP64) at the paper preset from the reference
    # initialisation: time steps until the reference's criterion stops them
    # (989 steps of 53 phases); the call synchronises to read the step count
    from npbench_amd.benchmarks.channel_flow import channel_flow as ch_init
    NYH, NXH, NITH = 101, 101, 50
    u, v, p, dx, dy, dt = ch_init.initialize(NYH, NXH)
    h0 = [dev.to_device(x, torch.float64) for x in (u, v, p)]
    h1 = [x.clone() for x in h0]
    nsteps = []

    def chf():
        for x, y in zip(h1, h0):
            x.copy_(y)
        nsteps.append(st.channel_flow(NITH, *h1[:2], dt, dx, dy, h1[2], 1.0, 0.1, 1.0,
                                      stream=clock.stream))
    chf()
    _, ev = clock.run(chf, reps, WARM_S)
    res[f"channel_flow_paper_{NYH}x{NXH}_nit{NITH}_f64"] = {
        "ms_per_call": round(ev / reps * 1e3, 4), "steps": nsteps[-1],
        "us_per_phase": round(ev / reps / (nsteps[-1] * (NITH + 3)) * 1e6, 3),
        "bound": "latency (one workgroup, a barrier per phase)"}
    del h0, h1

    # Gauss-Seidel 2-D (polybench seidel_2d, FP64) at the paper preset: 99
    # sweeps of 400 x 400 in wavefront slots (rows of successive sweeps two
    # apart run together), each row's recurrence along j serial
    from npbench_amd.benchmarks.polybench.seidel_2d import seidel_2d as sd_init
    TS, NS = 100, 400
    s0 = dev.to_device(sd_init.initialize(NS), torch.float64)
    s1 = s0.clone()

    def sd():
        s1.copy_(s0)
        st.seidel2d(s1, TS, stream=clock.stream)
    sd()
    _, ev = clock.run(sd, reps, WARM_S)
    res[f"seidel_2d_paper_N{NS}_T{TS}_f64"] = {
        "ms_per_call": round(ev / reps * 1e3, 4),
        "glups": round((TS - 1) * (NS - 2) ** 2 * reps / ev / 1e9, 3),
        "bound": "latency (a serial recurrence per row, rows in wavefront slots)"}
    del s0, s1
    return res


def vadv_only(st, dev, torch, clock, reps, path):
    """vadv 256 x 256 x 160 (FP64 and FP32 fields) alone, for profiler runs."""
    from npbench_amd import _native as nat
    from npbench_amd.benchmarks.weather_stencils.vadv.vadv import initialize_config as va_init
    I, J, K = 256, 256, 160
    dtr, *fields = va_init(I, J, K, seed=42)
    nat.set_tuning("vadv.path", path)
    out = {}
    for tag, dt in (("f64", torch.float64), ("f32", torch.float32)):
        d = [dev.to_device(a, dt) for a in fields]

        def va():
            st.vadv(*d, dtr, stream=clock.stream)
        _, ev = clock.run(va, reps, WARM_S)
        out[tag] = {"avg_us": round(ev / reps * 1e6, 3)}
        del d
    nat.set_tuning("vadv.path", 0)
    return out


def cpu_rates():
    """Oracle GLUP/s per kind on bounded samples (1 core): Jacobi 1024^2 x 10
    sweeps and heat 128^3 x 4 sweeps, FP64, the reference NumPy path."""
    from oracle import stencils as ref
    rng = np.random.default_rng(1)
    out = {}
    for kind, n, sweeps in (("jacobi_2d", 1024, 10), ("heat_3d", 128, 4)):
        shape = (n, n) if kind == "jacobi_2d" else (n, n, n)
        A, B = rng.random(shape), rng.random(shape)
        f = ref.jacobi_sweeps if kind == "jacobi_2d" else ref.heat_sweeps
        t0 = time.perf_counter()
        f(A, B, sweeps)
        out[kind] = (n - 2) ** len(shape) * sweeps / (time.perf_counter() - t0) / 1e9
    return out


def cpu_jacobi1d_ms():
    """The NumPy oracle's jacobi_1d at the paper preset (TSTEPS=4000,
    N=32000; 1 core): milliseconds per call, the median of 3."""
    from oracle import stencils as ref
    from npbench_amd.benchmarks.polybench.jacobi_1d import jacobi_1d as j1_init
    times = []
    for _ in range(3):
        A, B = j1_init.initialize(32000)
        t0 = time.perf_counter()
        ref.jacobi_1d_kernel(4000, A, B)
        times.append(time.perf_counter() - t0)
    return round(statistics.median(times) * 1e3, 1)


def cpu_fdtd2d_ms():
    """The NumPy oracle's fdtd_2d at the L preset (TMAX=150, 800 x 900; 1
    core): milliseconds per call, the median of 3."""
    from oracle import stencils as ref
    from npbench_amd.benchmarks.polybench.fdtd_2d import fdtd_2d as fd_init
    times = []
    for _ in range(3):
        ex, ey, hz, fict = fd_init.initialize(150, 800, 900)
        t0 = time.perf_counter()
        ref.fdtd_2d_kernel(150, ex, ey, hz, fict)
        times.append(time.perf_counter() - t0)
    return round(statistics.median(times) * 1e3, 1)


def cpu_cavity_ms(sample_nt=70):
    """The NumPy oracle's cavity_flow at the paper preset (101 x 101, nit=50;
    1 core) on a bounded sample of sample_nt of its 700 time steps, scaled to
    700: every step costs the same (fixed nit, no data-dependent work)."""
    from oracle import stencils as ref
    from npbench_amd.benchmarks.cavity_flow import cavity_flow as cf_init
    u, v, p, dx, dy, dt = cf_init.initialize(101, 101)
    t0 = time.perf_counter()
    ref.cavity_flow(101, 101, sample_nt, 50, u, v, dt, dx, dy, p, 1.0, 0.1)
    return round((time.perf_counter() - t0) * 1e3 * 700 / sample_nt, 1)


def cpu_channel_ms():
    """The NumPy oracle's channel_flow at the paper preset (101 x 101, nit=50,
    to convergence; 1 core): milliseconds for one call."""
    from oracle import stencils as ref
    from npbench_amd.benchmarks.channel_flow import channel_flow as ch_init
    u, v, p, dx, dy, dt = ch_init.initialize(101, 101)
    t0 = time.perf_counter()
    ref.channel_flow(50, u, v, dt, dx, dy, p, 1.0, 0.1, 1.0)
    return round((time.perf_counter() - t0) * 1e3, 1)


def cpu_seidel_ms(sample_sweeps=10):
    """The NumPy oracle's seidel_2d at the paper preset (N=400, 99 sweeps; 1
    core) on a bounded sample of sample_sweeps sweeps, scaled to 99: every
    sweep costs the same."""
    from oracle import stencils as ref
    from npbench_amd.benchmarks.polybench.seidel_2d import seidel_2d as sd_init
    A = sd_init.initialize(400)
    t0 = time.perf_counter()
    ref.seidel_2d_kernel(sample_sweeps + 1, 400, A)
    return round((time.perf_counter() - t0) * 1e3 * 99 / sample_sweeps, 1)


def cpu_adi_ms():
    """The NumPy oracle's ADI at the reference's paper preset (TSTEPS=100,
    N=200; 1 core): milliseconds per call, the median of 3."""
    from oracle import stencils as ref
    from npbench_amd.benchmarks.polybench.adi import adi as adi_init
    times = []
    for _ in range(3):
        u = adi_init.initialize(200)
        t0 = time.perf_counter()
        ref.adi_kernel(100, 200, u)
        times.append(time.perf_counter() - t0)
    return round(statistics.median(times) * 1e3, 1)


def cpu_baseline(n, sweeps):
    """NumPy FP64 oracle, median time per heat sweep on this host (1 core)."""
    from oracle import stencils as ref
    A, B = seeded_grid(n)
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


def only(args, st, dev, batch, tune, torch, clock, A64):
    """One part of the benchmark on its own (for profiler runs)."""
    if args.only == "headline":
        head, _ = batched_fp32(st, dev, batch, tune, torch, clock, args, A64)
        return {"part": "headline", "glups": head["glups"], "cfg": head["cfg"],
                "roofline": head["roofline"]}
    if args.only in ("fp64", "fp32"):
        dt = torch.float64 if args.only == "fp64" else torch.float32
        r = single_grid(st, dev, torch, clock, args, A64, dt, 0)
        return {"part": args.only, "glups": r["glups"], "avg_launch_us": r["launch_s"] * 1e6}
    if args.only == "vadv":
        return {"part": "vadv", **vadv_only(st, dev, torch, clock, args.steps, args.vadv_path)}
    if args.only == "serve":
        return {"part": "serve", **serving(torch, args)}
    if args.only == "copy":
        return {"part": "copy", **kernel_suite(st, dev, torch, clock, args.steps, copy_only=True)}
    return {"part": "kernels", **kernel_suite(st, dev, torch, clock, args.steps)}


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=BATCH)
    ap.add_argument("--graph", type=int, default=1, help="replay a captured HIP graph (1/0)")
    ap.add_argument("--cpu-sweeps", type=int, default=24,
                    help="heat sweeps in the CPU baseline sample (0 = skip every CPU leg: "
                         "oracle times and errors against the oracle)")
    ap.add_argument("--serve-requests", type=int, default=300)
    ap.add_argument("--serve-qps", type=float, default=400.0)
    ap.add_argument("--no-kernels", action="store_true", help="skip the per-kernel lines")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving run")
    ap.add_argument("--shape", type=str, default=None,
                    help="headline launch shape T,rows,seg instead of the autotuner's choice "
                         "(profiler runs)")
    ap.add_argument("--vadv-path", type=int, default=0,
                    help="with --only vadv: the vadv.path knob (0 by shape, 1 workspace, "
                         "2 c' and d' in LDS, 3 c' in registers)")
    ap.add_argument("--only", choices=["headline", "fp64", "fp32", "copy", "kernels", "serve",
                                       "vadv"],
                    help="run one part and print its figures (for profiler runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the library has no multi-GPU path "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)

    import torch
    from npbench_amd import batch, device as dev, stencils as st, tune
    dev.require_device(0)
    clock = Clock(dev, torch.cuda.current_stream())
    A64, B64 = seeded_grid(HEAT_N)
    if args.only:
        print(json.dumps(only(args, st, dev, batch, tune, torch, clock, A64)), flush=True)
        return

    head, lup_step = batched_fp32(st, dev, batch, tune, torch, clock, args, A64)
    f64 = single_grid(st, dev, torch, clock, args, A64, torch.float64, 0)
    f32 = single_grid(st, dev, torch, clock, args, A64, torch.float32, 0)
    err = fp32_error(st, dev, torch, A64, B64)
    lup_sweep = (HEAT_N - 2) ** 3

    line = {
        "metric": METRIC,
        "value": round(head["glups"], 3),
        "unit": "GLUP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["wall"] / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded uniform[0,1), numpy default_rng(42); input sha256 "
                f"{input_hash(A64)})",
        "config": {"workload": f"heat_3d 256^3 x {args.batch} grids, TSTEPS=101 (200 sweeps "
                               "each), FP32 storage, batched",
                   "N": HEAT_N, "tsteps": TSTEPS, "sweeps_per_grid": SWEEPS,
                   "batch": args.batch, "storage": "f32",
                   "shape": head["cfg"], "graph": bool(args.graph and head["cfg"]["T"] > 1),
                   "launches_per_step": head["launches_per_step"]},
        "roofline": head["roofline"],
        "fp64_plain": {
            "value": round(f64["glups"], 3), "unit": "GLUP/s",
            "workload": "heat_3d 256^3, TSTEPS=101, FP64 drop-in, one launch per sweep",
            "ms_per_step": round(f64["wall"] / args.steps * 1e3, 4),
            "roofline": roofline(16.0 * lup_sweep, f64["launch_s"], "heat3d_sweep_vec<double>")},
        "fp32_single_grid": {
            "value": round(f32["glups"], 3), "unit": "GLUP/s",
            "roofline": roofline(8.0 * lup_sweep, f32["launch_s"], "heat3d_sweep_vec<float>")},
        "speedup_vs_fp64_plain": round(head["glups"] / f64["glups"], 3),
        "max_rel_err_vs_fp64": err["inf_norm"],
        "fp32_error_vs_fp64": err,
    }
    if not args.no_serve:
        line["serving"] = serving(torch, args,
                                  cpu_rates=cpu_rates() if args.cpu_sweeps > 0 else None)
    if not args.no_kernels:
        line["kernels"] = kernel_suite(st, dev, torch, clock, 10, cpu=args.cpu_sweeps > 0)
        if args.cpu_sweeps > 0:
            adi = line["kernels"]["adi_paper_N200_T100_f64"]
            adi["cpu_oracle_ms"] = cpu_adi_ms()
            adi["speedup_vs_cpu_oracle"] = round(adi["cpu_oracle_ms"] / adi["ms_per_call"], 1)
            j1 = line["kernels"]["jacobi_1d_paper_N32000_T4000_f64"]
            j1["cpu_oracle_ms"] = cpu_jacobi1d_ms()
            j1["speedup_vs_cpu_oracle"] = round(j1["cpu_oracle_ms"] / j1["ms_per_call"], 1)
            fd = line["kernels"]["fdtd_2d_L_800x900_T150_f64"]
            fd["cpu_oracle_ms"] = cpu_fdtd2d_ms()
            fd["speedup_vs_cpu_oracle"] = round(fd["cpu_oracle_ms"] / fd["ms_per_call"], 1)
            cf = line["kernels"]["cavity_flow_paper_101x101_nt700_nit50_f64"]
            cf["cpu_oracle_ms"] = cpu_cavity_ms()
            cf["cpu_oracle_sample"] = "nt=70 of 700 steps, scaled by 10"
            cf["speedup_vs_cpu_oracle"] = round(cf["cpu_oracle_ms"] / cf["ms_per_call"], 1)
            ch = line["kernels"]["channel_flow_paper_101x101_nit50_f64"]
            ch["cpu_oracle_ms"] = cpu_channel_ms()
            ch["speedup_vs_cpu_oracle"] = round(ch["cpu_oracle_ms"] / ch["ms_per_call"], 1)
            sdl = line["kernels"]["seidel_2d_paper_N400_T100_f64"]
            sdl["cpu_oracle_ms"] = cpu_seidel_ms()
            sdl["cpu_oracle_sample"] = "10 of 99 sweeps, scaled by 9.9"
            sdl["speedup_vs_cpu_oracle"] = round(sdl["cpu_oracle_ms"] / sdl["ms_per_call"], 1)
    if args.cpu_sweeps > 0:
        line["cpu_baseline"] = cpu_baseline(HEAT_N, args.cpu_sweeps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
