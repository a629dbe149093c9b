"""Serving form of the explicit stencils: continuous batching on one GPU.

Requests (a 2-D Jacobi or 3-D heat grid and a number of ping-pong steps)
arrive over time.  The server

* admits a waiting request when its HBM footprint fits the memory budget and
  an active slot is free (admission by footprint, first come first served);
* keeps every admitted grid resident in HBM until it completes;
* advances all active grids together: each tick is ONE persistent tile-queue
  launch (libnpbx ``npbx_rounds_launch``) in which every active grid runs its
  next round of up to ``tblock`` sweeps, so requests that arrive while others
  are on the GPU join at the next tick and their work overlaps on the device;
* retires a grid after its last round, frees its memory, and records the
  latency from arrival to completion on the device.

Each grid follows the round schedule of ``npbx_round_depths`` for its own
sweep range, so its result is bitwise the result of running it alone
(``npbench_amd.stencils`` sweeps), whatever else shares its ticks.

Inputs are generated on the device from a per-request seed (uniform [0, 1)
FP64 drawn by torch's generator, stored in the request's dtype), which stands
in for a payload already resident in HBM.
"""

import collections
import ctypes
import dataclasses
import math
import time

import numpy as np
import torch

from . import _native as nat
from . import device as dev

KIND = {"jacobi_2d": nat.KIND_JACOBI2D, "heat_3d": nat.KIND_HEAT3D}
ELEM = {torch.float64: 8, torch.float32: 4, torch.bfloat16: 2}


@dataclasses.dataclass
class Request:
    rid: int
    kind: str          # "jacobi_2d" | "heat_3d"
    n: int             # side length
    steps: int         # ping-pong iterations (TSTEPS - 1): 2*steps sweeps
    arrival: float     # seconds from the start of the run
    seed: int

    @property
    def ndim(self):
        return 2 if self.kind == "jacobi_2d" else 3

    @property
    def sweeps(self):
        return 2 * self.steps

    @property
    def lups(self):
        return self.sweeps * (self.n - 2) ** self.ndim

    def footprint(self, dtype):
        """Grid bytes plus, for FP32 heat, the packed boundary columns."""
        cols = 16 * self.n ** 2 + 4 if (self.kind == "heat_3d" and dtype == torch.float32) else 0
        return 2 * self.n ** self.ndim * ELEM[dtype] + cols


def make_inputs(req, dtype, device="cuda"):
    """Seeded uniform[0,1) A and B (drawn in FP64, stored in ``dtype``)."""
    gen = torch.Generator(device=device)
    gen.manual_seed(req.seed)
    shape = (req.n,) * req.ndim
    A = torch.rand(shape, generator=gen, dtype=torch.float64, device=device)
    B = torch.rand(shape, generator=gen, dtype=torch.float64, device=device)
    return A.to(dtype), B.to(dtype)


def poisson_mix(count, qps, seed=42, jacobi_frac=0.5, jacobi_sides=(128, 4096),
                heat_sides=(32, 256), steps=(10, 200), multiple=8):
    """The serving mix: Poisson arrivals at ``qps``; a ``jacobi_frac`` share of
    2-D Jacobi requests with side log-uniform in ``jacobi_sides`` and the rest
    3-D heat with side uniform in ``heat_sides``; steps uniform in ``steps``.
    Sides are rounded down to a multiple of ``multiple`` (whole 16-byte
    vectors for the tile kernels)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for rid in range(count):
        t += rng.exponential(1.0 / qps)
        if rng.random() < jacobi_frac:
            kind = "jacobi_2d"
            lo, hi = np.log(jacobi_sides[0]), np.log(jacobi_sides[1])
            n = int(np.floor(np.exp(rng.uniform(lo, hi))))
        else:
            kind = "heat_3d"
            n = int(rng.integers(heat_sides[0], heat_sides[1] + 1))
        n = max(multiple, (n // multiple) * multiple)
        st = int(rng.integers(steps[0], steps[1] + 1))
        out.append(Request(rid, kind, n, st, t, int(rng.integers(1 << 31))))
    return out


def round_schedule(nsweeps, tblock):
    cap = max(4, nsweeps + 2)
    buf = (ctypes.c_int32 * cap)()
    cnt = ctypes.c_int32()
    nat.call("npbx_round_depths", int(nsweeps), int(tblock), buf, cap, ctypes.byref(cnt))
    return [buf[i] for i in range(cnt.value)]


class _Active:
    __slots__ = ("req", "A", "B", "cols", "X", "level", "rounds", "next_round", "admitted")

    def __init__(self, req, A, B, tblock, admitted, cols=None):
        self.req = req
        self.A, self.B = A, B
        self.cols = cols  # packed boundary columns (FP32 heat), or None
        self.level = 0
        self.X = A  # level 0 lives in A
        self.rounds = round_schedule(req.sweeps, tblock)
        self.next_round = 0
        self.admitted = admitted

    @property
    def done(self):
        return self.next_round >= len(self.rounds)


@dataclasses.dataclass
class Completed:
    req: Request
    admitted: float
    finished: float
    A: torch.Tensor = None
    B: torch.Tensor = None

    @property
    def latency(self):
        return self.finished - self.req.arrival


class Server:
    """Continuous-batching stencil server on one GPU (see module docstring).

    hbm_budget: bytes of grid storage that may be resident at once.
    max_active: grids advanced per tick (at most 48 per launch; more are
    split into several launches of the same tick).
    keep_results: keep completed grids' tensors (for checking); otherwise
    their memory is released at completion.
    """

    def __init__(self, dtype=torch.float32, tblock=2, hbm_budget=16 << 30, max_active=48,
                 seg=0, keep_results=False, inflight_ticks=2):
        dev.require_device()
        self.dtype = dtype
        self.code = dev.dtype_code(torch.empty(0, dtype=dtype))
        self.tblock = int(tblock)
        self.budget = int(hbm_budget)
        self.max_active = int(max_active)
        self.seg = int(seg)
        self.keep = keep_results
        self.inflight = max(1, int(inflight_ticks))
        self.stream = torch.cuda.Stream()
        q = ctypes.c_void_p()
        nat.call("npbx_queue_create", ctypes.byref(q))
        self._queue = q
        self.ticks = 0
        self.peak_used = 0      # most grid bytes resident at once
        self.peak_active = 0    # most grids advanced in one tick

    def close(self):
        if self._queue and self._queue.value:
            self.stream.synchronize()
            nat.call("npbx_queue_destroy", self._queue)
            self._queue = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -------------------------------------------------------------- a tick
    def _launch_tick(self, active):
        specs = (nat.RoundSpec * len(active))()
        for s, a in zip(specs, active):
            depth = a.rounds[a.next_round]
            Y = a.B if a.X is a.A else a.A
            s.kind = KIND[a.req.kind]
            s.dtype = self.code
            s.n = a.req.n
            s.A, s.B = dev.ptr(a.A), dev.ptr(a.B)
            s.X, s.Y = dev.ptr(a.X), dev.ptr(Y)
            s.level0 = a.level
            s.depth = depth
            s.cols = dev.ptr(a.cols) if a.cols is not None else None
            a.X = Y
            a.level += depth
            a.next_round += 1
        nat.call("npbx_rounds_launch", specs, len(active), self.tblock, self.seg,
                 self._queue, dev.stream_handle(self.stream))
        self.ticks += 1

    # -------------------------------------------------------------- the loop
    def run(self, requests, realtime=True):
        """Serves ``requests`` (sorted by arrival); returns the Completed list.

        realtime=True waits for each arrival time on the host clock (latency
        includes queueing); realtime=False admits requests as soon as
        capacity allows (a throughput run).
        """
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        waiting = collections.deque()
        active = []
        used = 0
        done = []
        inflight = collections.deque()  # (event, [finished _Active])
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0  # noqa: E731

        def retire(block):
            nonlocal used
            ev, finished = inflight.popleft()
            if block:
                ev.synchronize()
            t = now()
            for a in finished:
                done.append(Completed(a.req, a.admitted, t,
                                      a.A if self.keep else None, a.B if self.keep else None))
                used -= a.req.footprint(self.dtype)

        with torch.cuda.stream(self.stream):
            while pending or waiting or active or inflight:
                t = now()
                while pending and (not realtime or pending[0].arrival <= t):
                    waiting.append(pending.popleft())
                # retire ticks that completed (frees memory for admission)
                while inflight and inflight[0][0].query():
                    retire(False)
                # admission by footprint, first come first served
                while waiting and len(active) < self.max_active:
                    req = waiting[0]
                    fp = req.footprint(self.dtype)
                    if fp > self.budget:
                        raise ValueError(f"request {req.rid} needs {fp} B, more than the budget")
                    if used + fp > self.budget:
                        break
                    waiting.popleft()
                    A, B = make_inputs(req, self.dtype)
                    cols = None
                    if req.kind == "heat_3d" and self.dtype == torch.float32:
                        # boundary columns packed once per request (A and B's
                        # boundaries never change while it is served)
                        n_el = ctypes.c_int64()
                        nat.call("npbx_cols_elems", req.n, ctypes.byref(n_el))
                        cols = torch.empty(n_el.value, dtype=torch.float32, device="cuda")
                        nat.call("npbx_pack_cols", self.code, dev.ptr(A), dev.ptr(B), req.n,
                                 dev.ptr(cols), dev.stream_handle(self.stream))
                    used += fp
                    self.peak_used = max(self.peak_used, used)
                    active.append(_Active(req, A, B, self.tblock, now(), cols))
                if active:
                    if len(inflight) >= self.inflight:
                        retire(True)
                        continue
                    self.peak_active = max(self.peak_active, len(active))
                    self._launch_tick(active)
                    ev = torch.cuda.Event()
                    ev.record(self.stream)
                    finished = [a for a in active if a.done]
                    active = [a for a in active if not a.done]
                    inflight.append((ev, finished))
                elif inflight:
                    retire(True)
                elif pending and realtime:
                    time.sleep(max(0.0, min(pending[0].arrival - now(), 0.001)))
        self.stream.synchronize()
        return done


def summarize(completed, wall_s):
    lat = np.array([c.latency for c in completed]) * 1e3
    lups = sum(c.req.lups for c in completed)
    return {
        "requests": len(completed),
        "p50_ms": round(float(np.percentile(lat, 50)), 3) if len(lat) else None,
        "p99_ms": round(float(np.percentile(lat, 99)), 3) if len(lat) else None,
        "mean_ms": round(float(lat.mean()), 3) if len(lat) else None,
        "glups": round(lups / wall_s / 1e9, 2) if wall_s > 0 else None,
        "wall_s": round(wall_s, 4),
    }


def total_gpu_time_estimate(requests, glups):
    """Seconds of device time the requests need at ``glups``."""
    return sum(r.lups for r in requests) / (glups * 1e9) if glups else math.inf
