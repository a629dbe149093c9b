"""Batched grids: many independent Jacobi 2-D / heat 3-D grids of uneven size
advanced together by persistent tile-queue launches (libnpbx plans).

A ``Plan`` fixes the grids (device tensors A, B of one storage dtype), each
grid's sweep range and the temporal-blocking depth.  ``launch()`` enqueues
every round; ``capture()`` records them into a HIP graph on a side stream so
later launches replay the graph.  Results are bitwise those of running each
grid alone with the per-grid entry points.
"""

import ctypes

from . import _native as nat
from . import device as dev

KIND = {"jacobi_2d": nat.KIND_JACOBI2D, "heat_3d": nat.KIND_HEAT3D}


class Grid:
    """One problem: ``kind`` ("jacobi_2d" or "heat_3d"), arrays A and B
    (cubic/square device tensors, the reference's ping-pong pair), and the
    sweep range [first, first + nsweeps)."""

    __slots__ = ("kind", "A", "B", "first", "nsweeps")

    def __init__(self, kind, A, B, nsweeps, first=0):
        if kind not in KIND:
            raise ValueError(f"unknown grid kind {kind!r}")
        ndim = 2 if kind == "jacobi_2d" else 3
        dev.check_device_array("A", A)
        dev.check_device_array("B", B, shape=A.shape, dtype=A.dtype)
        if A.dim() != ndim or len(set(A.shape)) != 1:
            raise ValueError(f"{kind}: expected a cubic {ndim}-D grid, got {tuple(A.shape)}")
        self.kind, self.A, self.B = kind, A, B
        self.first, self.nsweeps = int(first), int(nsweeps)

    @property
    def n(self):
        return self.A.shape[0]

    @property
    def lups(self):
        inner = self.n - 2
        return self.nsweeps * inner ** (2 if self.kind == "jacobi_2d" else 3)


class Plan:
    def __init__(self, grids, tblock=2, seg=0):
        grids = list(grids)
        if not grids:
            raise ValueError("a plan needs at least one grid")
        self.grids = grids  # keeps the tensors alive while the plan exists
        specs = (nat.GridSpec * len(grids))()
        for s, g in zip(specs, grids):
            s.kind = KIND[g.kind]
            s.dtype = dev.dtype_code(g.A)
            s.n = g.n
            s.A = dev.ptr(g.A)
            s.B = dev.ptr(g.B)
            s.first = g.first
            s.nsweeps = g.nsweeps
        handle = ctypes.c_void_p()
        nat.call("npbx_plan_create", specs, len(grids), int(tblock), int(seg),
                 ctypes.byref(handle))
        self._h = handle
        self.tblock = tblock
        self._graph_stream = None

    def info(self):
        launches, tiles, lups = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        nat.call("npbx_plan_info", self._h, ctypes.byref(launches), ctypes.byref(tiles),
                 ctypes.byref(lups))
        return {"launches": launches.value, "tiles": tiles.value, "lups": lups.value}

    def launch(self, stream=None):
        """Enqueues every round (or replays the captured graph) on ``stream``."""
        nat.call("npbx_plan_launch", self._h, dev.stream_handle(stream))

    def capture(self, stream):
        """Records the rounds into a HIP graph on ``stream`` (not the default
        stream); the graph is replayed by later ``launch`` calls."""
        nat.call("npbx_plan_capture", self._h, dev.stream_handle(stream))
        self._graph_stream = stream

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            nat.call("npbx_plan_destroy", self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
