"""GPU parity of the channel flow (npbx_channel_flow) against the FP64 oracle.

The kernel runs the reference's arithmetic, so u, v and p must be bitwise the
oracle's whenever the step count agrees (the stopping test sums in another
order than NumPy); every case checks the step count first.  Cases run from
rest (the reference's start) to convergence: tiny grids (hundreds of steps),
both pressure placements (LDS up to 10240 points, device memory above), the
reference's S and M presets, launches capped at a few steps (the host
relaunches until the criterion stops one), and the NumPy drop-in.
"""

import numpy as np
import pytest

from oracle import stencils as ref

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def st():
    from npbench_amd import device, stencils
    device.require_device()
    return stencils


def assert_bitwise(x, y, what):
    if not np.array_equal(x, y):
        bad = np.argwhere(x != y)
        first = tuple(bad[0])
        raise AssertionError(f"{what}: {len(bad)} elements differ, first at {first}: "
                             f"{x[first]!r} vs {y[first]!r}")


def check(st, ny, nx, nit, v0=None, p0=None, max_steps=None, F=1.0, what=""):
    from npbench_amd import device
    dx, dy, dt = 2 / (nx - 1), 2 / (ny - 1), .1 / ((nx - 1) * (ny - 1))
    u = np.zeros((ny, nx))
    v = np.zeros((ny, nx)) if v0 is None else v0
    p = np.ones((ny, nx)) if p0 is None else p0
    want = [u.copy(), v.copy(), p.copy()]
    steps = ref.channel_flow(nit, want[0], want[1], dt, dx, dy, want[2], 1.0, 0.1, F)
    d = [device.to_device(a, torch.float64) for a in (u, v, p)]
    print(f"channel_flow {what} {ny}x{nx} nit={nit} max_steps={max_steps}", flush=True)
    got = st.channel_flow(nit, d[0], d[1], dt, dx, dy, d[2], 1.0, 0.1, F, max_steps=max_steps)
    assert got == steps, f"{what}: {got} steps, the oracle {steps}"
    for name, g, w in zip(("u", "v", "p"), d, want):
        assert_bitwise(g.cpu().numpy(), w, f"{what} {name}")


@pytest.mark.parametrize("ny,nx,nit", [(3, 3, 1), (3, 4, 1), (4, 5, 2), (5, 3, 3), (6, 7, 4),
                                       (40, 37, 3), (100, 102, 2), (103, 100, 2)])
def test_channel_fp64_bitwise(st, ny, nx, nit):
    rng = np.random.default_rng(ny * 100 + nx)
    check(st, ny, nx, nit, 0.01 * rng.random((ny, nx)), 1 + 0.1 * rng.random((ny, nx)),
          what="seeded")


@pytest.mark.parametrize("ny,nx,nit", [(61, 61, 5), (121, 121, 10)])
def test_channel_reference_presets_bitwise(st, ny, nx, nit):
    check(st, ny, nx, nit, what="preset")


@pytest.mark.parametrize("ny,nx", [(6, 7), (110, 100)])
def test_channel_relaunches(st, ny, nx):
    """Launches of at most 7 steps chain to the same result."""
    rng = np.random.default_rng(3)
    check(st, ny, nx, 2, 0.01 * rng.random((ny, nx)), 1 + 0.1 * rng.random((ny, nx)),
          max_steps=7, what="max_steps=7")


def test_channel_zero_force_stops_after_one_step(st):
    """F = 0 from rest: udiff is 0/0 (NaN), which ends the reference's loop."""
    with np.errstate(invalid="ignore"):
        check(st, 8, 9, 2, F=0.0, what="F=0")


def test_channel_dropin_numpy_roundtrip(st):
    from npbench_amd.benchmarks.channel_flow import channel_flow, channel_flow_mi355x
    u, v, p, dx, dy, dt = channel_flow.initialize(21, 17)
    want = [u.copy(), v.copy(), p.copy()]
    steps = ref.channel_flow(3, want[0], want[1], dt, dx, dy, want[2], 1.0, 0.1, 1.0)
    assert channel_flow_mi355x.channel_flow(3, u, v, dt, dx, dy, p, 1.0, 0.1, 1.0) == steps
    for g, w, name in zip((u, v, p), want, ("u", "v", "p")):
        assert_bitwise(g, w, name)


def test_channel_c_loop_entry_point(st):
    """npbx_channel_flow (the whole loop inside the library, for C callers)
    gives the oracle's step count and fields, like the launch-by-launch
    path."""
    from npbench_amd import _native, device
    import ctypes
    ny, nx, nit = 40, 37, 3
    dx, dy, dt = 2 / (nx - 1), 2 / (ny - 1), .1 / ((nx - 1) * (ny - 1))
    rng = np.random.default_rng(12)
    u, v, p = np.zeros((ny, nx)), 0.01 * rng.random((ny, nx)), 1 + 0.1 * rng.random((ny, nx))
    want = [u.copy(), v.copy(), p.copy()]
    steps = ref.channel_flow(nit, want[0], want[1], dt, dx, dy, want[2], 1.0, 0.1, 1.0)
    d = [device.to_device(a, torch.float64) for a in (u, v, p)]
    got = ctypes.c_int64(0)
    print("channel_flow C loop 40x37", flush=True)
    _native.call("npbx_channel_flow", _native.F64, nit, device.ptr(d[0]), device.ptr(d[1]), dt,
                 dx, dy, device.ptr(d[2]), 1.0, 0.1, 1.0, nx, ny, 5, ctypes.byref(got),
                 device.stream_handle(None))
    assert got.value == steps
    for name, g, w in zip(("u", "v", "p"), d, want):
        assert_bitwise(g.cpu().numpy(), w, name)


def test_channel_flow_yields_to_a_harness_timeout(st):
    """The harness's timeout (an interrupt of the main thread) stops a long
    channel flow between two bounded launches: one step of 500 pressure
    iterations per launch on the paper grid (about 2 ms a launch, on the
    order of a thousand launches to convergence) under a 0.3 s limit."""
    from npbench_amd import device
    from npbench_amd.infrastructure import timeout
    from npbench_amd.benchmarks.channel_flow import channel_flow
    u, v, p, dx, dy, dt = channel_flow.initialize(101, 101)
    d = [device.to_device(a, torch.float64) for a in (u, v, p)]

    @timeout.exit_after(0.3)
    def run():
        return st.channel_flow(500, d[0], d[1], dt, dx, dy, d[2], 1.0, 0.1, 1.0, max_steps=1)
    print("channel_flow under a 0.3 s timeout", flush=True)
    with pytest.raises(KeyboardInterrupt):
        run()
    torch.cuda.synchronize()
