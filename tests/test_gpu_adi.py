"""GPU parity of the ADI solver (npbx_adi) against the FP64 oracle.

The HIP path performs, per element, the reference's operations in its order
(right-hand side, forward elimination with IEEE division, back substitution),
so every result must be bitwise the oracle's.  Sizes cover one line, lines
that end inside and on a 16-step chunk, one and several workgroups, every
launch path (all time steps in one launch up to 512 lines; one launch per
half step over 256-line workgroups, or over one-wave workgroups with q in
LDS up to 272 steps and in device memory beyond), the first time step's
original boundary columns, and the reference's own initialisation at its S
and paper presets.
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


@pytest.fixture
def adi_path():
    """Sets the library's adi.path knob for one test and restores it."""
    from npbench_amd import _native
    old = _native.get_tuning("adi.path")
    yield lambda path: _native.set_tuning("adi.path", path)
    _native.set_tuning("adi.path", old)


def assert_bitwise(x, y, what):
    if not np.array_equal(x, y):
        bad = np.argwhere(x != y)
        first = tuple(bad[0])
        raise AssertionError(f"{what}: {len(bad)} elements differ, first at {first}: "
                             f"{x[first]!r} vs {y[first]!r}")


def run_gpu(st, u, tsteps):
    from npbench_amd import device
    d = device.to_device(u, torch.float64)
    st.adi(d, tsteps)
    torch.cuda.synchronize()
    return d.cpu().numpy()


CASES = [(3, 1), (4, 2), (5, 3), (17, 2), (18, 1), (19, 2), (33, 3), (100, 4), (200, 2),
         (257, 1), (258, 2), (259, 1), (300, 2), (514, 1), (515, 1), (600, 1)]


@pytest.mark.parametrize("n,tsteps", CASES)
@pytest.mark.parametrize("path", [0, 1, 2, 3])
def test_adi_fp64_bitwise(st, adi_path, n, tsteps, path):
    if path == 1 and n - 2 > 512:
        pytest.skip("the one-launch path holds at most 512 lines")
    u = np.random.default_rng(13 * n + tsteps).random((n, n))
    want = ref.adi_kernel(tsteps, n, u.copy())
    adi_path(path)
    print(f"adi n={n} tsteps={tsteps} path={path}", flush=True)
    assert_bitwise(run_gpu(st, u, tsteps), want, f"adi n={n} tsteps={tsteps} path={path}")


@pytest.mark.parametrize("tsteps,n", [(5, 100), (100, 200)])
def test_adi_reference_presets_bitwise(st, tsteps, n):
    """The reference's initialisation at its S and paper presets; values grow
    to about 1e212 over the paper preset's 100 steps and stay finite."""
    from npbench_amd.benchmarks.polybench.adi import adi
    u = adi.initialize(n)
    want = ref.adi_kernel(tsteps, n, u.copy())
    print(f"adi preset n={n} tsteps={tsteps}", flush=True)
    got = run_gpu(st, u, tsteps)
    assert np.all(np.isfinite(want))
    assert_bitwise(got, want, f"adi preset n={n}")


def test_adi_dropin_numpy_roundtrip(st):
    """The drop-in takes the host array, updates it in place and returns it."""
    from npbench_amd.benchmarks.polybench.adi import adi_mi355x
    n, tsteps = 40, 3
    u = np.random.default_rng(2).random((n, n))
    want = ref.adi_kernel(tsteps, n, u.copy())
    out = adi_mi355x.kernel(tsteps, n, u)
    assert out is u
    assert_bitwise(u, want, "drop-in")


def test_adi_repeated_calls_and_sizes(st):
    """Calls of different sizes in a row share the library's scratch."""
    for n, tsteps in ((260, 2), (20, 3), (600, 1), (37, 2)):
        u = np.random.default_rng(n).random((n, n))
        want = ref.adi_kernel(tsteps, n, u.copy())
        print(f"adi sequence n={n}", flush=True)
        assert_bitwise(run_gpu(st, u, tsteps), want, f"adi sequence n={n}")


def test_adi_short_division_is_the_division_operator(st):
    """The forward passes divide with the IEEE division sequence shortened by
    a reciprocal precomputed from the denominator: on 2^24 seeded quotients
    (half next to a rounding midpoint) inside its no-rescale range it gives
    the division operator's bits."""
    import ctypes
    from npbench_amd import _native
    checked, wrong = ctypes.c_int64(), ctypes.c_int64()
    pairs = 1 << 24
    _native.call("npbx_adi_division_check", pairs, 20261202, ctypes.byref(checked),
                 ctypes.byref(wrong))
    assert checked.value > 0.9 * pairs
    assert wrong.value == 0


def test_adi_wide_range_inputs_take_the_division_operator(st):
    """Inputs spanning 2^-1000 .. 2^1000 push some steps outside the short
    division's range; those chunks are redone with the operator, bitwise."""
    n, tsteps = 40, 2
    rng = np.random.default_rng(9)
    u = rng.random((n, n)) * np.exp2(rng.integers(-1000, 1000, size=(n, n)))
    u[5, 7] = 0.0
    u[9, 11] = 5e-324  # subnormal
    want = ref.adi_kernel(tsteps, n, u.copy())
    assert_bitwise(run_gpu(st, u, tsteps), want, "adi wide range")


def test_adi_argument_checks(st, adi_path):
    from npbench_amd import _native, device
    u = device.to_device(np.ones((600, 600)), torch.float64)
    adi_path(1)
    with pytest.raises(_native.NpbxError):
        st.adi(u, 1)
    with pytest.raises(ValueError):
        st.adi(device.to_device(np.ones((4, 5)), torch.float64), 1)
    with pytest.raises(TypeError):
        st.adi(device.to_device(np.ones((4, 4)), torch.float32), 1)
