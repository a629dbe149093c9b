"""GPU parity of the 2-D Gauss-Seidel sweeps (npbx_seidel2d) against the FP64
oracle.

Rows run in wavefront slots with the reference's arithmetic, so results must
be bitwise the oracle's.  Sizes cover the smallest grid, rows shorter and
longer than a 256-column block (one, two and three blocks), sweeps spanning
two slot schedules (at most 64 sweeps each), full 64-row slots (the paper
preset), launches split at arbitrary slots (the "seidel.blocks" budget), the
division operator's path (zeros), the reference's initialisation at its S, M,
L and paper presets, and the largest accepted N.  Large cases are checked
against the oracle's C restatement (pinned bit for bit to the NumPy oracle on
the CPU).
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


def check(st, tsteps, A, what, oracle=None):
    from npbench_amd import device
    n = A.shape[0]
    want = A.copy()
    (oracle or ref.seidel_2d_kernel)(tsteps, n, want)
    d = device.to_device(A, torch.float64)
    print(f"seidel2d {what}", flush=True)
    st.seidel2d(d, tsteps)
    torch.cuda.synchronize()
    assert_bitwise(d.cpu().numpy(), want, what)


@pytest.mark.parametrize("n", [3, 4, 5, 17, 64, 65, 66, 130, 257, 258, 259, 600])
@pytest.mark.parametrize("tsteps", [1, 2, 3, 7])
def test_seidel2d_fp64_bitwise(st, n, tsteps):
    rng = np.random.default_rng(n * 100 + tsteps)
    check(st, tsteps, rng.random((n, n)), f"n={n} tsteps={tsteps}")


def test_seidel2d_two_launches(st):
    rng = np.random.default_rng(9)
    check(st, 70, rng.random((20, 20)), "n=20 tsteps=70 (69 sweeps)")


@pytest.mark.parametrize("tsteps,n", [(8, 50), (15, 100), (40, 200)])
def test_seidel2d_reference_presets_bitwise(st, tsteps, n):
    from npbench_amd.benchmarks.polybench.seidel_2d import seidel_2d
    check(st, tsteps, seidel_2d.initialize(n), f"preset n={n} tsteps={tsteps}")


def test_seidel2d_zero_field(st):
    """Zeros take the division-operator path of every row."""
    check(st, 3, np.zeros((9, 9)), "zeros")


def test_seidel2d_dropin_numpy_roundtrip(st):
    from npbench_amd.benchmarks.polybench.seidel_2d import seidel_2d_mi355x
    rng = np.random.default_rng(8)
    A = rng.random((30, 30))
    want = A.copy()
    ref.seidel_2d_kernel(5, 30, want)
    assert seidel_2d_mi355x.kernel(5, 30, A) is None
    assert_bitwise(A, want, "drop-in")


@pytest.fixture
def seidel_blocks():
    """Sets the slot-blocks-per-launch knob for one test and restores it."""
    from npbench_amd import _native
    old = _native.get_tuning("seidel.blocks")
    yield lambda b: _native.set_tuning("seidel.blocks", b)
    _native.set_tuning("seidel.blocks", old)


@pytest.mark.parametrize("tsteps,n", [(100, 400), (66, 300)])
def test_seidel2d_full_slots_bitwise(st, tsteps, n):
    """Slots of all 64 rows (lanes 0..63 of the recurrence wave), at the
    paper preset and at N=300 with 65 sweeps."""
    from npbench_amd.benchmarks.polybench.seidel_2d import seidel_2d
    from oracle import c_oracle
    check(st, tsteps, seidel_2d.initialize(n), f"full slots n={n} tsteps={tsteps}",
          oracle=c_oracle.seidel_2d_kernel)


@pytest.mark.parametrize("blocks", [1, 7, 64])
def test_seidel2d_launches_split_at_any_slot(st, seidel_blocks, blocks):
    """Launches of 1, 7 and 64 slot-blocks (a slot of N=300 has 2 blocks):
    the split never changes a bit."""
    from oracle import c_oracle
    seidel_blocks(blocks)
    rng = np.random.default_rng(blocks)
    check(st, 70, rng.random((300, 300)), f"n=300 tsteps=70 seidel.blocks={blocks}",
          oracle=c_oracle.seidel_2d_kernel)


def test_seidel2d_largest_n(st):
    """N = 16384, the largest accepted, two sweeps."""
    from oracle import c_oracle
    rng = np.random.default_rng(16384)
    check(st, 3, rng.random((16384, 16384)), "n=16384 tsteps=3",
          oracle=c_oracle.seidel_2d_kernel)

