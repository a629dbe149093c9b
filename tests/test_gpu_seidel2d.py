"""GPU parity of the 2-D Gauss-Seidel sweeps (npbx_seidel2d) against the FP64
oracle.

Rows run in wavefront slots with the reference's arithmetic, so results must
be bitwise the oracle's.  Sizes cover the smallest grid, rows shorter and
longer than a 256-column block (one, two and three blocks), sweeps spanning
two launches (at most 64 per launch), the division operator's path (zeros),
and the reference's initialisation at its S, M and L presets.
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


def check(st, tsteps, A, what):
    from npbench_amd import device
    n = A.shape[0]
    want = A.copy()
    ref.seidel_2d_kernel(tsteps, n, want)
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
