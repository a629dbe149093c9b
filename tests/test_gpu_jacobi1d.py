"""GPU parity of the 1-D Jacobi rounds (npbx_jacobi1d) against the FP64 oracle.

Each round sweeps a one-wave register tile up to 64 times with the
reference's arithmetic, so results must be bitwise the oracle's.  Sizes cover
one interior point, one tile and tiles ending inside and exactly at the grid's
end (128 outputs per tile), sweep counts below, at and above a round (rounds
of up to 64 sweeps and the final single sweep), and the reference's
initialisation at its S and paper presets.
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


def run_gpu(st, A, B, tsteps):
    from npbench_amd import device
    a, b = device.to_device(A, torch.float64), device.to_device(B, torch.float64)
    st.jacobi1d(a, b, tsteps)
    torch.cuda.synchronize()
    return a.cpu().numpy(), b.cpu().numpy()


@pytest.mark.parametrize("n", [3, 4, 5, 100, 129, 130, 131, 258, 259, 5000])
@pytest.mark.parametrize("tsteps", [1, 2, 3, 17, 33, 34, 70])
def test_jacobi1d_fp64_bitwise(st, n, tsteps):
    rng = np.random.default_rng(n * 100 + tsteps)
    A, B = rng.random(n), rng.random(n)
    ra, rb = A.copy(), B.copy()
    ref.jacobi_1d_kernel(tsteps, ra, rb)
    print(f"jacobi1d n={n} tsteps={tsteps}", flush=True)
    ga, gb = run_gpu(st, A, B, tsteps)
    assert_bitwise(ga, ra, f"A n={n} tsteps={tsteps}")
    assert_bitwise(gb, rb, f"B n={n} tsteps={tsteps}")


@pytest.mark.parametrize("tsteps,n", [(800, 3200), (4000, 32000)])
def test_jacobi1d_reference_presets_bitwise(st, tsteps, n):
    from npbench_amd.benchmarks.polybench.jacobi_1d import jacobi_1d
    A, B = jacobi_1d.initialize(n)
    ra, rb = A.copy(), B.copy()
    ref.jacobi_1d_kernel(tsteps, ra, rb)
    print(f"jacobi1d preset n={n} tsteps={tsteps}", flush=True)
    ga, gb = run_gpu(st, A, B, tsteps)
    assert_bitwise(ga, ra, "A")
    assert_bitwise(gb, rb, "B")


def test_jacobi1d_dropin_numpy_roundtrip(st):
    from npbench_amd.benchmarks.polybench.jacobi_1d import jacobi_1d_mi355x
    rng = np.random.default_rng(5)
    A, B = rng.random(300), rng.random(300)
    ra, rb = A.copy(), B.copy()
    ref.jacobi_1d_kernel(9, ra, rb)
    assert jacobi_1d_mi355x.kernel(9, A, B) is None
    assert_bitwise(A, ra, "A")
    assert_bitwise(B, rb, "B")


def test_jacobi1d_argument_checks(st):
    from npbench_amd import device
    a = device.to_device(np.ones(10), torch.float64)
    with pytest.raises(ValueError):
        st.jacobi1d(a, device.to_device(np.ones(11), torch.float64), 3)
    with pytest.raises(TypeError):
        st.jacobi1d(device.to_device(np.ones(10), torch.float32), a, 3)
