"""GPU parity of the 2-D FDTD steps (npbx_fdtd2d) against the FP64 oracle.

Two launches per time step perform the reference's statements in order with
its arithmetic, so results must be bitwise the oracle's.  Sizes cover single
rows and columns, grids narrower and wider than a 256-lane row block, and the
reference's initialisation at its S and M presets.
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


def check(st, tmax, ex, ey, hz, fict, what):
    from npbench_amd import device
    want = [ex.copy(), ey.copy(), hz.copy()]
    ref.fdtd_2d_kernel(tmax, *want, fict)
    d = [device.to_device(a, torch.float64) for a in (ex, ey, hz, fict)]
    print(f"fdtd2d {what}", flush=True)
    st.fdtd2d(*d, tmax)
    torch.cuda.synchronize()
    for name, g, w in zip(("ex", "ey", "hz"), d[:3], want):
        assert_bitwise(g.cpu().numpy(), w, f"{what} {name}")


@pytest.mark.parametrize("nx,ny", [(1, 1), (1, 7), (6, 1), (2, 2), (3, 255), (4, 256), (5, 257),
                                   (33, 300), (129, 513)])
@pytest.mark.parametrize("tmax", [1, 2, 7])
def test_fdtd2d_fp64_bitwise(st, nx, ny, tmax):
    rng = np.random.default_rng(nx * 1000 + ny + tmax)
    ex, ey, hz = (rng.random((nx, ny)) for _ in range(3))
    check(st, tmax, ex, ey, hz, rng.random(tmax + 3), f"nx={nx} ny={ny} tmax={tmax}")


@pytest.mark.parametrize("tmax,nx,ny", [(20, 200, 220), (60, 400, 450)])
def test_fdtd2d_reference_presets_bitwise(st, tmax, nx, ny):
    from npbench_amd.benchmarks.polybench.fdtd_2d import fdtd_2d
    check(st, tmax, *fdtd_2d.initialize(tmax, nx, ny), f"preset {nx}x{ny} tmax={tmax}")


def test_fdtd2d_dropin_numpy_roundtrip(st):
    from npbench_amd.benchmarks.polybench.fdtd_2d import fdtd_2d_mi355x
    rng = np.random.default_rng(8)
    ex, ey, hz = (rng.random((30, 40)) for _ in range(3))
    fict = rng.random(5)
    want = [ex.copy(), ey.copy(), hz.copy()]
    ref.fdtd_2d_kernel(5, *want, fict)
    assert fdtd_2d_mi355x.kernel(5, ex, ey, hz, fict) is None
    for g, w, name in zip((ex, ey, hz), want, ("ex", "ey", "hz")):
        assert_bitwise(g, w, name)


def test_fdtd2d_argument_checks(st):
    from npbench_amd import device
    f = device.to_device(np.ones((4, 5)), torch.float64)
    g = device.to_device(np.ones((4, 5)), torch.float64)
    h = device.to_device(np.ones((4, 5)), torch.float64)
    with pytest.raises(ValueError):
        st.fdtd2d(f, g, h, device.to_device(np.ones(2), torch.float64), 3)  # fict too short
    with pytest.raises(ValueError):
        st.fdtd2d(f, g, device.to_device(np.ones((4, 6)), torch.float64),
                  device.to_device(np.ones(3), torch.float64), 3)
