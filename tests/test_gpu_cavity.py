"""GPU parity of the lid-driven cavity flow (npbx_cavity_flow) against the
FP64 oracle.

One workgroup runs every phase of every time step with the reference's
arithmetic, so results must be bitwise the oracle's.  Sizes cover the
smallest grid (3 x 3), grids whose interior is narrower and wider than the
1024-thread workgroup, both pressure placements (LDS up to 10240 points,
device memory above), odd and even iteration counts (the pressure ends in
either buffer), nit = 0, and the reference's initialisation at its S and M
presets.
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


def grid(nx, ny):
    return 2 / (nx - 1), 2 / (ny - 1), .1 / ((nx - 1) * (ny - 1))


def check(st, nx, ny, nt, nit, u, v, p, rho, nu, what):
    from npbench_amd import device
    dx, dy, dt = grid(nx, ny)
    want = [u.copy(), v.copy(), p.copy()]
    ref.cavity_flow(nx, ny, nt, nit, want[0], want[1], dt, dx, dy, want[2], rho, nu)
    d = [device.to_device(a, torch.float64) for a in (u, v, p)]
    print(f"cavity_flow {what}", flush=True)
    st.cavity_flow(nx, ny, nt, nit, d[0], d[1], dt, dx, dy, d[2], rho, nu)
    torch.cuda.synchronize()
    for name, g, w in zip(("u", "v", "p"), d, want):
        assert_bitwise(g.cpu().numpy(), w, f"{what} {name}")


@pytest.mark.parametrize("ny,nx", [(3, 3), (3, 7), (6, 3), (5, 9), (34, 33), (40, 37), (101, 101),
                                   (80, 128), (102, 101), (130, 97)])
@pytest.mark.parametrize("nt,nit", [(1, 1), (3, 4), (2, 5), (2, 0)])
def test_cavity_fp64_bitwise(st, ny, nx, nt, nit):
    rng = np.random.default_rng(ny * 1000 + nx * 10 + nit)
    u, v, p = rng.random((ny, nx)), rng.random((ny, nx)) - 0.5, rng.random((ny, nx))
    check(st, nx, ny, nt, nit, u, v, p, 1.3, 0.07, f"{ny}x{nx} nt={nt} nit={nit}")


@pytest.mark.parametrize("ny,nx,nt,nit", [(61, 61, 25, 5), (121, 121, 50, 10)])
def test_cavity_reference_presets_bitwise(st, ny, nx, nt, nit):
    from npbench_amd.benchmarks.cavity_flow import cavity_flow
    u, v, p, dx, dy, dt = cavity_flow.initialize(ny, nx)
    check(st, nx, ny, nt, nit, u, v, p, 1.0, 0.1, f"preset {ny}x{nx} nt={nt} nit={nit}")


def test_cavity_paper_preset_bitwise(st):
    """The paper preset (101 x 101, 700 steps of 50 iterations) runs as
    three launches of at most ~280 steps: the split is invisible."""
    from npbench_amd.benchmarks.cavity_flow import cavity_flow
    u, v, p, dx, dy, dt = cavity_flow.initialize(101, 101)
    check(st, 101, 101, 700, 50, u, v, p, 1.0, 0.1, "preset paper")


def test_cavity_wide_grid(st):
    """An interior row longer than the workgroup (the walk's row stride is 0)."""
    rng = np.random.default_rng(5)
    ny, nx = 4, 1500
    check(st, nx, ny, 2, 3, rng.random((ny, nx)), rng.random((ny, nx)), rng.random((ny, nx)),
          1.0, 0.1, f"{ny}x{nx}")


def test_cavity_repeat_calls_restart_b(st):
    """b starts at zero on every call: a second, smaller call on other fields
    matches a fresh oracle run."""
    rng = np.random.default_rng(6)
    for ny, nx in ((50, 60), (20, 30)):
        check(st, nx, ny, 2, 3, rng.random((ny, nx)), rng.random((ny, nx)),
              rng.random((ny, nx)), 1.0, 0.1, f"repeat {ny}x{nx}")


def test_cavity_dropin_numpy_roundtrip(st):
    from npbench_amd.benchmarks.cavity_flow import cavity_flow, cavity_flow_mi355x
    u, v, p, dx, dy, dt = cavity_flow.initialize(41, 31)
    want = [u.copy(), v.copy(), p.copy()]
    ref.cavity_flow(31, 41, 6, 4, want[0], want[1], dt, dx, dy, want[2], 1.0, 0.1)
    assert cavity_flow_mi355x.cavity_flow(31, 41, 6, 4, u, v, dt, dx, dy, p, 1.0, 0.1) is None
    for g, w, name in zip((u, v, p), want, ("u", "v", "p")):
        assert_bitwise(g, w, name)


def test_cavity_argument_checks(st):
    from npbench_amd import device
    f = device.to_device(np.zeros((4, 5)), torch.float64)
    g = device.to_device(np.zeros((4, 5)), torch.float64)
    with pytest.raises(ValueError):
        st.cavity_flow(5, 4, 1, 1, f, g, 0.1, 0.5, 0.5,
                       device.to_device(np.zeros((5, 4)), torch.float64), 1.0, 0.1)
    with pytest.raises(TypeError):
        st.cavity_flow(5, 4, 1, 1, f, g, 0.1, 0.5, 0.5,
                       device.to_device(np.zeros((4, 5)), torch.float32), 1.0, 0.1)
