"""The LDS-blocked 2-D Jacobi kernel (knob "jacobi.deep": up to that many
sweeps per launch; "jacobi.deep_waves": 4, 8 or 16 waves per workgroup)
against the one-launch-per-sweep path: bitwise equal for every storage type,
grid sizes from the smallest to several blocks per side (blocks of 64 rows x
48 columns), depths 2 .. 8, every wave count and sweep ranges with odd and
even first sweeps; and the FP64 result against the oracle at the measured
configuration (1000^2, 200 sweeps)."""

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
def deep():
    from npbench_amd import _native
    old = _native.get_tuning("jacobi.deep"), _native.get_tuning("jacobi.deep_waves")

    def set_(depth, waves=16):
        _native.set_tuning("jacobi.deep", depth)
        _native.set_tuning("jacobi.deep_waves", waves)

    yield set_
    set_(*old)


DT = {"f64": torch.float64, "f32": torch.float32, "bf16": torch.bfloat16}


def run(st, A, B, dt, nsweeps, first):
    from npbench_amd import device
    a, b = device.to_device(A, DT[dt]), device.to_device(B, DT[dt])
    st.jacobi2d_sweeps(a, b, nsweeps, first=first)
    torch.cuda.synchronize()
    return a, b


@pytest.mark.parametrize("dt", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("n", [3, 4, 5, 17, 50, 66, 67, 130, 257])
@pytest.mark.parametrize("depth,nsweeps,first,waves", [(2, 5, 0, 16), (3, 7, 1, 8), (8, 2, 0, 4),
                                                       (8, 17, 1, 16), (8, 40, 0, 16),
                                                       (8, 40, 0, 4), (5, 23, 1, 8)])
def test_deep_matches_plain_sweeps(st, deep, dt, n, depth, nsweeps, first, waves):
    rng = np.random.default_rng(n * 7 + depth + nsweeps)
    A, B = rng.random((n, n)), rng.random((n, n))
    deep(0)
    pa, pb = run(st, A, B, dt, nsweeps, first)
    deep(depth, waves)
    print(f"jacobi deep {dt} n={n} depth={depth} waves={waves} sweeps={nsweeps} first={first}",
          flush=True)
    da, db = run(st, A, B, dt, nsweeps, first)
    assert torch.equal(da, pa) and torch.equal(db, pb)


def test_deep_fp64_1000_is_the_oracle(st, deep):
    from npbench_amd.benchmarks.polybench.jacobi_2d.jacobi_2d import initialize_uniform
    A, B = initialize_uniform(1000, seed=42)
    ra, rb = A.copy(), B.copy()
    ref.jacobi_2d_kernel(101, ra, rb)
    deep(8)
    print("jacobi deep f64 1000^2 200 sweeps", flush=True)
    a, b = run(st, A, B, "f64", 200, 0)
    assert np.array_equal(a.cpu().numpy(), ra) and np.array_equal(b.cpu().numpy(), rb)
