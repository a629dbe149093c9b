"""GPU parity at the full sizes of the measured configurations, against the
FP64 oracle on the configurations' own seeded inputs:

* 2-D Jacobi 1000 x 1000, TSTEPS=101 (200 sweeps), uniform[0,1) A and B;
* 3-D heat 256^3, uniform[0,1) A, B = A.copy(): the FP64 drop-in, and the
  FP32 batched tile plan of the headline launch shape;
* hdiff 256 x 256 x 160, coeff = 0.25 U: FP64, FP32 and BF16 storage;
* vadv 256 x 256 x 160, wcon = 2U - 1: FP64 and FP32 fields (FP64
  elimination), with the reference's wcon in [0, 1) beside it;
* the serving mix at its own distribution (2-D Jacobi sides log-uniform in
  [128, 4096], 3-D heat sides in [32, 256], 10-200 steps, FP32).

FP64 is bitwise.  Reduced-precision forms print their error against the
FP64 oracle (inf-norm relative: max|x - x64| / max|x64|) and assert the
tolerance written in each test.
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


def dev(a, dtype=None):
    from npbench_amd import device
    return device.to_device(a, dtype)


def host(t):
    torch.cuda.synchronize()
    return t.double().cpu().numpy()


def rel_inf(x, y):
    s = np.max(np.abs(y))
    return np.max(np.abs(x - y)) / s if s > 0 else np.max(np.abs(x - y))


def assert_bitwise(x, y, what):
    if not np.array_equal(x, y):
        bad = np.argwhere(x != y)
        first = tuple(bad[0])
        raise AssertionError(f"{what}: {len(bad)} elements differ, first at {first}: "
                             f"{x[first]!r} vs {y[first]!r}; rel_inf={rel_inf(x, y):.3e}")


def quantize(a, dtype):
    return torch.from_numpy(a).to(dtype).double().numpy()


# ------------------------------------------------ 2-D Jacobi 1000^2, 200 sweeps
@pytest.fixture(scope="module")
def jacobi_c1():
    from npbench_amd.benchmarks.polybench.jacobi_2d.jacobi_2d import initialize_uniform
    A, B = initialize_uniform(1000, seed=42)
    ra, rb = A.copy(), B.copy()
    ref.jacobi_2d_kernel(101, ra, rb)
    return A, B, ra, rb


def test_jacobi_1000_fp64_dropin_bitwise(st, jacobi_c1):
    print("jacobi_2d 1000^2 TSTEPS=101 fp64 drop-in", flush=True)
    A, B, ra, rb = jacobi_c1
    a, b = A.copy(), B.copy()
    st.jacobi_2d(101, a, b, precision="fp64")
    assert_bitwise(a, ra, "A")
    assert_bitwise(b, rb, "B")


# BF16 storage rounds every sweep's result to 8 significand bits; over 200
# averaging sweeps the rounding drifts to several percent of the field
# (0.07 in a CPU emulation of the same arithmetic), so BF16 is reported and
# held only to a loose bound; FP32 meets the 1e-5 target.
@pytest.mark.parametrize("precision,tol", [("fp32", 1e-5), ("bf16", 0.15)])
def test_jacobi_1000_reduced_precision(st, jacobi_c1, precision, tol):
    print(f"jacobi_2d 1000^2 TSTEPS=101 {precision}", flush=True)
    A, B, ra, rb = jacobi_c1
    dt = torch.float32 if precision == "fp32" else torch.bfloat16
    da, db = dev(A, dt), dev(B, dt)
    st.jacobi2d_sweeps(da, db, 200)
    err = max(rel_inf(host(da), ra), rel_inf(host(db), rb))
    print(f"jacobi_2d 1000^2 {precision} rel_inf vs FP64 oracle = {err:.3e}")
    assert err <= tol


# ------------------------------------------------ 3-D heat 256^3
@pytest.fixture(scope="module")
def heat_c2():
    from npbench_amd.benchmarks.polybench.heat_3d.heat_3d import initialize_uniform
    A, B = initialize_uniform(256, seed=42)
    ra, rb = A.copy(), B.copy()
    ref.heat_3d_kernel(3, ra, rb)  # 4 sweeps
    return A, B, ra, rb


def test_heat_256_fp64_dropin_bitwise(st, heat_c2):
    """The FP64 drop-in (one launch per sweep) at 256^3, 4 sweeps."""
    print("heat_3d 256^3 TSTEPS=3 fp64 drop-in", flush=True)
    A, B, ra, rb = heat_c2
    a, b = A.copy(), B.copy()
    st.heat_3d(3, a, b, precision="fp64")
    assert_bitwise(a, ra, "A")
    assert_bitwise(b, rb, "B")


def test_heat_256_fp32_batched_headline_shape(st, heat_c2):
    """Two 256^3 FP32 grids in one batch plan with the headline's launch
    shape (depth-3 wave-private tiles of 14 rows, 64 planes per tile, grids of
    equal boundaries): bitwise the plain FP32 sweeps, and within 1e-5 of the
    FP64 oracle."""
    from npbench_amd import _native, batch
    A, B, ra, rb = heat_c2
    plain = [dev(A, torch.float32), dev(B, torch.float32)]
    st.heat3d_sweeps(plain[0], plain[1], 4)
    old_rows, old_seg = _native.get_tuning("heat.rows"), _native.get_tuning("heat.seg")
    try:
        _native.set_tuning("heat.rows", 14)
        _native.set_tuning("heat.seg", 64)
        grids = [batch.Grid("heat_3d", dev(A, torch.float32), dev(B, torch.float32), 4)
                 for _ in range(2)]
        print("heat_3d 256^3 x 2 fp32 plan T=3 R=14 seg=64", flush=True)
        plan = batch.Plan(grids, tblock=3, seg=64)
        plan.launch()
        torch.cuda.synchronize()
        for g in grids:
            assert torch.equal(g.A, plain[0]) and torch.equal(g.B, plain[1])
        plan.close()
    finally:
        _native.set_tuning("heat.rows", old_rows)
        _native.set_tuning("heat.seg", old_seg)
    err = max(rel_inf(host(plain[0]), ra), rel_inf(host(plain[1]), rb))
    print(f"heat_3d 256^3 fp32 rel_inf vs FP64 oracle = {err:.3e}")
    assert err <= 1e-5


# ------------------------------------------------ hdiff 256 x 256 x 160
@pytest.fixture(scope="module")
def hdiff_c3():
    from npbench_amd.benchmarks.weather_stencils.hdiff.hdiff import initialize_config
    inf, out, coeff = initialize_config(256, 256, 160, seed=42)
    r = out.copy()
    ref.hdiff(inf, r, coeff)
    return inf, out, coeff, r


def test_hdiff_256x256x160_fp64_dropin_bitwise(st, hdiff_c3):
    print("hdiff 256x256x160 fp64 drop-in", flush=True)
    inf, out, coeff, r = hdiff_c3
    o = out.copy()
    st.hdiff_dropin(inf, o, coeff, precision="fp64")
    assert_bitwise(o, r, "out_field")


# The limiter's decision r*grad > 0 is discontinuous: FP32-rounded inputs
# flip a few decisions at this size, so the result differs from the FP64
# oracle by up to ~8e-3 relative at those points (the oracle fed the same
# rounded inputs shows the same).  The kernel is therefore held to the oracle
# run on the same rounded inputs; the error against the FP64-input oracle is
# printed with the share of points beyond 1e-5.
@pytest.mark.parametrize("precision,kernel_tol", [("fp32", 1e-6), ("bf16", 4e-3)])
def test_hdiff_256x256x160_reduced_precision(st, hdiff_c3, precision, kernel_tol):
    print(f"hdiff 256x256x160 {precision}", flush=True)
    inf, out, coeff, r = hdiff_c3
    dt = torch.float32 if precision == "fp32" else torch.bfloat16
    rq = out.copy()
    ref.hdiff(quantize(inf, dt), rq, quantize(coeff, dt))
    d_out = dev(out, dt)
    st.hdiff(dev(inf, dt), d_out, dev(coeff, dt))
    got = host(d_out)
    err_k = rel_inf(got, rq)
    err64 = rel_inf(got, r)
    beyond = np.mean(np.abs(got - r) > 1e-5 * np.max(np.abs(r)))
    print(f"hdiff {precision}: rel_inf vs same-input oracle {err_k:.3e}, vs FP64 oracle "
          f"{err64:.3e}, share of points beyond 1e-5: {beyond:.2e}")
    assert err_k <= kernel_tol


# ------------------------------------------------ vadv 256 x 256 x 160
@pytest.fixture(scope="module")
def vadv_c4():
    from npbench_amd.benchmarks.weather_stencils.vadv.vadv import initialize_config
    dtr, *f = initialize_config(256, 256, 160, seed=42)
    r = f[0].copy()
    ref.vadv(r, *f[1:], dtr)
    return dtr, f, r


def test_vadv_256x256x160_fp64_dropin_bitwise(st, vadv_c4):
    print("vadv 256x256x160 fp64 drop-in", flush=True)
    dtr, f, r = vadv_c4
    us = f[0].copy()
    st.vadv_dropin(us, *f[1:], dtr, precision="fp64")
    assert_bitwise(us, r, "utens_stage")


@pytest.mark.parametrize("wcon_dist", ["config", "reference"])
def test_vadv_256x256x160_fp32_fields(st, vadv_c4, wcon_dist):
    """FP32 fields, FP64 elimination.  The configuration's wcon in [-1, 1)
    has near-zero pivots that amplify the FP32 rounding of the inputs (the
    oracle fed the rounded inputs differs from the FP64 oracle by the same
    amount); the reference's wcon in [0, 1) keeps the pivots away from zero.
    The kernel is held to the oracle on the same FP32 inputs (the final FP32
    rounding); the error against the FP64 oracle is printed, and for the
    reference's distribution held to 1e-5."""
    if wcon_dist == "config":
        dtr, f, r = vadv_c4
    else:
        from npbench_amd.benchmarks.weather_stencils.vadv.vadv import initialize
        dtr, *f = initialize(256, 256, 160)
        r = f[0].copy()
        ref.vadv(r, *f[1:], dtr)
    print(f"vadv 256x256x160 fp32 fields, wcon {wcon_dist}", flush=True)
    q = [quantize(a, torch.float32) for a in f]
    rq = q[0].copy()
    ref.vadv(rq, *q[1:], dtr)
    d = [dev(a, torch.float32) for a in f]
    st.vadv(*d, dtr)
    got = host(d[0])
    err_k = rel_inf(got, rq)
    err64 = rel_inf(got, r)
    print(f"vadv fp32 fields ({wcon_dist}): rel_inf vs same-input oracle {err_k:.3e}, "
          f"vs FP64 oracle {err64:.3e}")
    assert err_k <= 1.2e-7
    if wcon_dist == "reference":
        assert err64 <= 1e-5


# ------------------------------------------------ the serving mix
def test_serving_mix_at_its_distribution_alone_vs_mix():
    """20 requests of the serving mix's own distribution (seed 42, 400 QPS,
    FP32, depth-2 rounds on the serving heat tile): every result bitwise the
    same request run alone; the three smallest checked against the FP64
    oracle."""
    from npbench_amd import _native, serve, stencils
    reqs = serve.poisson_mix(20, qps=400.0, seed=42)
    print("serving mix:", [(r.kind, r.n, r.steps) for r in reqs], flush=True)
    assert any(r.kind == "heat_3d" for r in reqs) and any(r.kind == "jacobi_2d" for r in reqs)
    old_rows, old_seg = _native.get_tuning("heat.rows"), _native.get_tuning("heat.seg")
    try:
        _native.set_tuning("heat.rows", 10)
        srv = serve.Server(dtype=torch.float32, tblock=2, seg=32, keep_results=True)
        done = srv.run(reqs, realtime=True)
        assert sorted(c.req.rid for c in done) == list(range(len(reqs)))
        assert srv.peak_active >= 2, "requests never overlapped on the device"
        for c in done:
            A, B = serve.make_inputs(c.req, torch.float32)
            f = stencils.jacobi2d_sweeps if c.req.kind == "jacobi_2d" else stencils.heat3d_sweeps
            f(A, B, c.req.sweeps)
            torch.cuda.synchronize()
            assert torch.equal(c.A, A) and torch.equal(c.B, B), f"request {c.req.rid}"
        smallest = sorted(done, key=lambda c: c.req.lups)[:3]
        for c in smallest:
            A, B = serve.make_inputs(c.req, torch.float64)
            a, b = A.cpu().numpy(), B.cpu().numpy()
            (ref.jacobi_sweeps if c.req.kind == "jacobi_2d" else ref.heat_sweeps)(
                a, b, c.req.sweeps)
            err = max(rel_inf(host(c.A), a), rel_inf(host(c.B), b))
            print(f"request {c.req.rid} {c.req.kind} n={c.req.n} steps={c.req.steps}: "
                  f"rel_inf vs FP64 oracle {err:.3e}")
            assert err <= 1e-5
        srv.close()
    finally:
        _native.set_tuning("heat.rows", old_rows)
        _native.set_tuning("heat.seg", old_seg)
