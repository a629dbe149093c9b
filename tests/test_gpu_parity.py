"""GPU parity of the HIP kernels against the FP64 oracle, through the C ABI.

FP64 (the drop-in) must be bitwise equal to the oracle (which is bitwise the
reference NumPy code), stricter than the 1e-12 relative bound; a failure
message reports the first differing element and the inf-norm relative error.  FP32 and BF16
storage forms are checked against the oracle on the original FP64 inputs with
the tolerances written in each test.
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


def rel_inf(x, ref_):
    d = np.max(np.abs(x - ref_))
    s = np.max(np.abs(ref_))
    return d / s if s > 0 else d


def assert_bitwise(x, y, what):
    if not np.array_equal(x, y):
        bad = np.argwhere(x != y)
        first = tuple(bad[0])
        raise AssertionError(f"{what}: {len(bad)} elements differ, first at {first}: "
                             f"{x[first]!r} vs {y[first]!r}; rel_inf={rel_inf(x, y):.3e}")


# ---------------------------------------------------------------- Jacobi 2-D
@pytest.mark.parametrize("n,tsteps", [(3, 3), (4, 2), (5, 4), (17, 5), (66, 3), (130, 6),
                                      (1000, 3)])
def test_jacobi2d_fp64_bitwise(st, n, tsteps):
    print(f"jacobi2d fp64 n={n} tsteps={tsteps}", flush=True)
    rng = np.random.default_rng(n * 31 + tsteps)
    A = rng.random((n, n))
    B = rng.random((n, n))
    ra, rb = A.copy(), B.copy()
    ref.jacobi_2d_kernel(tsteps, ra, rb)
    da, db = dev(A), dev(B)
    st.jacobi2d_sweeps(da, db, 2 * (tsteps - 1))
    assert_bitwise(host(da), ra, "A")
    assert_bitwise(host(db), rb, "B")


@pytest.mark.parametrize("first,nsweeps", [(0, 1), (1, 1), (1, 4), (0, 7)])
def test_jacobi2d_sweep_parity_offsets(st, first, nsweeps):
    n = 45
    rng = np.random.default_rng(first * 7 + nsweeps)
    A = rng.random((n, n))
    B = rng.random((n, n))
    ra, rb = A.copy(), B.copy()
    # sweep s reads A when s is even
    for s in range(first, first + nsweeps):
        if s % 2 == 0:
            ref.jacobi_sweeps(ra, rb, 1)
        else:
            ref.jacobi_sweeps(rb, ra, 1)
    da, db = dev(A), dev(B)
    st.jacobi2d_sweeps(da, db, nsweeps, first=first)
    assert_bitwise(host(da), ra, "A")
    assert_bitwise(host(db), rb, "B")


@pytest.mark.parametrize("precision,tol", [("fp32", 1e-5), ("bf16", 2e-2)])
def test_jacobi2d_reduced_precision(st, precision, tol):
    n, tsteps = 257, 6
    rng = np.random.default_rng(5)
    A = rng.random((n, n))
    B = rng.random((n, n))
    ra, rb = A.copy(), B.copy()
    ref.jacobi_2d_kernel(tsteps, ra, rb)
    dt = torch.float32 if precision == "fp32" else torch.bfloat16
    da, db = dev(A, dt), dev(B, dt)
    st.jacobi2d_sweeps(da, db, 2 * (tsteps - 1))
    err = max(rel_inf(host(da), ra), rel_inf(host(db), rb))
    print(f"jacobi2d {precision} rel_inf={err:.3e}")
    assert err <= tol


# ---------------------------------------------------------------- heat 3-D
@pytest.mark.parametrize("n,tsteps", [(3, 3), (4, 2), (5, 3), (17, 4), (40, 3), (70, 2),
                                      (130, 2)])
def test_heat3d_fp64_bitwise(st, n, tsteps):
    print(f"heat3d fp64 n={n} tsteps={tsteps}", flush=True)
    rng = np.random.default_rng(n * 13 + tsteps)
    A = rng.random((n, n, n))
    B = rng.random((n, n, n))  # independent boundaries exercise the ping-pong
    ra, rb = A.copy(), B.copy()
    ref.heat_3d_kernel(tsteps, ra, rb)
    da, db = dev(A), dev(B)
    st.heat3d_sweeps(da, db, 2 * (tsteps - 1))
    assert_bitwise(host(da), ra, "A")
    assert_bitwise(host(db), rb, "B")


@pytest.mark.parametrize("precision,tol", [("fp32", 1e-5), ("bf16", 2e-2)])
def test_heat3d_reduced_precision(st, precision, tol):
    n, tsteps = 66, 5
    rng = np.random.default_rng(9)
    A = rng.random((n, n, n))
    B = A.copy()
    ra, rb = A.copy(), B.copy()
    ref.heat_3d_kernel(tsteps, ra, rb)
    dt = torch.float32 if precision == "fp32" else torch.bfloat16
    da, db = dev(A, dt), dev(B, dt)
    st.heat3d_sweeps(da, db, 2 * (tsteps - 1))
    err = max(rel_inf(host(da), ra), rel_inf(host(db), rb))
    print(f"heat3d {precision} rel_inf={err:.3e}")
    assert err <= tol


# ---------------------------------------------------------------- hdiff
def _hdiff_inputs(I, J, K, seed):
    rng = np.random.default_rng(seed)
    inf = rng.random((I + 4, J + 4, K))
    out = rng.random((I, J, K))
    coeff = 0.25 * rng.random((I, J, K))
    return inf, out, coeff


@pytest.mark.parametrize("I,J,K", [(1, 1, 1), (5, 7, 3), (64, 64, 60), (13, 70, 160),
                                   (33, 17, 65)])
def test_hdiff_fp64_bitwise(st, I, J, K):
    print(f"hdiff fp64 {I}x{J}x{K}", flush=True)
    inf, out, coeff = _hdiff_inputs(I, J, K, I * J + K)
    r = out.copy()
    ref.hdiff(inf, r, coeff)
    d_in, d_out, d_c = dev(inf), dev(out), dev(coeff)
    st.hdiff(d_in, d_out, d_c)
    assert_bitwise(host(d_out), r, "out_field")
    assert_bitwise(host(d_in), inf, "in_field untouched")


@pytest.mark.parametrize("nsteps", [2, 3, 5])
def test_hdiff_iterated_fp64_bitwise(st, nsteps):
    I, J, K = 20, 24, 33
    inf, out, coeff = _hdiff_inputs(I, J, K, nsteps)
    r = out.copy()
    ref.hdiff_iterated(inf, r, coeff, nsteps)
    d_in, d_out, d_c = dev(inf), dev(out), dev(coeff)
    st.hdiff(d_in, d_out, d_c, nsteps=nsteps)
    assert_bitwise(host(d_out), r, "out_field")


def quantize(a, dtype):
    """Round FP64 values to the storage dtype (RNE) and back to FP64."""
    return torch.from_numpy(a).to(dtype).double().numpy()


# The flux limiter is discontinuous: quantising the inputs to BF16 flips
# limiter decisions, so the BF16 result differs from the FP64-input oracle
# by O(1) relative (about 0.32 here, a property of the problem, also seen
# with the oracle on BF16 inputs).  Kernel error is therefore measured
# against the oracle run on the same quantised inputs; the error against the
# FP64-input oracle is printed and, for FP32, held to the 1e-5 target.
@pytest.mark.parametrize("precision,kernel_tol,f64_tol", [("fp32", 1e-6, 1e-5),
                                                          ("bf16", 4e-3, None)])
def test_hdiff_reduced_precision(st, precision, kernel_tol, f64_tol):
    I, J, K = 48, 40, 160
    inf, out, coeff = _hdiff_inputs(I, J, K, 3)
    r = out.copy()
    ref.hdiff(inf, r, coeff)
    dt = torch.float32 if precision == "fp32" else torch.bfloat16
    rq = out.copy()
    ref.hdiff(quantize(inf, dt), rq, quantize(coeff, dt))
    d_out = dev(out, dt)
    st.hdiff(dev(inf, dt), d_out, dev(coeff, dt))
    got = host(d_out)
    err_kernel = rel_inf(got, rq)
    err_f64 = rel_inf(got, r)
    print(f"hdiff {precision} rel_inf vs quantised-input oracle={err_kernel:.3e}, "
          f"vs FP64 oracle={err_f64:.3e}")
    assert err_kernel <= kernel_tol
    if f64_tol is not None:
        assert err_f64 <= f64_tol


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16"])
@pytest.mark.parametrize("rows", [4, 5, 6, 7, 16, 40])
@pytest.mark.parametrize("I,J,K", [(9, 61, 16), (33, 130, 24), (70, 5, 8)])
def test_hdiff_two_step_strips(st, hdiff_tblock, precision, rows, I, J, K):
    """The two-step fused pass (hdiff.tblock 2) with strips of 4 to 40
    output rows (hdiff.fuse_rows; every remainder mod 3, the period of its
    rotating row windows), bands of 60 output columns cut by the grid edge
    and K-vectors past K: bitwise the one-launch-per-step iteration at 4 and
    5 steps (a last round of one step)."""
    from npbench_amd import _native
    dt = {"fp64": torch.float64, "fp32": torch.float32, "bf16": torch.bfloat16}[precision]
    inf, out, coeff = _hdiff_inputs(I, J, K, 5 * I + J + K + rows)
    d_in, d_c = dev(inf, dt), dev(coeff, dt)
    old = _native.get_tuning("hdiff.fuse_rows")
    try:
        for nsteps in (4, 5):
            print(f"hdiff two-step {precision} rows={rows} {I}x{J}x{K} nsteps={nsteps}", flush=True)
            hdiff_tblock(1)
            it = dev(out, dt)
            st.hdiff(d_in, it, d_c, nsteps=nsteps)
            hdiff_tblock(2)
            _native.set_tuning("hdiff.fuse_rows", rows)
            fu = dev(out, dt)
            st.hdiff(d_in, fu, d_c, nsteps=nsteps)
            torch.cuda.synchronize()
            same = torch.equal(fu.view(torch.int16), it.view(torch.int16)) \
                if dt == torch.bfloat16 else torch.equal(fu, it)
            assert same, f"nsteps {nsteps}"
    finally:
        _native.set_tuning("hdiff.fuse_rows", old)


@pytest.fixture
def hdiff_shape():
    """Sets the library's hdiff.shape knob for one test and restores it."""
    from npbench_amd import _native
    old = _native.get_tuning("hdiff.shape")
    yield lambda v: _native.set_tuning("hdiff.shape", v)
    _native.set_tuning("hdiff.shape", old)


def _unaligned(a, dt):
    """A device copy of `a` whose data starts one element past a 16-byte
    boundary, which sends hdiff to its one-point-per-thread kernel."""
    buf = torch.empty(a.size + 1, dtype=dt, device="cuda")
    v = buf[1:].view(a.shape)
    v.copy_(torch.from_numpy(a).to(dt))
    return v


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16"])
@pytest.mark.parametrize("I,J,K", [(1, 1, 8), (5, 7, 16), (17, 33, 40), (9, 3, 24),
                                   (64, 64, 160)])
@pytest.mark.parametrize("nsteps", [1, 3])
def test_hdiff_march_shapes_agree(st, hdiff_shape, precision, I, J, K, nsteps):
    """Every marching shape (hdiff.shape 1 .. 13: 1 to 4 columns per lane,
    6 to 18 rows per thread, 4- to 16-byte k-vectors) gives
    bitwise the one-point-per-thread kernel's result, for odd J (the last
    column group shifted left), J below the group width, one-row strips and
    iterated steps; FP64 also bitwise the oracle's."""
    dt = {"fp64": torch.float64, "fp32": torch.float32, "bf16": torch.bfloat16}[precision]
    inf, out, coeff = _hdiff_inputs(I, J, K, I + 3 * J + K + nsteps)
    d_in, d_c = dev(inf, dt), dev(coeff, dt)
    ref_out = _unaligned(out, dt)
    st.hdiff(_unaligned(inf, dt), ref_out, _unaligned(coeff, dt), nsteps=nsteps)
    want = ref_out.contiguous()
    if precision == "fp64":
        r = out.copy()
        ref.hdiff_iterated(inf, r, coeff, nsteps) if nsteps > 1 else ref.hdiff(inf, r, coeff)
        assert_bitwise(host(want), r, "point kernel vs oracle")
    for shape in range(1, 14):
        print(f"hdiff shape={shape} {precision} {I}x{J}x{K} nsteps={nsteps}", flush=True)
        hdiff_shape(shape)
        o = dev(out, dt)
        st.hdiff(d_in, o, d_c, nsteps=nsteps)
        torch.cuda.synchronize()
        same = torch.equal(o.view(torch.int16), want.view(torch.int16)) \
            if dt == torch.bfloat16 else torch.equal(o, want)
        assert same, f"shape {shape}"


# ---------------------------------------------------------------- vadv
def _vadv_inputs(I, J, K, seed):
    rng = np.random.default_rng(seed)
    utens_stage = rng.random((I, J, K))
    u_stage = rng.random((I, J, K))
    wcon = 2.0 * rng.random((I + 1, J, K)) - 1.0
    u_pos = rng.random((I, J, K))
    utens = rng.random((I, J, K))
    return utens_stage, u_stage, wcon, u_pos, utens


# K <= 160 with whole 16-byte vectors per column runs the on-chip kernel
# (c' and d' in LDS; 40, 80 and 160 levels), other shapes the workspace
# kernel; columns not a multiple of 64 leave a partial last wave.
@pytest.mark.parametrize("I,J,K", [(1, 1, 2), (1, 1, 3), (5, 7, 3), (60, 60, 40),
                                   (17, 9, 160), (3, 300, 11), (3, 37, 12), (7, 11, 80),
                                   (5, 13, 42), (2, 64, 160), (2, 3, 200), (4, 5, 158)])
def test_vadv_fp64_bitwise(st, I, J, K):
    print(f"vadv fp64 {I}x{J}x{K}", flush=True)
    f = _vadv_inputs(I, J, K, I + J + K)
    r = f[0].copy()
    ref.vadv(r, *f[1:], 0.15)
    d = [dev(a) for a in f]
    st.vadv(*d, 0.15)
    assert_bitwise(host(d[0]), r, "utens_stage")


@pytest.fixture
def vadv_path():
    """Sets the library's vadv.path knob for one test and restores it."""
    from npbench_amd import _native
    old = _native.get_tuning("vadv.path")
    yield lambda v: _native.set_tuning("vadv.path", v)
    _native.set_tuning("vadv.path", old)


@pytest.mark.parametrize("I,J,K", [(2, 64, 160), (3, 37, 84), (5, 13, 100), (4, 5, 158),
                                   (1, 1, 96), (9, 30, 144)])
def test_vadv_paths_agree(st, vadv_path, I, J, K):
    """Every kernel (vadv.path 1: workspace, 2: c' and d' in LDS, 3: c' in
    registers, 4: coalesced staged loads with c' and d' in registers; 3 and 4
    for 80 < K <= 160) gives the oracle's bits with FP64 fields, and the
    same bits as each other with FP32 fields."""
    f = _vadv_inputs(I, J, K, 3 * I + J + K)
    r = f[0].copy()
    ref.vadv(r, *f[1:], 0.15)
    got32 = []
    for path in (1, 2, 3, 4):
        print(f"vadv path={path} {I}x{J}x{K}", flush=True)
        vadv_path(path)
        d = [dev(a) for a in f]
        st.vadv(*d, 0.15)
        assert_bitwise(host(d[0]), r, f"utens_stage, path {path}")
        d32 = [dev(a, torch.float32) for a in f]
        st.vadv(*d32, 0.15)
        torch.cuda.synchronize()
        got32.append(d32[0])
    assert all(torch.equal(got32[0], g) for g in got32[1:])


@pytest.mark.parametrize("path", [0, 3, 4])
@pytest.mark.parametrize("dtr,wscale", [(1e-250, 0.0), (1e-250, 1e-255), (1e250, 1.0),
                                         (0.15, 1e-300), (1e-240, 0.0), (1e235, 1.0)])
def test_vadv_extreme_pivots(st, vadv_path, path, dtr, wscale):
    """Pivots of magnitude about 1e-250 or 1e250 (tiny wcon leaves them at
    dtr), where the register kernel's short reciprocal hands over to the
    division operator: bitwise the oracle at extreme exponents."""
    I, J, K = 2, 64, 160
    vadv_path(path)
    print(f"vadv extreme pivots path={path} dtr={dtr} wscale={wscale}", flush=True)
    f = list(_vadv_inputs(I, J, K, 5))
    f[2] = wscale * f[2]
    r = f[0].copy()
    ref.vadv(r, *f[1:], dtr)
    d = [dev(a) for a in f]
    st.vadv(*d, dtr)
    assert_bitwise(host(d[0]), r, "utens_stage")


@pytest.mark.parametrize("I,J,K", [(3, 37, 12), (7, 11, 80), (5, 13, 44), (2, 64, 160),
                                   (17, 9, 160), (3, 37, 96), (5, 13, 144),
                                   (2, 3, 200), (4, 5, 36), (3, 5, 7)])
def test_vadv_fp32_fields_shapes(st, I, J, K):
    """FP32 fields through every kernel shape (on-chip 40, 80, 160 levels;
    the staged kernel at K = 160 with its chunk count fixed at compile time
    and at K = 96, 144 with a runtime count, each with a partial last
    workgroup; the workspace kernel above 160 levels or with partial
    vectors): within the final FP32 rounding of the oracle on the same FP32
    inputs."""
    print(f"vadv fp32 fields {I}x{J}x{K}", flush=True)
    f = _vadv_inputs(I, J, K, 7 * I + J + K)
    q = [quantize(a, torch.float32) for a in f]
    rq = q[0].copy()
    ref.vadv(rq, *q[1:], 0.15)
    d = [dev(a, torch.float32) for a in f]
    st.vadv(*d, 0.15)
    assert rel_inf(host(d[0]), rq) <= 1.2e-7


@pytest.mark.parametrize("wcon_lo,f64_tol", [(0.0, 1e-5), (-1.0, None)])
def test_vadv_fp32_fields(st, wcon_lo, f64_tol):
    """FP32 field storage with FP64 elimination.

    wcon in [0, 1) is the reference init (vadv.py:13-17); wcon in [-1, 1) is
    the serving/bench configuration, whose near-zero pivots amplify the FP32
    quantisation of the inputs to ~1e-3 relative in the oracle itself, so
    there only the kernel error (against the oracle on the same FP32-rounded
    inputs, bounded by the final FP32 rounding) is asserted.
    """
    I, J, K = 64, 64, 160
    rng = np.random.default_rng(11)
    f = [rng.random((I, J, K)), rng.random((I, J, K)),
         (1.0 - wcon_lo) * rng.random((I + 1, J, K)) + wcon_lo,
         rng.random((I, J, K)), rng.random((I, J, K))]
    r = f[0].copy()
    ref.vadv(r, *f[1:], 0.15)
    q = [quantize(a, torch.float32) for a in f]
    rq = q[0].copy()
    ref.vadv(rq, *q[1:], 0.15)
    d = [dev(a, torch.float32) for a in f]
    st.vadv(*d, 0.15)
    got = host(d[0])
    err_kernel = rel_inf(got, rq)
    err_f64 = rel_inf(got, r)
    print(f"vadv fp32-fields wcon_lo={wcon_lo}: rel_inf vs quantised-input oracle="
          f"{err_kernel:.3e}, vs FP64 oracle={err_f64:.3e}")
    assert err_kernel <= 1.2e-7
    if f64_tol is not None:
        assert err_f64 <= f64_tol


# ---------------------------------------------------------------- drop-in
def test_dropin_numpy_arrays_roundtrip(st):
    """The drop-in accepts host float64 arrays and updates them in place."""
    n, tsteps = 50, 4
    rng = np.random.default_rng(1)
    A = rng.random((n, n))
    B = rng.random((n, n))
    ra, rb = A.copy(), B.copy()
    ref.jacobi_2d_kernel(tsteps, ra, rb)
    st.jacobi_2d(tsteps, A, B)
    assert_bitwise(A, ra, "A")
    assert_bitwise(B, rb, "B")


def _bf16_rne_bits(x):
    """Direct FP64 -> BF16 round-to-nearest-even for normal numbers in range,
    computed on the FP64 bit pattern (7 of 52 fraction bits kept)."""
    u = np.asarray(x, dtype=np.float64).view(np.uint64).astype(np.uint64)
    sign = (u >> np.uint64(63)) & np.uint64(1)
    mag = u & np.uint64(0x7FFFFFFFFFFFFFFF)
    lsb = (mag >> np.uint64(45)) & np.uint64(1)
    mag = (mag + np.uint64((1 << 44) - 1) + lsb) >> np.uint64(45)  # exp(11) | frac(7)
    exp = (mag >> np.uint64(7)).astype(np.int64) - 1023 + 127
    frac = mag & np.uint64(0x7F)
    return ((sign << np.uint64(15)) | (exp.astype(np.uint64) << np.uint64(7)) | frac).astype(
        np.uint16)


def test_cast_bf16_rne(st):
    x = np.array([1.0, 1.00390625, 1.005859375, 1.0029296875, -2.5, 3.0e38, 1.5e-30,
                  1.0 + 2.0**-8 + 2.0**-30,     # just above a tie: FP32 first would tie
                  -(1.0 + 2.0**-8 + 2.0**-40), 1.0 + 2.0**-8 - 2.0**-40],
                 dtype=np.float64)
    b = st.cast(dev(x), torch.bfloat16)
    got = b.cpu().view(torch.int16).numpy().view(np.uint16)
    np.testing.assert_array_equal(got, _bf16_rne_bits(x))
    # FP32 -> BF16 agrees with torch's RNE conversion
    x32 = np.random.default_rng(0).standard_normal(4096).astype(np.float32)
    b32 = st.cast(dev(x32), torch.bfloat16)
    want = torch.from_numpy(x32).to(torch.bfloat16)
    assert torch.equal(b32.cpu().view(torch.int16), want.view(torch.int16))


def test_bf16_rounding_of_every_float(st):
    """All 2^32 FP32 bit patterns through npbx_cast, which rounds as every
    BF16 store of the library does (the conversion instruction): bitwise
    torch's round-to-nearest-even for every non-NaN pattern (denormals,
    ties, overflow to infinity included) and a NaN for every NaN."""
    chunk = 1 << 28
    base = torch.arange(chunk, dtype=torch.int32, device="cuda")
    for c in range(16):
        print(f"bf16 rounding: patterns {c << 28:#x} ..", flush=True)
        bits = base + ((c << 28) - (1 << 32 if c >= 8 else 0))
        x = bits.view(torch.float32)
        got = st.cast(x, torch.bfloat16).view(torch.int16)
        want = x.to(torch.bfloat16).view(torch.int16)
        nan = (bits & 0x7FFFFFFF) > 0x7F800000
        assert torch.equal(got[~nan], want[~nan]), f"chunk {c}"
        assert bool(((got[nan].to(torch.int32) & 0x7FFF) > 0x7F80).all()), f"chunk {c}: NaN"
        del bits, x, got, want, nan


@pytest.fixture
def hdiff_tblock():
    """Sets the library's hdiff.tblock knob for one test and restores it."""
    from npbench_amd import _native
    old = _native.get_tuning("hdiff.tblock")
    yield lambda tb: _native.set_tuning("hdiff.tblock", tb)
    _native.set_tuning("hdiff.tblock", old)


@pytest.mark.parametrize("precision,tb", [("fp64", 2), ("fp64", 4), ("fp32", 2), ("fp32", 4),
                                          ("bf16", 2)])
@pytest.mark.parametrize("I,J,K", [(1, 1, 8), (20, 24, 40), (70, 33, 16), (64, 96, 24)])
@pytest.mark.parametrize("nsteps", [2, 3, 5, 8])
def test_hdiff_fused_matches_iterated(st, hdiff_tblock, precision, tb, I, J, K, nsteps):
    """Fused steps (hdiff.tblock 2: the two-step register march; 4: 32 x 32
    blocks staged in LDS) give
    bitwise the one-launch-per-step iteration, for blocks cut by the grid
    edge, rounds shorter than tblock and every storage type; FP64 is also
    bitwise the oracle's iteration."""
    print(f"hdiff fused {precision} tb={tb} {I}x{J}x{K} nsteps={nsteps}", flush=True)
    dt = {"fp64": torch.float64, "fp32": torch.float32, "bf16": torch.bfloat16}[precision]
    inf, out, coeff = _hdiff_inputs(I, J, K, I + J + K + nsteps)
    d_in, d_c = dev(inf, dt), dev(coeff, dt)
    hdiff_tblock(1)
    it = dev(out, dt)
    st.hdiff(d_in, it, d_c, nsteps=nsteps)
    hdiff_tblock(tb)
    fu = dev(out, dt)
    st.hdiff(d_in, fu, d_c, nsteps=nsteps)
    torch.cuda.synchronize()
    eq = torch.equal(fu.view(torch.int16), it.view(torch.int16)) if dt == torch.bfloat16 \
        else torch.equal(fu, it)
    assert eq, "fused != iterated"
    if precision == "fp64":
        r = out.copy()
        ref.hdiff_iterated(inf, r, coeff, nsteps)
        assert_bitwise(host(fu), r, "out_field")
