"""Pins the FP64 oracle (CPU only).

The reference has no golden vectors for this path, so each oracle function is
pinned by what the mathematics fixes, chosen so that a plausible mistake (a
dropped term, a wrong weight, a swapped index or sign, an off-by-one slice)
fails at least one check:

  jacobi_2d  affine fixed point; discrete Fourier eigenmode with its exact
             eigenvalue 0.2*(1 + 2cos a + 2cos b); untouched boundaries;
             bitwise against the scalar brute force and the golden vectors.
  heat_3d    affine fixed point; eigenmode 1 + 0.25*(cos a + cos b + cos c - 3);
             the quadratic i^2 gains exactly 0.25 per sweep inside the cone
             the fixed boundary has not reached (all arithmetic exact);
             brute force and golden vectors.
  hdiff      integer linear and quadratic inputs give out == centre exactly;
             coeff = 0 gives the centre; a single spike gives the limiter's
             closed form (c, -c, 1, -c, c) and not the unlimited 1 + 6c;
             x^4 gives in + 24*coeff; brute force and golden vectors.
  vadv       an exact solution built from the Crank-Nicolson structure
             (equal implicit/explicit weights) for arbitrary wcon; wcon = 0
             gives utens + utens_stage; the elimination matches a banded LU
             solve (scipy.linalg.solve_banded); linearity; brute force and
             golden vectors (including K = 2).
  adi        each half step solves the tridiagonal systems it stands for
             (scipy.linalg.solve_banded, with coefficients that tell a, b, c,
             d, e, f apart); whole time steps against solves composed by the
             test; the reference's coefficients (b = 1 + mul2); untouched
             rows 0 and N-1 and the boundary columns set to 1; brute force
             and golden vectors.
  jacobi_1d  discrete Fourier eigenmode with eigenvalue 0.33333 (1 + 2 cos a);
             untouched ends; linearity; brute force and golden vectors.
  fdtd_2d    one step's response to an hz impulse, worked out by hand (pins
             both coefficients, every sign and neighbour direction); constant
             fields are a fixed point; ey's first row holds the last source
             value; linearity; fields that vary along i only keep ex at zero
             and every column equal; brute force and golden vectors.
  cavity     from rest, the lid's first viscous step puts nu*dt/dy^2 on the
             row below it and nothing else (dx != dy); a pressure ramp along
             x or y drives only u or only v by dt/(2 rho d)*2; pressure
             iterations from p = 1 relax the row under the p = 0 lid to
             (2dy^2 + dx^2)/(2(dx^2 + dy^2)) on the second iteration, not the
             first; velocity fields with exact difference quotients give the
             source term's closed forms (the divergence and the cross term)
             through p = -dx^2 dy^2/(2(dx^2 + dy^2)) b; linear fields advect
             by u du/dx and v dv/dy; brute force and golden vectors.
  channel    from rest with p = 1 the first step gives u = F dt on every
             row between the walls and nothing else (the force term; p stays
             1); a field rolled along x gives the rolled result with the same
             step count (the periodic columns wrap, no column is special);
             brute force (to convergence, hundreds of steps) and golden
             vectors with their step counts.
  seidel_2d  one sweep of a 4 x 4 grid holding 81 at (1, 1), worked by hand:
             (1, 1) -> 9, (1, 2) -> 1 (its left neighbour already swept), row
             2 sees the swept row 1 (sum 10): 10/9 and (10 + 10/9)/9 (pins
             the sweep order, the seven-neighbour sum and the recurrence);
             a constant field is a fixed point ((8c + c)/9 = c); brute force
             and golden vectors.

The last section plants mistakes in the oracle (monkeypatched copies) and
checks the pins above catch them.
"""

import glob
import os

import numpy as np
import pytest

from oracle import scalar, stencils as S

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ------------------------------------------------------------------ helpers
def jacobi_affine_ok(kernel=S.jacobi_2d_kernel, n=20, tsteps=4):
    i, j = np.indices((n, n), dtype=np.float64)
    A = 0.5 + 0.25 * i - 0.125 * j
    B = A.copy()
    a, b = A.copy(), B.copy()
    kernel(tsteps, a, b)
    return np.allclose(a, A, rtol=0, atol=1e-13) and np.allclose(b, B, rtol=0, atol=1e-13)


def jacobi_mode_err(kernel=S.jacobi_2d_kernel, n=33, p=2, q=3, tsteps=4):
    i, j = np.indices((n, n), dtype=np.float64)
    al, be = p * np.pi / (n - 1), q * np.pi / (n - 1)
    mode = np.sin(al * i) * np.sin(be * j)
    lam = 0.2 * (1.0 + 2.0 * np.cos(al) + 2.0 * np.cos(be))
    A = mode.copy()
    B = mode.copy()  # boundary of B is the mode's boundary (zero)
    kernel(tsteps, A, B)
    s = 2 * (tsteps - 1)
    return max(np.max(np.abs(A - lam ** s * mode)), np.max(np.abs(B - lam ** (s - 1) * mode)))


def heat_mode_err(kernel=S.heat_3d_kernel, n=17, pqr=(1, 2, 3), tsteps=3):
    idx = np.indices((n, n, n), dtype=np.float64)
    angles = [m * np.pi / (n - 1) for m in pqr]
    mode = np.ones((n, n, n))
    for ax, a in zip(idx, angles):
        mode = mode * np.sin(a * ax)
    lam = 1.0 + 0.25 * sum(np.cos(a) - 1.0 for a in angles)
    A = mode.copy()
    B = mode.copy()
    kernel(tsteps, A, B)
    s = 2 * (tsteps - 1)
    return max(np.max(np.abs(A - lam ** s * mode)), np.max(np.abs(B - lam ** (s - 1) * mode)))


def heat_cone_ok(kernel=S.heat_3d_kernel, n=12, tsteps=3):
    i = np.indices((n, n, n))[0].astype(np.float64)
    A = i * i
    B = A.copy()
    a, b = A.copy(), B.copy()
    kernel(tsteps, a, b)
    sweeps = 2 * (tsteps - 1)
    idx = np.indices((n, n, n))
    dist = np.minimum(idx, n - 1 - idx).min(axis=0)
    mask = dist >= sweeps
    assert mask.any()
    return np.array_equal(a[mask], (A + 0.25 * sweeps)[mask])


def hdiff_spike_ok(fn=S.hdiff, I=9, J=4, K=2, c=0.125):
    in_field = np.zeros((I + 4, J + 4, K))
    p0 = 2 + I // 2  # in-field row of the spike, away from the halo
    in_field[p0, :, :] = 1.0  # depends on i only: a spike along i
    coeff = np.full((I, J, K), c)
    out = np.empty((I, J, K))
    fn(in_field, out, coeff)
    want = np.zeros((I, J, K))
    x0 = p0 - 2
    want[x0] = 1.0                     # limited: both fluxes at the peak are zeroed
    want[x0 - 1] = want[x0 + 1] = -c
    want[x0 - 2] = want[x0 + 2] = c
    return np.array_equal(out, want)


def vadv_exact_err(fn=S.vadv, I=4, J=5, K=12, wlo=-1.0, seed=0):
    rng = np.random.default_rng(seed)
    dtr = 0.15
    X = rng.random((I, J, K))
    wcon = (1.0 - wlo) * rng.random((I + 1, J, K)) + wlo
    u_stage = 0.7 - X
    u_pos = rng.random((I, J, K))
    utens = rng.random((I, J, K))
    utens_stage = dtr * X - dtr * u_pos - utens
    out = utens_stage.copy()
    fn(out, u_stage, wcon, u_pos, utens, dtr)
    want = dtr * (X - u_pos)
    return np.max(np.abs(out - want)) / np.max(np.abs(want))


# ------------------------------------------------------------------ Jacobi
def test_jacobi_affine_fixed_point():
    assert jacobi_affine_ok()


@pytest.mark.parametrize("p,q", [(1, 1), (2, 3), (5, 1)])
def test_jacobi_eigenmode(p, q):
    assert jacobi_mode_err(p=p, q=q) < 1e-13


def test_jacobi_boundaries_untouched():
    rng = np.random.default_rng(0)
    A, B = rng.random((15, 15)), rng.random((15, 15))
    a, b = A.copy(), B.copy()
    S.jacobi_2d_kernel(5, a, b)
    for x, y in ((a, A), (b, B)):
        assert np.array_equal(x[0], y[0]) and np.array_equal(x[-1], y[-1])
        assert np.array_equal(x[:, 0], y[:, 0]) and np.array_equal(x[:, -1], y[:, -1])


@pytest.mark.parametrize("n,tsteps", [(3, 3), (4, 1), (5, 2), (8, 4)])
def test_jacobi_bruteforce_bitwise(n, tsteps):
    rng = np.random.default_rng(n + tsteps)
    A, B = rng.random((n, n)), rng.random((n, n))
    a, b = A.copy(), B.copy()
    S.jacobi_2d_kernel(tsteps, a, b)
    ra, rb = scalar.jacobi_2d_kernel(tsteps, A, B)
    assert np.array_equal(a, ra) and np.array_equal(b, rb)


def test_jacobi_transpose_symmetry():
    rng = np.random.default_rng(3)
    A, B = rng.random((12, 12)), rng.random((12, 12))
    a, b = A.copy(), B.copy()
    S.jacobi_2d_kernel(4, a, b)
    at, bt = A.T.copy(), B.T.copy()
    S.jacobi_2d_kernel(4, at, bt)
    assert np.allclose(a, at.T, rtol=1e-14, atol=0)


# ------------------------------------------------------------------ heat
def test_heat_affine_fixed_point():
    i, j, k = np.indices((10, 10, 10), dtype=np.float64)
    A = 1.0 + 0.5 * i - 0.25 * j + 0.125 * k
    a, b = A.copy(), A.copy()
    S.heat_3d_kernel(4, a, b)
    assert np.array_equal(a, A) and np.array_equal(b, A)


@pytest.mark.parametrize("pqr", [(1, 1, 1), (1, 2, 3), (4, 1, 2)])
def test_heat_eigenmode(pqr):
    assert heat_mode_err(pqr=pqr) < 1e-13


def test_heat_quadratic_cone_exact():
    assert heat_cone_ok()


@pytest.mark.parametrize("n,tsteps", [(3, 2), (4, 3), (6, 2)])
def test_heat_bruteforce_bitwise(n, tsteps):
    rng = np.random.default_rng(n * tsteps)
    A, B = rng.random((n, n, n)), rng.random((n, n, n))
    a, b = A.copy(), B.copy()
    S.heat_3d_kernel(tsteps, a, b)
    ra, rb = scalar.heat_3d_kernel(tsteps, A, B)
    assert np.array_equal(a, ra) and np.array_equal(b, rb)


# ------------------------------------------------------------------ hdiff
def _grid_fn(I, J, K, f):
    p, q, k = np.indices((I + 4, J + 4, K), dtype=np.float64)
    return f(p, q, k)


@pytest.mark.parametrize("f", [lambda p, q, k: 3 * p - 2 * q + k,
                               lambda p, q, k: p * p + q * q - k,
                               lambda p, q, k: p * q + 5.0])
def test_hdiff_polynomial_inputs_keep_centre(f):
    I, J, K = 6, 5, 3
    in_field = _grid_fn(I, J, K, f)
    coeff = np.random.default_rng(1).random((I, J, K))
    out = np.empty((I, J, K))
    S.hdiff(in_field, out, coeff)
    assert np.array_equal(out, in_field[2:-2, 2:-2])


def test_hdiff_zero_coeff_is_centre():
    rng = np.random.default_rng(2)
    in_field = rng.random((9, 8, 4))
    out = np.empty((5, 4, 4))
    S.hdiff(in_field, out, np.zeros((5, 4, 4)))
    assert np.array_equal(out, in_field[2:-2, 2:-2])


def test_hdiff_limiter_spike():
    assert hdiff_spike_ok()


def test_hdiff_quartic_closed_form():
    I, J, K = 7, 3, 2
    in_field = _grid_fn(I, J, K, lambda p, q, k: (p - 4.0) ** 4)
    coeff = np.random.default_rng(4).random((I, J, K))
    out = np.empty((I, J, K))
    S.hdiff(in_field, out, coeff)
    # lap = -(12 x^2 + 2): flux r = -12(2x+1) opposes the gradient everywhere,
    # so nothing is limited and the flux divergence is -24.
    want = in_field[2:-2, 2:-2] + 24.0 * coeff
    assert np.allclose(out, want, rtol=1e-15, atol=1e-12)


@pytest.mark.parametrize("shape", [(1, 1, 1), (3, 4, 5), (5, 2, 3)])
def test_hdiff_bruteforce_bitwise(shape):
    I, J, K = shape
    rng = np.random.default_rng(sum(shape))
    in_field, coeff = rng.random((I + 4, J + 4, K)), rng.random((I, J, K))
    out = np.empty((I, J, K))
    S.hdiff(in_field, out, coeff)
    assert np.array_equal(out, scalar.hdiff(in_field, coeff))


def test_hdiff_iterated_one_step_is_reference():
    rng = np.random.default_rng(8)
    in_field, coeff = rng.random((8, 9, 3)), rng.random((4, 5, 3))
    a, b = np.empty((4, 5, 3)), np.empty((4, 5, 3))
    S.hdiff(in_field, a, coeff)
    S.hdiff_iterated(in_field, b, coeff, 1)
    assert np.array_equal(a, b)
    c = np.empty((4, 5, 3))
    S.hdiff_iterated(in_field, c, coeff, 3)
    assert np.array_equal(c, scalar.hdiff_iterated(in_field, coeff, 3))


# ------------------------------------------------------------------ vadv
@pytest.mark.parametrize("wlo", [0.0, -1.0])
def test_vadv_exact_solution(wlo):
    assert vadv_exact_err(wlo=wlo) < 1e-12


def test_vadv_zero_wind():
    rng = np.random.default_rng(5)
    I, J, K = 3, 4, 9
    f = [rng.random((I, J, K)) for _ in range(4)]
    us, u, up, ut = f
    out = us.copy()
    S.vadv(out, u, np.zeros((I + 1, J, K)), up, ut, 0.15)
    assert np.allclose(out, ut + us, rtol=1e-14, atol=1e-14)


def test_vadv_matches_banded_solver():
    """The elimination against LAPACK's banded solve of the same system.

    The tridiagonal system is assembled here from the reference's
    coefficient definitions (vadv_numpy.py:15-68); the exact-solution test
    above pins that assembly independently of this restatement.
    """
    la = pytest.importorskip("scipy.linalg")
    rng = np.random.default_rng(6)
    I, J, K, dtr = 2, 3, 10, 0.15
    us, u, up, ut = (rng.random((I, J, K)) for _ in range(4))
    w = rng.random((I + 1, J, K))
    out = us.copy()
    S.vadv(out, u, w, up, ut, dtr)
    for i in range(I):
        for j in range(J):
            s = w[i + 1, j] + w[i, j]
            a = np.zeros(K)
            c = np.zeros(K)
            a[1:] = -0.25 * s[1:] * 0.5
            c[:-1] = 0.25 * s[1:] * 0.5
            b = dtr - a - c
            uu = u[i, j]
            corr = np.zeros(K)
            corr[1:] += -a[1:] * (uu[:-1] - uu[1:])
            corr[:-1] += -c[:-1] * (uu[1:] - uu[:-1])
            d = dtr * up[i, j] + ut[i, j] + us[i, j] + corr
            ab = np.zeros((3, K))
            ab[0, 1:] = c[:-1]
            ab[1] = b
            ab[2, :-1] = a[1:]
            x = la.solve_banded((1, 1), ab, d)
            np.testing.assert_allclose(out[i, j], dtr * (x - up[i, j]), rtol=1e-11, atol=1e-12)


def test_vadv_linearity():
    rng = np.random.default_rng(7)
    I, J, K = 3, 3, 8
    w = rng.random((I + 1, J, K))

    def run(fields):
        out = fields[0].copy()
        S.vadv(out, fields[1], w, fields[2], fields[3], 0.15)
        return out

    p = [rng.random((I, J, K)) for _ in range(4)]
    q = [rng.random((I, J, K)) for _ in range(4)]
    comb = [2.0 * x - 3.0 * y for x, y in zip(p, q)]
    assert np.allclose(run(comb), 2.0 * run(p) - 3.0 * run(q), rtol=1e-12, atol=1e-11)


@pytest.mark.parametrize("shape", [(1, 1, 2), (2, 3, 3), (3, 2, 7)])
def test_vadv_bruteforce_bitwise(shape):
    I, J, K = shape
    rng = np.random.default_rng(K)
    us, u, up, ut = (rng.random((I, J, K)) for _ in range(4))
    w = 2.0 * rng.random((I + 1, J, K)) - 1.0
    out = us.copy()
    S.vadv(out, u, w, up, ut, 0.15)
    assert np.array_equal(out, scalar.vadv(us, u, w, up, ut, 0.15))


def test_vadv_coefficient_identities():
    """The identities that let the GPU kernel hand one coefficient per level
    from its pivot wave to its right-hand-side wave: cs_k = gcv_k BET_M is
    the pivot chain's c_k = gcv_k BET_P (BET_M = BET_P), and as_k = acol_k =
    gav_k BET_M = -cs_{k-1} bitwise, because gav_k = -0.25 s_k and gcv_{k-1}
    = 0.25 s_k differ only in sign and round-to-nearest is symmetric.  Checked
    over wcon sums of every magnitude, subnormals, signed zeros and
    infinities (vadv_numpy.py:32-40)."""
    assert S.BET_M == S.BET_P
    rng = np.random.default_rng(3)
    mant = rng.random(200_000) + 1.0
    expo = rng.integers(-1074, 1024, 200_000)
    s = np.ldexp(mant, expo) * rng.choice([-1.0, 1.0], 200_000)
    s = np.concatenate([s, [0.0, -0.0, np.inf, -np.inf, 5e-324, -5e-324, 2.2250738585072014e-308,
                            1.7976931348623157e308, -1.7976931348623157e308]])
    with np.errstate(over="ignore", under="ignore"):
        gav = -0.25 * s
        gcv_prev = 0.25 * s
        as_ = gav * S.BET_M
        acol = gav * S.BET_P
        cs_prev = gcv_prev * S.BET_M
    neg = -cs_prev
    assert np.array_equal(as_.view(np.uint64), neg.view(np.uint64))
    assert np.array_equal(acol.view(np.uint64), neg.view(np.uint64))
    # the correction term's -as_ is then cs_{k-1} itself
    assert np.array_equal((-as_).view(np.uint64), cs_prev.view(np.uint64))


def test_vadv_rejects_k1():
    z = np.zeros((1, 1, 1))
    with pytest.raises(ValueError):
        S.vadv(z.copy(), z, np.zeros((2, 1, 1)), z, z, 0.15)


# ------------------------------------------------------------------ adi
# distinct weights, diagonally dominant: a swapped or wrong weight shows
ADI_GENERAL = dict(a=-0.3, b=1.9, c=-0.45, d=-0.2, e=1.7, f=-0.35)


def _tridiag_solve(lo, mid, hi, rhs, first, last):
    """x[1:-1] of lo x[j-1] + mid x[j] + hi x[j+1] = rhs[j] (j = 1 .. m),
    x[0] = first, x[m+1] = last, by a banded LU solve."""
    import scipy.linalg as la
    m = len(rhs)
    ab = np.zeros((3, m))
    ab[0, 1:] = hi
    ab[1, :] = mid
    ab[2, :-1] = lo
    r = np.array(rhs, dtype=np.float64)
    r[0] -= lo * first
    r[-1] -= hi * last
    return la.solve_banded((1, 1), ab, r)


def adi_columns_expected(u, a, b, c, d, f):
    n = u.shape[0]
    v = np.full((n, n), np.nan)
    v[0, 1:-1] = v[-1, 1:-1] = 1.0
    for i in range(1, n - 1):
        rhs = -d * u[1:-1, i - 1] + (1 + 2 * d) * u[1:-1, i] - f * u[1:-1, i + 1]
        v[1:-1, i] = _tridiag_solve(a, b, c, rhs, 1.0, 1.0)
    return v


def adi_rows_expected(v, u0, a, c, d, e, f):
    n = v.shape[0]
    u = u0.copy()
    u[1:-1, 0] = u[1:-1, -1] = 1.0
    for i in range(1, n - 1):
        rhs = -a * v[i - 1, 1:-1] + (1 + 2 * a) * v[i, 1:-1] - c * v[i + 1, 1:-1]
        u[i, 1:-1] = _tridiag_solve(d, e, f, rhs, 1.0, 1.0)
    return u


@pytest.mark.parametrize("n", [3, 4, 9, 24])
def test_adi_column_sweep_solves_its_systems(n):
    k = ADI_GENERAL
    u = np.random.default_rng(n).random((n, n))
    v = np.full((n, n), np.nan)
    S.adi_column_sweep(u, v, (k["a"], k["b"], k["c"], k["d"], k["f"]))
    want = adi_columns_expected(u, k["a"], k["b"], k["c"], k["d"], k["f"])
    assert np.allclose(v[:, 1:-1], want[:, 1:-1], rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("n", [3, 4, 9, 24])
def test_adi_row_sweep_solves_its_systems(n):
    k = ADI_GENERAL
    rng = np.random.default_rng(100 + n)
    v, u0 = rng.random((n, n)), rng.random((n, n))
    u = u0.copy()
    S.adi_row_sweep(v, u, (k["a"], k["c"], k["d"], k["e"], k["f"]))
    want = adi_rows_expected(v, u0, k["a"], k["c"], k["d"], k["e"], k["f"])
    assert np.allclose(u, want, rtol=1e-13, atol=1e-14)
    assert np.array_equal(u[0], u0[0]) and np.array_equal(u[-1], u0[-1])


def test_adi_coefficients():
    a, b, c, d, e, f = S.adi_coefficients(4, 10)
    # DT = 1/4, DX = DY = 1/10: mul1 = 2*DT/DX^2 = 50, mul2 = DT/DY^2 = 25
    assert a == pytest.approx(-25.0, rel=1e-14) and c == a
    assert d == pytest.approx(-12.5, rel=1e-14) and f == d
    assert b == pytest.approx(26.0, rel=1e-14) and e == b  # b = 1 + mul2


@pytest.mark.parametrize("n,tsteps", [(5, 1), (8, 2), (12, 3)])
def test_adi_kernel_is_composed_solves(n, tsteps):
    """Whole time steps against the test's own solves: the first column
    sweep reads u's original boundary columns, later ones read 1."""
    a, b, c, d, e, f = S.adi_coefficients(tsteps, n)
    u0 = np.random.default_rng(n).random((n, n))
    want = u0.copy()
    for _ in range(tsteps):
        v = adi_columns_expected(want, a, b, c, d, f)
        want = adi_rows_expected(v, want, a, c, d, e, f)
    got = S.adi_kernel(tsteps, n, u0.copy())
    scale = np.max(np.abs(want))
    assert np.max(np.abs(got - want)) <= 1e-11 * scale


def test_adi_boundaries():
    n = 10
    u0 = np.random.default_rng(3).random((n, n))
    u = S.adi_kernel(2, n, u0.copy())
    assert np.array_equal(u[0], u0[0]) and np.array_equal(u[-1], u0[-1])
    assert np.all(u[1:-1, 0] == 1.0) and np.all(u[1:-1, -1] == 1.0)


@pytest.mark.parametrize("n,tsteps", [(3, 1), (4, 3), (7, 2), (11, 4)])
def test_adi_bruteforce_bitwise(n, tsteps):
    u = np.random.default_rng(7 * n + tsteps).random((n, n))
    assert np.array_equal(S.adi_kernel(tsteps, n, u.copy()), scalar.adi(tsteps, n, u))


# ------------------------------------------------------------------ jacobi 1-D
def jacobi1_mode_err(kernel=S.jacobi_1d_kernel, n=41, k=3, tsteps=5):
    i = np.arange(n, dtype=np.float64)
    al = k * np.pi / (n - 1)
    mode = np.sin(al * i)  # zero at both ends
    lam = 0.33333 * (1.0 + 2.0 * np.cos(al))
    A, B = mode.copy(), mode.copy()
    kernel(tsteps, A, B)
    s = 2 * (tsteps - 1)
    return max(np.max(np.abs(A - lam ** s * mode)), np.max(np.abs(B - lam ** (s - 1) * mode)))


@pytest.mark.parametrize("k", [1, 4, 9])
def test_jacobi1d_eigenmode(k):
    assert jacobi1_mode_err(k=k) < 1e-13


def test_jacobi1d_ends_untouched_and_linear():
    rng = np.random.default_rng(4)
    A, B, C, D = (rng.random(30) for _ in range(4))
    a, b = A.copy(), B.copy()
    S.jacobi_1d_kernel(6, a, b)
    assert a[0] == A[0] and a[-1] == A[-1] and b[0] == B[0] and b[-1] == B[-1]
    c, d = C.copy(), D.copy()
    S.jacobi_1d_kernel(6, c, d)
    e, f = A + 2.0 * C, B + 2.0 * D
    S.jacobi_1d_kernel(6, e, f)
    assert np.allclose(e, a + 2.0 * c, rtol=1e-13, atol=1e-13)
    assert np.allclose(f, b + 2.0 * d, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("n,tsteps", [(3, 3), (4, 2), (12, 6), (40, 9)])
def test_jacobi1d_bruteforce_bitwise(n, tsteps):
    rng = np.random.default_rng(n + tsteps)
    A, B = rng.random(n), rng.random(n)
    a, b = A.copy(), B.copy()
    S.jacobi_1d_kernel(tsteps, a, b)
    ra, rb = scalar.jacobi_1d(tsteps, A, B)
    assert np.array_equal(a, ra) and np.array_equal(b, rb)


def test_planted_jacobi1d_mistakes_are_caught(monkeypatch):
    """A dropped neighbour and a shifted neighbour both break the eigenmode."""
    for mistake in ("drop", "shift"):
        def sweep(src, dst, mistake=mistake):
            right = src[2:] if mistake == "drop" else src[1:-1]
            acc = src[:-2] + src[1:-1] if mistake != "drop" else src[1:-1] + src[1:-1]
            dst[1:-1] = 0.33333 * (acc + right)
        monkeypatch.setattr(S, "_jacobi1_sweep", sweep)
        assert jacobi1_mode_err() > 1e-6, mistake


# ------------------------------------------------------------------ fdtd 2-D
def fdtd_impulse(kernel=None):
    kernel = kernel or S.fdtd_2d_kernel  # looked up per call (planted mistakes patch it)
    n = 7
    ex, ey, hz = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
    hz[3, 3] = 1.0
    kernel(1, ex, ey, hz, np.zeros(1))
    return ex, ey, hz


def test_fdtd2d_impulse_response():
    """hz = delta at (3, 3), e = 0, source 0, one step: ey[3,3] = -0.5,
    ey[4,3] = 0.5, ex[3,3] = -0.5, ex[3,4] = 0.5; then hz[3,3] = 1 - 0.7 * 2,
    hz[2,3] = hz[3,2] = -0.7 * -0.5, hz[4,3] = hz[3,4] = -0.7 * ... (see below)."""
    ex, ey, hz = fdtd_impulse()
    want_ey = np.zeros((7, 7))
    want_ey[3, 3], want_ey[4, 3] = -0.5, 0.5
    want_ex = np.zeros((7, 7))
    want_ex[3, 3], want_ex[3, 4] = -0.5, 0.5
    assert np.array_equal(ey, want_ey) and np.array_equal(ex, want_ex)
    # hz[i,j] -= 0.7 ((ex[i,j+1] - ex[i,j]) + ey[i+1,j] - ey[i,j])
    want_hz = np.zeros((7, 7))
    want_hz[3, 3] = 1.0 - 0.7 * 2.0        # (0.5 + 0.5) + (0.5 + 0.5)
    want_hz[2, 3] = -0.7 * -0.5            # ey[3,3] - ey[2,3] enters as +ey[i+1,j]
    want_hz[3, 2] = -0.7 * -0.5            # ex[3,3] enters as +ex[i,j+1]
    want_hz[4, 3] = -0.7 * (0.0 - 0.5)     # -ey[4,3]
    want_hz[3, 4] = -0.7 * (0.0 - 0.5)     # -ex[3,4]
    assert np.allclose(hz, want_hz, rtol=0, atol=1e-15)


def test_fdtd2d_fixed_point_and_source_row():
    n, tmax = 9, 5
    ex, ey, hz = np.full((n, n + 2), 0.25), np.full((n, n + 2), 2.0), np.full((n, n + 2), -1.5)
    fict = np.full(tmax, 2.0)
    a, b, c = ex.copy(), ey.copy(), hz.copy()
    S.fdtd_2d_kernel(tmax, a, b, c, fict)
    assert np.array_equal(a, ex) and np.array_equal(b, ey) and np.array_equal(c, hz)
    rng = np.random.default_rng(2)
    a, b, c = (rng.random((6, 8)) for _ in range(3))
    f = rng.random(tmax)
    S.fdtd_2d_kernel(tmax, a, b, c, f)
    assert np.all(b[0] == f[tmax - 1])


def test_fdtd2d_linearity_and_plane_waves():
    rng = np.random.default_rng(3)
    x = [rng.random((8, 9)) for _ in range(3)] + [rng.random(6)]
    y = [rng.random((8, 9)) for _ in range(3)] + [rng.random(6)]
    def run(f):
        a, b, c = (v.copy() for v in f[:3])
        S.fdtd_2d_kernel(6, a, b, c, f[3])
        return a, b, c
    rx, ry = run(x), run(y)
    rz = run([u + 3.0 * v for u, v in zip(x, y)])
    for p, q, r in zip(rx, ry, rz):
        assert np.allclose(r, p + 3.0 * q, rtol=1e-12, atol=1e-12)
    # fields that vary along i only: ex stays 0 and all columns evolve alike,
    # except near the last column (hz's is never updated; that reaches one
    # column further in per step)
    col = rng.random((8, 1))
    ex, ey, hz = np.zeros((8, 20)), np.repeat(col, 20, axis=1), np.repeat(2 * col, 20, axis=1)
    S.fdtd_2d_kernel(6, ex, ey, hz, rng.random(6))
    assert np.all(ex[:, :10] == 0.0)
    assert np.all(ey[:, :10] == ey[:, :1]) and np.all(hz[:, :10] == hz[:, :1])


@pytest.mark.parametrize("nx,ny,tmax", [(1, 1, 2), (1, 5, 3), (5, 1, 3), (4, 6, 4), (7, 5, 6)])
def test_fdtd2d_bruteforce_bitwise(nx, ny, tmax):
    rng = np.random.default_rng(nx * 31 + ny)
    ex, ey, hz = (rng.random((nx, ny)) for _ in range(3))
    fict = rng.random(tmax)
    a, b, c = ex.copy(), ey.copy(), hz.copy()
    S.fdtd_2d_kernel(tmax, a, b, c, fict)
    ra, rb, rc = scalar.fdtd_2d(tmax, ex, ey, hz, fict)
    assert np.array_equal(a, ra) and np.array_equal(b, rb) and np.array_equal(c, rc)


def test_planted_fdtd2d_mistakes_are_caught(monkeypatch):
    """A wrong coefficient, a flipped neighbour and a dropped term all move
    the impulse response."""
    def make(mistake):
        def kernel(TMAX, ex, ey, hz, fict):
            for t in range(TMAX):
                ey[0, :] = fict[t]
                up = hz[:-1, :] if mistake != "flip" else hz[1:, :]
                ey[1:, :] = ey[1:, :] - 0.5 * (hz[1:, :] - up)
                ex[:, 1:] = ex[:, 1:] - (0.6 if mistake == "coef" else 0.5) * (hz[:, 1:] - hz[:, :-1])
                curl = ex[:-1, 1:] - ex[:-1, :-1]
                if mistake != "drop":
                    curl = curl + ey[1:, :-1]
                curl = curl - ey[:-1, :-1]
                hz[:-1, :-1] = hz[:-1, :-1] - 0.7 * curl
        return kernel
    for mistake in ("coef", "flip", "drop"):
        monkeypatch.setattr(S, "fdtd_2d_kernel", make(mistake))
        with pytest.raises(AssertionError):
            test_fdtd2d_impulse_response()


# ------------------------------------------------------------------ cavity flow
def cavity_run(ny, nx, nt, nit, u=None, v=None, p=None, rho=1.0, nu=0.1):
    """Runs S.cavity_flow (looked up per call: planted mistakes patch it) on
    the reference grid spacing and returns (u, v, p, dx, dy, dt)."""
    dx, dy, dt = 2 / (nx - 1), 2 / (ny - 1), .1 / ((nx - 1) * (ny - 1))
    u = np.zeros((ny, nx)) if u is None else u.copy()
    v = np.zeros((ny, nx)) if v is None else v.copy()
    p = np.zeros((ny, nx)) if p is None else p.copy()
    S.cavity_flow(nx, ny, nt, nit, u, v, dt, dx, dy, p, rho, nu)
    return u, v, p, dx, dy, dt


def test_cavity_lid_diffusion():
    """From rest, step 1 only sets the lid; in step 2 nothing but the lid's
    viscous term moves the row below it: u = nu dt / dy^2 there (dx != dy),
    and v and p stay zero."""
    ny, nx, nu = 6, 9, 0.1
    u, v, p, dx, dy, dt = cavity_run(ny, nx, 1, 3, nu=nu)
    assert np.all(u[-1] == 1.0) and np.all(u[:-1] == 0.0) and not v.any() and not p.any()
    u, v, p, dx, dy, dt = cavity_run(ny, nx, 2, 3, nu=nu)
    assert dx != dy
    assert np.all(u[-2, 1:-1] == nu * (dt / dy ** 2))
    assert np.all(u[1:-2] == 0.0) and np.all(u[:, [0, -1]][:-1] == 0.0) and np.all(u[-1] == 1.0)
    assert not v.any() and not p.any()


def test_cavity_pressure_gradient():
    """With no flow and no pressure iterations, a pressure ramp along x
    (columns) moves only u and one along y (rows) only v, each by
    -dt/(2 rho d) (p[+1] - p[-1])."""
    ny, nx, rho = 9, 5, 2.0
    i, j = np.indices((ny, nx), dtype=np.float64)
    u, v, p, dx, dy, dt = cavity_run(ny, nx, 1, 0, p=3 * j, rho=rho, nu=0.0)
    assert np.allclose(u[1:-1, 1:-1], -dt / (2 * rho * dx) * 6, rtol=1e-15, atol=0)
    assert not v.any() and np.array_equal(p, 3 * j)
    u, v, p, dx, dy, dt = cavity_run(ny, nx, 1, 0, p=3 * i, rho=rho, nu=0.0)
    assert np.allclose(v[1:-1, 1:-1], -dt / (2 * rho * dy) * 6, rtol=1e-15, atol=0)
    assert not u[:-1].any()


def test_cavity_pressure_iterations():
    """No flow (b = 0), p = 1 everywhere: the first iteration keeps p = 1 but
    zeroes the lid row; the second relaxes the row below to
    (2 dy^2 + dx^2) / (2 (dx^2 + dy^2)) and leaves the rest at 1, boundary
    copies included (corner (0, 0) and column nx-1 too)."""
    ny, nx = 7, 11
    one = np.ones((ny, nx))
    u, v, p, dx, dy, dt = cavity_run(ny, nx, 1, 1, p=one)
    assert np.allclose(p[:-1], 1.0, rtol=0, atol=2e-16) and np.all(p[-1] == 0.0)
    u, v, p, dx, dy, dt = cavity_run(ny, nx, 1, 2, p=one)
    relaxed = (2 * dy ** 2 + dx ** 2) / (2 * (dx ** 2 + dy ** 2))
    assert abs(relaxed - 1.0) > 0.05
    assert np.allclose(p[-2], relaxed, rtol=1e-15, atol=0)
    assert np.allclose(p[:-2], 1.0, rtol=0, atol=4e-16) and np.all(p[-1] == 0.0)


def test_cavity_source_term():
    """Velocity fields whose difference quotients are exact (dx = 1/2,
    dy = 1/4): p after one iteration from 0 is -c b with
    c = dx^2 dy^2 / (2 (dx^2 + dy^2)), where b = rho (1/dt (ux + vy) - ux^2 -
    2 uy vx - vy^2)."""
    ny, nx, rho = 9, 5, 1.5
    i, j = np.indices((ny, nx), dtype=np.float64)
    dx, dy = 0.5, 0.25
    c = dx ** 2 * dy ** 2 / (2 * (dx ** 2 + dy ** 2))
    # the cross term: u varies along y only, v along x only
    u, v, p, dx_, dy_, dt = cavity_run(ny, nx, 1, 1, u=3 * i, v=5 * j, rho=rho)
    assert (dx_, dy_) == (dx, dy)
    b = rho * (-2 * (6 / (2 * dy)) * (10 / (2 * dx)))
    assert np.allclose(p[1:-1, 1:-1], -c * b, rtol=1e-14, atol=0)
    # the divergence and its squares: u along x, v along y
    u, v, p, _, _, dt = cavity_run(ny, nx, 1, 1, u=3 * j, v=5 * i, rho=rho)
    ux, vy = 6 / (2 * dx), 10 / (2 * dy)
    b = rho * ((ux + vy) / dt - ux ** 2 - vy ** 2)
    assert np.allclose(p[1:-1, 1:-1], -c * b, rtol=1e-14, atol=0)


def test_cavity_advection():
    """Linear fields u = a x, v = b y with no pressure and no viscosity: one
    step gives u (1 - a dt/dx') and v (1 - b dt/dy') in grid units (the upwind
    difference of a linear field is exact), and the boundary conditions."""
    ny, nx = 9, 5
    i, j = np.indices((ny, nx), dtype=np.float64)
    u, v, p, dx, dy, dt = cavity_run(ny, nx, 1, 0, u=0.5 * j, v=0.25 * i, nu=0.0)
    assert np.allclose(u[1:-1, 1:-1], (0.5 * j * (1 - 0.5 * dt / dx))[1:-1, 1:-1], rtol=1e-15, atol=0)
    assert np.allclose(v[1:-1, 1:-1], (0.25 * i * (1 - 0.25 * dt / dy))[1:-1, 1:-1], rtol=1e-15,
                       atol=0)
    assert np.all(u[-1] == 1.0) and not u[:-1, [0, -1]].any() and not u[0].any()
    assert not v[[0, -1]].any() and not v[:, [0, -1]].any()


@pytest.mark.parametrize("ny,nx,nt,nit", [(3, 3, 2, 1), (5, 7, 3, 4), (6, 4, 2, 3), (8, 9, 3, 0),
                                          (7, 5, 3, 5)])
def test_cavity_bruteforce_bitwise(ny, nx, nt, nit):
    rng = np.random.default_rng(ny * 37 + nx)
    u, v, p = rng.random((ny, nx)), rng.random((ny, nx)) - 0.5, rng.random((ny, nx))
    a, b, c, dx, dy, dt = cavity_run(ny, nx, nt, nit, u, v, p, rho=1.25, nu=0.07)
    ra, rb, rc = scalar.cavity_flow(nx, ny, nt, nit, u, v, dt, dx, dy, p, 1.25, 0.07)
    assert np.array_equal(a, ra) and np.array_equal(b, rb) and np.array_equal(c, rc)


CAVITY_PINS = (test_cavity_lid_diffusion, test_cavity_pressure_gradient,
               test_cavity_pressure_iterations, test_cavity_source_term, test_cavity_advection)


def _caught(pins):
    for pin in pins:
        try:
            pin()
        except AssertionError:
            return True
    return False


def test_planted_cavity_mistakes_are_caught(monkeypatch):
    """A halved cross term, swapped Poisson weights, a pn copied once instead
    of per iteration, and a flipped advection neighbour each fail a pin."""
    good_b, good_p, good_flow = S._cavity_b, S._cavity_pressure, S.cavity_flow

    def half_cross(b, rho, dt, u, v, dx, dy):
        good_b(b, rho, dt, u, v, dx, dy)
        cross = (u[2:, 1:-1] - u[:-2, 1:-1]) / (2 * dy) * (v[1:-1, 2:] - v[1:-1, :-2]) / (2 * dx)
        b[1:-1, 1:-1] += rho * cross

    def swapped_weights(nit, p, dx, dy, b):
        good_p(nit, p, dy, dx, b)

    def stale_pn(nit, p, dx, dy, b):
        pn = p.copy()
        den = 2 * (dx ** 2 + dy ** 2)
        for _ in range(nit):
            p[1:-1, 1:-1] = (((pn[1:-1, 2:] + pn[1:-1, :-2]) * dy ** 2
                              + (pn[2:, 1:-1] + pn[:-2, 1:-1]) * dx ** 2) / den
                             - dx ** 2 * dy ** 2 / den * b[1:-1, 1:-1])
            p[:, -1] = p[:, -2]
            p[0, :] = p[1, :]
            p[:, 0] = p[:, 1]
            p[-1, :] = 0

    def flipped_upwind(nx, ny, nt, nit, u, v, dt, dx, dy, p, rho, nu):
        # mirror the field along x, step it, mirror back: the upwind
        # neighbour becomes the east one
        uf, vf, pf = u[:, ::-1].copy(), v[:, ::-1].copy(), p[:, ::-1].copy()
        good_flow(nx, ny, nt, nit, uf, vf, dt, dx, dy, pf, rho, nu)
        u[...], v[...], p[...] = uf[:, ::-1], vf[:, ::-1], pf[:, ::-1]

    for name, bad in (("_cavity_b", half_cross), ("_cavity_pressure", swapped_weights),
                      ("_cavity_pressure", stale_pn), ("cavity_flow", flipped_upwind)):
        with monkeypatch.context() as m:
            m.setattr(S, name, bad)
            assert _caught(CAVITY_PINS), name
    assert all(pin() is None for pin in CAVITY_PINS)


# ------------------------------------------------------------------ channel flow
def channel_run(ny, nx, nit, u=None, v=None, p=None, F=1.0):
    """S.channel_flow (looked up per call) on the reference grid spacing;
    returns (u, v, p, steps, dt)."""
    dx, dy, dt = 2 / (nx - 1), 2 / (ny - 1), .1 / ((nx - 1) * (ny - 1))
    u = np.zeros((ny, nx)) if u is None else u.copy()
    v = np.zeros((ny, nx)) if v is None else v.copy()
    p = np.ones((ny, nx)) if p is None else p.copy()
    steps = S.channel_flow(nit, u, v, dt, dx, dy, p, 1.0, 0.1, F)
    return u, v, p, steps, dt


def test_channel_first_step_is_the_force():
    """With u = v = 0 and p = 1, step one has no source, no gradient and no
    advection: u = F dt between the walls, v = 0, p = 1.  The stopping test
    then sees udiff = 1, so a second step runs; a zero-F run stops at once
    (0/0 is NaN, which ends the loop)."""
    ny, nx, F = 5, 7, 2.5
    dx, dy, dt = 2 / (nx - 1), 2 / (ny - 1), .1 / ((nx - 1) * (ny - 1))
    u, v, p = np.zeros((ny, nx)), np.zeros((ny, nx)), np.ones((ny, nx))
    calls = []
    real_b = S._channel_b

    def spy(*a):
        calls.append(a[4].copy())
        return real_b(*a)
    S._channel_b = spy
    try:
        steps = S.channel_flow(3, u, v, dt, dx, dy, p, 1.0, 0.1, F)
    finally:
        S._channel_b = real_b
    assert steps >= 2
    assert np.all(calls[1][1:-1] == F * dt) and not calls[1][[0, -1]].any()
    with np.errstate(invalid="ignore"):
        u, v, p, steps, _ = channel_run(ny, nx, 3, F=0.0)
    assert steps == 1 and not u.any() and not v.any() and np.all(p == 1.0)


def test_channel_is_periodic_in_x():
    """Rolling the input along x rolls the output: every column, 0 and nx-1
    included, is computed with wrapped neighbours.  (Bit-exact per point;
    the step count is compared too.)"""
    ny, nx, nit = 5, 8, 2
    rng = np.random.default_rng(4)
    v0, p0 = 0.01 * rng.random((ny, nx)), 1 + 0.1 * rng.random((ny, nx))
    u, v, p, steps, _ = channel_run(ny, nx, nit, v=v0, p=p0)
    for k in (1, 3):
        ur, vr, pr, sr, _ = channel_run(ny, nx, nit, v=np.roll(v0, k, 1), p=np.roll(p0, k, 1))
        assert sr == steps
        assert np.allclose(ur, np.roll(u, k, 1), rtol=1e-12, atol=1e-15)
        assert np.allclose(vr, np.roll(v, k, 1), rtol=1e-12, atol=1e-15)
        assert np.allclose(pr, np.roll(p, k, 1), rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("ny,nx,nit", [(3, 4, 1), (4, 5, 2), (5, 3, 3), (6, 7, 4)])
def test_channel_bruteforce_bitwise(ny, nx, nit):
    rng = np.random.default_rng(ny * 41 + nx)
    v0, p0 = 0.01 * rng.random((ny, nx)), 1 + 0.1 * rng.random((ny, nx))
    u, v, p, steps, dt = channel_run(ny, nx, nit, v=v0, p=p0)
    dx, dy = 2 / (nx - 1), 2 / (ny - 1)
    ru, rv, rp, rs = scalar.channel_flow(nit, np.zeros((ny, nx)), v0, dt, dx, dy, p0, 1.0, 0.1, 1.0)
    assert steps == rs and steps > 100
    assert np.array_equal(u, ru) and np.array_equal(v, rv) and np.array_equal(p, rp)


def test_planted_channel_mistakes_are_caught(monkeypatch):
    """A dropped force term and a non-periodic (clamped) x-neighbour each
    fail a pin."""
    good = S.channel_flow

    def no_force(nit, u, v, dt, dx, dy, p, rho, nu, F):
        return good(nit, u, v, dt, dx, dy, p, rho, nu, 0.0 * F + 1e-300)

    def clamped(nit, u, v, dt, dx, dy, p, rho, nu, F):
        # pad with copies of the edge columns instead of wrapping: columns
        # 0 and nx-1 then see themselves as their outer neighbour
        ny, nx = u.shape
        big = [np.concatenate([a[:, :1], a, a[:, -1:]], axis=1) for a in (u, v, p)]
        steps = good(nit, *big[:2], dt, dx, dy, big[2], rho, nu, F)
        u[...], v[...], p[...] = big[0][:, 1:-1], big[1][:, 1:-1], big[2][:, 1:-1]
        return steps

    for bad in (no_force, clamped):
        with monkeypatch.context() as m:
            m.setattr(S, "channel_flow", bad)
            with pytest.raises(AssertionError):
                with np.errstate(all="ignore"):
                    test_channel_first_step_is_the_force()
                    test_channel_is_periodic_in_x()


# ------------------------------------------------------------------ seidel 2-D
def test_seidel2d_delta_by_hand():
    A = np.zeros((4, 4))
    A[1, 1] = 81.0
    S.seidel_2d_kernel(2, 4, A)
    want = np.zeros((4, 4))
    want[1, 1], want[1, 2] = 9.0, 1.0
    want[2, 1] = 10.0 / 9.0
    want[2, 2] = (10.0 + 10.0 / 9.0) / 9.0
    assert np.array_equal(A, want)


def test_seidel2d_constant_is_fixed_and_boundary_untouched():
    for c in (1.0, 0.5, 3.0):
        A = np.full((9, 9), c)
        S.seidel_2d_kernel(5, 9, A)
        assert np.all(A == c)
    rng = np.random.default_rng(2)
    A = rng.random((8, 8))
    B = A.copy()
    S.seidel_2d_kernel(4, 8, B)
    for edge in (np.s_[0, :], np.s_[-1, :], np.s_[:, 0], np.s_[:, -1]):
        assert np.array_equal(A[edge], B[edge])


@pytest.mark.parametrize("n,tsteps", [(3, 2), (4, 3), (5, 4), (9, 3), (12, 5)])
def test_seidel2d_bruteforce_bitwise(n, tsteps):
    A = np.random.default_rng(n * 7 + tsteps).random((n, n))
    a = A.copy()
    S.seidel_2d_kernel(tsteps, n, a)
    assert np.array_equal(a, scalar.seidel_2d(tsteps, n, A))


def test_planted_seidel2d_mistakes_are_caught(monkeypatch):
    """A Jacobi-style sweep (old values of the row above), a divisor of 8
    and a dropped neighbour each move the hand-worked delta response."""
    def make(mistake):
        def kernel(TSTEPS, N, A):
            for _ in range(TSTEPS - 1):
                old = A.copy()
                for i in range(1, N - 1):
                    up = old[i - 1] if mistake == "jacobi" else A[i - 1]
                    s = up[:-2]
                    if mistake != "drop":
                        s = s + up[1:-1]
                    s = s + up[2:] + A[i, 2:] + A[i + 1, :-2] + A[i + 1, 1:-1] + A[i + 1, 2:]
                    A[i, 1:-1] += s
                    for j in range(1, N - 1):
                        A[i, j] = (A[i, j] + A[i, j - 1]) / (8.0 if mistake == "eight" else 9.0)
        return kernel
    for mistake in ("jacobi", "eight", "drop"):
        monkeypatch.setattr(S, "seidel_2d_kernel", make(mistake))
        with pytest.raises(AssertionError):
            test_seidel2d_delta_by_hand()


# ------------------------------------------------------------------ golden
def _golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_golden_files_present():
    names = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))
    assert len(names) >= 10, names


@pytest.mark.parametrize("name", ["jacobi_2d_ref", "jacobi_2d_uniform"])
def test_golden_jacobi(name):
    g = _golden(name)
    a, b = g["A"].copy(), g["B"].copy()
    S.jacobi_2d_kernel(int(g["TSTEPS"]), a, b)
    assert np.array_equal(a, g["A_out"]) and np.array_equal(b, g["B_out"])


@pytest.mark.parametrize("name", ["heat_3d_ref", "heat_3d_uniform"])
def test_golden_heat(name):
    g = _golden(name)
    a, b = g["A"].copy(), g["B"].copy()
    S.heat_3d_kernel(int(g["TSTEPS"]), a, b)
    assert np.array_equal(a, g["A_out"]) and np.array_equal(b, g["B_out"])


@pytest.mark.parametrize("name", ["hdiff_ref", "hdiff_config", "hdiff_iterated3"])
def test_golden_hdiff(name):
    g = _golden(name)
    out = g["out_field"].copy()
    S.hdiff_iterated(g["in_field"], out, g["coeff"], int(g.get("nsteps", 1)))
    assert np.array_equal(out, g["out"])


@pytest.mark.parametrize("name", ["jacobi_1d_ref", "jacobi_1d_uniform"])
def test_golden_jacobi_1d(name):
    g = _golden(name)
    a, b = g["A"].copy(), g["B"].copy()
    S.jacobi_1d_kernel(int(g["TSTEPS"]), a, b)
    assert np.array_equal(a, g["A_out"]) and np.array_equal(b, g["B_out"])


@pytest.mark.parametrize("name", ["fdtd_2d_ref", "fdtd_2d_uniform"])
def test_golden_fdtd_2d(name):
    g = _golden(name)
    ex, ey, hz = g["ex"].copy(), g["ey"].copy(), g["hz"].copy()
    S.fdtd_2d_kernel(int(g["TMAX"]), ex, ey, hz, g["fict"])
    assert np.array_equal(ex, g["ex_out"]) and np.array_equal(ey, g["ey_out"])
    assert np.array_equal(hz, g["hz_out"])


@pytest.mark.parametrize("name", ["cavity_flow_ref", "cavity_flow_uniform"])
def test_golden_cavity_flow(name):
    g = _golden(name)
    u, v, p = g["u"].copy(), g["v"].copy(), g["p"].copy()
    S.cavity_flow(int(g["nx"]), int(g["ny"]), int(g["nt"]), int(g["nit"]), u, v, float(g["dt"]),
                  float(g["dx"]), float(g["dy"]), p, float(g["rho"]), float(g["nu"]))
    assert np.array_equal(u, g["u_out"]) and np.array_equal(v, g["v_out"])
    assert np.array_equal(p, g["p_out"])


@pytest.mark.parametrize("name", ["channel_flow_ref", "channel_flow_uniform"])
def test_golden_channel_flow(name):
    g = _golden(name)
    u, v, p = g["u"].copy(), g["v"].copy(), g["p"].copy()
    steps = S.channel_flow(int(g["nit"]), u, v, float(g["dt"]), float(g["dx"]), float(g["dy"]), p,
                           float(g["rho"]), float(g["nu"]), float(g["F"]))
    assert steps == int(g["stepcount"])
    assert np.array_equal(u, g["u_out"]) and np.array_equal(v, g["v_out"])
    assert np.array_equal(p, g["p_out"])


@pytest.mark.parametrize("name", ["seidel_2d_ref", "seidel_2d_uniform"])
def test_golden_seidel_2d(name):
    g = _golden(name)
    A = g["A"].copy()
    S.seidel_2d_kernel(int(g["TSTEPS"]), int(g["N"]), A)
    assert np.array_equal(A, g["out"])


@pytest.mark.parametrize("name", ["adi_ref", "adi_uniform"])
def test_golden_adi(name):
    g = _golden(name)
    u = g["u"].copy()
    S.adi_kernel(int(g["TSTEPS"]), int(g["N"]), u)
    assert np.array_equal(u, g["out"])


@pytest.mark.parametrize("name", ["vadv_ref", "vadv_config", "vadv_k2"])
def test_golden_vadv(name):
    g = _golden(name)
    out = g["utens_stage"].copy()
    S.vadv(out, g["u_stage"], g["wcon"], g["u_pos"], g["utens"], float(g["dtr_stage"]))
    assert np.array_equal(out, g["out"])


# ------------------------------------------------------------------ planted mistakes
def _mutant_jacobi_sweep(weight=0.2, drop=None, shift=False):
    def sweep(src, dst):
        c = src[1:-1, 1:-1]
        terms = [src[1:-1, :-2], src[1:-1, 2:], src[2:, 1:-1],
                 src[:-2, 1:-1] if not shift else src[2:, 1:-1]]
        acc = c
        for n, t in enumerate(terms):
            if n != drop:
                acc = acc + t
        dst[1:-1, 1:-1] = weight * acc
    return sweep


@pytest.mark.parametrize("mutant", [dict(weight=0.25), dict(drop=3), dict(shift=True)])
def test_planted_jacobi_mistakes_are_caught(monkeypatch, mutant):
    monkeypatch.setattr(S, "_jacobi_sweep", _mutant_jacobi_sweep(**mutant))
    caught = (not jacobi_affine_ok()) or jacobi_mode_err() > 1e-13
    assert caught


def _mutant_heat_sweep(kind):
    def sweep(src, dst):
        c = src[1:-1, 1:-1, 1:-1]
        c2 = 2.0 * c
        ti = 0.125 * ((src[2:, 1:-1, 1:-1] - c2) + src[:-2, 1:-1, 1:-1])
        tj = 0.125 * ((src[1:-1, 2:, 1:-1] - c2) + src[1:-1, :-2, 1:-1])
        km = src[1:-1, 1:-1, :-2] if kind != "k-index" else src[1:-1, 1:-1, 2:]
        tk = 0.125 * ((src[1:-1, 1:-1, 2:] - c2) + km)
        if kind == "drop-k":
            tk = 0.0 * tk
        if kind == "no-centre":
            dst[1:-1, 1:-1, 1:-1] = (ti + tj) + tk
        elif kind == "weight":
            dst[1:-1, 1:-1, 1:-1] = ((ti + tj) + 1.25 * tk) + c
        else:
            dst[1:-1, 1:-1, 1:-1] = ((ti + tj) + tk) + c
    return sweep


@pytest.mark.parametrize("kind", ["drop-k", "no-centre", "weight", "k-index"])
def test_planted_heat_mistakes_are_caught(monkeypatch, kind):
    monkeypatch.setattr(S, "_heat_sweep", _mutant_heat_sweep(kind))
    caught = heat_mode_err() > 1e-13 or not heat_cone_ok()
    assert caught


@pytest.mark.parametrize("limiter", [
    lambda r, g: r,                               # limiter dropped
    lambda r, g: np.where(r * g >= 0, 0, r),      # '>=' for '>'
    lambda r, g: np.where(r * g < 0, 0, r),       # sign of the test flipped
])
def test_planted_hdiff_limiter_mistakes_are_caught(monkeypatch, limiter):
    monkeypatch.setattr(S, "_limit", limiter)
    assert not hdiff_spike_ok()


def _mutant_hdiff(mistake):
    """A copy of the oracle's hdiff with one plausible slicing mistake."""
    def fn(in_field, out_field, coeff):
        I, J, K = out_field.shape
        f = in_field
        nb = f[2:I + 4, 1:J + 3] + f[0:I + 2, 1:J + 3]
        nb = nb + f[1:I + 3, 2:J + 4]
        nb = nb + f[1:I + 3, 0:J + 2]
        lap = 4.0 * f[1:I + 3, 1:J + 3] - nb
        r = lap[1:, 1:J + 1] - lap[:-1, 1:J + 1]
        g = f[2:I + 3, 2:J + 2] - f[1:I + 2, 2:J + 2]
        flx = S._limit(r, g)
        r = lap[1:I + 1, 1:] - lap[1:I + 1, :-1]
        g = f[2:I + 2, 2:J + 3] - f[2:I + 2, 1:J + 2]
        fly = S._limit(r, g)
        if mistake == "flx-order":
            div = flx[:-1] - flx[1:]
        else:
            div = flx[1:] - flx[:-1]
        div = div + fly[:, 1:]
        div = div - fly[:, :-1]
        centre = f[2:I + 2, 3:J + 3] if mistake == "centre-j" else f[2:I + 2, 2:J + 2]
        out_field[:, :, :] = centre - coeff * div
    return fn


@pytest.mark.parametrize("mistake", ["flx-order", "centre-j"])
def test_planted_hdiff_slice_mistakes_are_caught(mistake):
    fn = _mutant_hdiff(mistake)
    I, J, K = 6, 5, 3
    in_field = _grid_fn(I, J, K, lambda p, q, k: 3 * p - 2 * q + k)
    out = np.empty((I, J, K))
    fn(in_field, out, np.random.default_rng(1).random((I, J, K)))
    caught = (not hdiff_spike_ok(fn=fn)) or not np.array_equal(out, in_field[2:-2, 2:-2])
    assert caught


def test_planted_vadv_mistakes_are_caught(monkeypatch):
    # a wrong implicit weight breaks the exact solution
    monkeypatch.setattr(S, "BET_P", 0.6)
    assert vadv_exact_err() > 1e-6
    monkeypatch.setattr(S, "BET_P", 0.5)
    # an explicit correction with the wrong sign breaks it as well
    monkeypatch.setattr(S, "BET_M", -0.5)
    assert vadv_exact_err() > 1e-6


@pytest.mark.parametrize("mistake", ["swap_neighbours", "rhs_weight", "p_order"])
def test_planted_adi_mistakes_are_caught(monkeypatch, mistake):
    """Mutant column sweeps (neighbours u[j,i-1] and u[j,i+1] swapped; the
    centre weight 1+2d written 1+d; p's recurrence with b and a swapped) fail
    the banded-solve check."""
    def mutant(u, v, coef):
        a, b, c, d, f = coef
        n = u.shape[0]
        lo, mid = (b, a) if mistake == "p_order" else (a, b)
        p, den = S._adi_p(lo, mid, c, n - 2)
        v[0, 1:n - 1] = v[n - 1, 1:n - 1] = 1.0
        q = np.empty((n - 1, n - 2))
        q[0] = 1.0
        w = 1.0 + (d if mistake == "rhs_weight" else 2.0 * d)
        west, east = (u[:, 2:], u[:, :-2]) if mistake == "swap_neighbours" else (u[:, :-2], u[:, 2:])
        for j in range(1, n - 1):
            rhs = -d * west[j] + w * u[j, 1:n - 1] - f * east[j]
            q[j] = (rhs - a * q[j - 1]) / den[j]
        for j in range(n - 2, 0, -1):
            v[j, 1:n - 1] = p[j] * v[j + 1, 1:n - 1] + q[j]
    monkeypatch.setattr(S, "adi_column_sweep", mutant)
    with pytest.raises(AssertionError):
        test_adi_column_sweep_solves_its_systems(9)


# ------------------------------------------------------------------ C oracle
def test_c_seidel2d_delta_by_hand():
    """The C restatement gives the hand-worked delta response."""
    from oracle import c_oracle
    A = np.zeros((4, 4))
    A[1, 1] = 81.0
    c_oracle.seidel_2d_kernel(2, 4, A)
    want = np.zeros((4, 4))
    want[1, 1], want[1, 2] = 9.0, 1.0
    want[2, 1] = 10.0 / 9.0
    want[2, 2] = (10.0 + 10.0 / 9.0) / 9.0
    assert np.array_equal(A, want)


@pytest.mark.parametrize("name", ["seidel_2d_ref", "seidel_2d_uniform"])
def test_c_seidel2d_golden(name):
    from oracle import c_oracle
    g = _golden(name)
    A = g["A"].copy()
    c_oracle.seidel_2d_kernel(int(g["TSTEPS"]), int(g["N"]), A)
    assert np.array_equal(A, g["out"])


@pytest.mark.parametrize("n,tsteps", [(3, 1), (3, 2), (4, 3), (9, 5), (64, 4), (130, 3),
                                      (257, 2)])
def test_c_seidel2d_matches_numpy_oracle_bitwise(n, tsteps):
    from oracle import c_oracle
    A = np.random.default_rng(n + 1000 * tsteps).random((n, n))
    a, b = A.copy(), A.copy()
    S.seidel_2d_kernel(tsteps, n, a)
    c_oracle.seidel_2d_kernel(tsteps, n, b)
    assert np.array_equal(a, b)
