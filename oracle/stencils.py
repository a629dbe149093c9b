"""Vectorised NumPy FP64 restatement of the reference stencil kernels.

Test infrastructure only (see ``oracle/__init__.py``).  Each function keeps
the element-wise order of operations of the reference NumPy code it follows,
so that on identical FP64 inputs the results are bit-identical to it.

Reference files (paths relative to the npbench source tree):
  jacobi_2d_kernel  npbench/benchmarks/polybench/jacobi_2d/jacobi_2d_numpy.py:4-10
  heat_3d_kernel    npbench/benchmarks/polybench/heat_3d/heat_3d_numpy.py:4-20
  hdiff             npbench/benchmarks/weather_stencils/hdiff/hdiff_numpy.py:5-29
  vadv              npbench/benchmarks/weather_stencils/vadv/vadv_numpy.py:4-78
  adi_kernel        npbench/benchmarks/polybench/adi/adi_numpy.py:6-52
  jacobi_1d_kernel  npbench/benchmarks/polybench/jacobi_1d/jacobi_1d_numpy.py:4-8
  fdtd_2d_kernel    npbench/benchmarks/polybench/fdtd_2d/fdtd_2d_numpy.py:4-11
  cavity_flow       npbench/benchmarks/cavity_flow/cavity_flow_numpy.py:13-83
  channel_flow      npbench/benchmarks/channel_flow/channel_flow_numpy.py:13-152
  seidel_2d_kernel  npbench/benchmarks/polybench/seidel_2d/seidel_2d_numpy.py:4-13
"""

import numpy as np


def _f64(*arrays):
    for a in arrays:
        if not (isinstance(a, np.ndarray) and a.dtype == np.float64):
            raise TypeError("the oracle computes in float64 and needs float64 "
                            "ndarrays")


# vadv_numpy.py:4-5 -- implicit and explicit weights of the vertical scheme.
BET_M = 0.5
BET_P = 0.5


# --------------------------------------------------------------------------
# 2-D 5-point Jacobi
# --------------------------------------------------------------------------
def _jacobi_sweep(src, dst):
    """One sweep ``dst[interior] = 0.2 * (c + west + east + south + north)``.

    Follows jacobi_2d_numpy.py:7-8: the sum is accumulated left to right as
    ``((((c + W) + E) + S) + N)`` with W = [i, j-1], E = [i, j+1],
    S = [i+1, j], N = [i-1, j], then scaled by 0.2.  Boundary rows and
    columns of ``dst`` are never written.
    """
    c = src[1:-1, 1:-1]
    acc = c + src[1:-1, :-2]
    acc = acc + src[1:-1, 2:]
    acc = acc + src[2:, 1:-1]
    acc = acc + src[:-2, 1:-1]
    dst[1:-1, 1:-1] = 0.2 * acc


def jacobi_sweeps(A, B, nsweeps):
    """Run ``nsweeps`` alternating sweeps: odd sweeps A->B, even sweeps B->A."""
    _f64(A, B)
    for s in range(nsweeps):
        if s % 2 == 0:
            _jacobi_sweep(A, B)
        else:
            _jacobi_sweep(B, A)


def jacobi_2d_kernel(TSTEPS, A, B):
    """jacobi_2d_numpy.kernel: ``TSTEPS - 1`` ping-pong iterations in place."""
    _f64(A, B)
    for _ in range(1, TSTEPS):
        _jacobi_sweep(A, B)
        _jacobi_sweep(B, A)


# --------------------------------------------------------------------------
# 3-D 7-point heat
# --------------------------------------------------------------------------
def _heat_sweep(src, dst):
    """One explicit heat sweep on the interior of a cubic grid.

    Follows heat_3d_numpy.py:7-13: per axis ``0.125 * ((x[+1] - 2.0*c) + x[-1])``,
    the three axis terms summed i, j, k left to right, then the centre added.
    """
    c = src[1:-1, 1:-1, 1:-1]
    c2 = 2.0 * c
    ti = 0.125 * ((src[2:, 1:-1, 1:-1] - c2) + src[:-2, 1:-1, 1:-1])
    tj = 0.125 * ((src[1:-1, 2:, 1:-1] - c2) + src[1:-1, :-2, 1:-1])
    tk = 0.125 * ((src[1:-1, 1:-1, 2:] - c2) + src[1:-1, 1:-1, :-2])
    dst[1:-1, 1:-1, 1:-1] = ((ti + tj) + tk) + c


def heat_sweeps(A, B, nsweeps):
    """Run ``nsweeps`` alternating heat sweeps: odd A->B, even B->A."""
    _f64(A, B)
    for s in range(nsweeps):
        if s % 2 == 0:
            _heat_sweep(A, B)
        else:
            _heat_sweep(B, A)


def heat_3d_kernel(TSTEPS, A, B):
    """heat_3d_numpy.kernel: ``TSTEPS - 1`` ping-pong iterations in place."""
    _f64(A, B)
    for _ in range(1, TSTEPS):
        _heat_sweep(A, B)
        _heat_sweep(B, A)


# --------------------------------------------------------------------------
# Horizontal diffusion with flux limiter
# --------------------------------------------------------------------------
def _limit(r, grad):
    """Flux limiter of hdiff_numpy.py:12-17/20-25: zero where r*grad > 0."""
    return np.where(r * grad > 0, 0, r)


def hdiff(in_field, out_field, coeff):
    """hdiff_numpy.hdiff: one application, ``out_field`` written in place.

    Shapes: in_field (I+4, J+4, K), out_field and coeff (I, J, K).
    lap is defined on (I+2, J+2), flx on (I+1, J), fly on (I, J+1)
    (hdiff_numpy.py:7-25); the limiter zeroes a flux whose product with the
    local gradient is strictly positive.
    """
    _f64(in_field, out_field, coeff)
    I, J, K = out_field.shape
    f = in_field
    # hdiff_numpy.py:7-9 -- 4*c - (((i+1) + (i-1)) + (j+1)) + (j-1))
    nb = f[2:I + 4, 1:J + 3] + f[0:I + 2, 1:J + 3]
    nb = nb + f[1:I + 3, 2:J + 4]
    nb = nb + f[1:I + 3, 0:J + 2]
    lap = 4.0 * f[1:I + 3, 1:J + 3] - nb

    # hdiff_numpy.py:11-17
    r = lap[1:, 1:J + 1] - lap[:-1, 1:J + 1]
    g = f[2:I + 3, 2:J + 2] - f[1:I + 2, 2:J + 2]
    flx = _limit(r, g)

    # hdiff_numpy.py:19-25
    r = lap[1:I + 1, 1:] - lap[1:I + 1, :-1]
    g = f[2:I + 2, 2:J + 3] - f[2:I + 2, 1:J + 2]
    fly = _limit(r, g)

    # hdiff_numpy.py:27-29 -- ((dflx_i + fly[j+1]) - fly[j])
    div = flx[1:] - flx[:-1]
    div = div + fly[:, 1:]
    div = div - fly[:, :-1]
    out_field[:, :, :] = f[2:I + 2, 2:J + 2] - coeff * div


def hdiff_iterated(in_field, out_field, coeff, nsteps):
    """Build-defined iteration of hdiff (the reference has no time loop).

    Step n feeds the previous output back into the interior of a copy of
    ``in_field`` whose 2-cell halo stays fixed at its original values; the
    final output lands in ``out_field``.  ``nsteps == 1`` is exactly the
    reference hdiff.  ``in_field`` is not modified.
    """
    if nsteps < 1:
        raise ValueError("nsteps must be >= 1")
    I, J, _ = out_field.shape
    work = np.array(in_field, dtype=np.float64, copy=True)
    res = np.empty(out_field.shape, dtype=np.float64)
    for n in range(nsteps):
        hdiff(work, res, coeff)
        if n + 1 < nsteps:
            work[2:I + 2, 2:J + 2] = res
    out_field[:, :, :] = res


# --------------------------------------------------------------------------
# Implicit vertical advection (Thomas solve per (i, j) column)
# --------------------------------------------------------------------------
def vadv(utens_stage, u_stage, wcon, u_pos, utens, dtr_stage):
    """vadv_numpy.vadv: per-column tridiagonal solve, ``utens_stage`` in place.

    For column (i, j) and level k, with s_k = wcon[i+1,j,k] + wcon[i,j,k]:
      a_k = (-0.25 * s_k) * BET_P,  c_k = (0.25 * s_{k+1}) * BET_P,
      b_k = dtr - a_k - c_k (a_k absent at k = 0, c_k absent at k = K-1),
      d_k = dtr*u_pos + utens + utens_stage + corr_k, with the explicit
      BET_M correction ``-as*(u[k-1]-u[k]) - cs*(u[k+1]-u[k])``.
    Forward sweep takes the reciprocal of the pivot then multiplies
    (vadv_numpy.py:28-30, 51-53, 67-68); back substitution writes
    ``dtr * (x_k - u_pos_k)`` (vadv_numpy.py:70-78).  Requires K >= 2.
    """
    _f64(utens_stage, u_stage, wcon, u_pos, utens)
    I, J, K = utens_stage.shape
    if K < 2:
        raise ValueError("vadv needs K >= 2")
    ccol = np.empty((I, J, K), dtype=np.float64)
    dcol = np.empty((I, J, K), dtype=np.float64)

    def wsum(k):
        return wcon[1:, :, k] + wcon[:-1, :, k]

    def rhs(k, corr):
        return ((dtr_stage * u_pos[:, :, k] + utens[:, :, k]) +
                utens_stage[:, :, k]) + corr

    # k = 0 (vadv_numpy.py:15-30)
    gcv = 0.25 * wsum(1)
    cs = gcv * BET_M
    c0 = gcv * BET_P
    bcol = dtr_stage - c0
    corr = -cs * (u_stage[:, :, 1] - u_stage[:, :, 0])
    d0 = rhs(0, corr)
    inv = 1.0 / bcol
    ccol[:, :, 0] = c0 * inv
    dcol[:, :, 0] = d0 * inv

    # 0 < k < K-1 (vadv_numpy.py:32-53)
    for k in range(1, K - 1):
        gav = -0.25 * wsum(k)
        gcv = 0.25 * wsum(k + 1)
        as_ = gav * BET_M
        cs = gcv * BET_M
        acol = gav * BET_P
        ck = gcv * BET_P
        bcol = dtr_stage - acol - ck
        corr = (-as_ * (u_stage[:, :, k - 1] - u_stage[:, :, k]) -
                cs * (u_stage[:, :, k + 1] - u_stage[:, :, k]))
        dk = rhs(k, corr)
        inv = 1.0 / (bcol - ccol[:, :, k - 1] * acol)
        ccol[:, :, k] = ck * inv
        dcol[:, :, k] = (dk - dcol[:, :, k - 1] * acol) * inv

    # k = K-1 (vadv_numpy.py:55-68)
    k = K - 1
    gav = -0.25 * wsum(k)
    as_ = gav * BET_M
    acol = gav * BET_P
    bcol = dtr_stage - acol
    corr = -as_ * (u_stage[:, :, k - 1] - u_stage[:, :, k])
    dk = rhs(k, corr)
    inv = 1.0 / (bcol - ccol[:, :, k - 1] * acol)
    dcol[:, :, k] = (dk - dcol[:, :, k - 1] * acol) * inv

    # back substitution (vadv_numpy.py:70-78)
    x = dcol[:, :, K - 1].copy()
    utens_stage[:, :, K - 1] = dtr_stage * (x - u_pos[:, :, K - 1])
    for k in range(K - 2, -1, -1):
        x = dcol[:, :, k] - ccol[:, :, k] * x
        utens_stage[:, :, k] = dtr_stage * (x - u_pos[:, :, k])


# --------------------------------------------------------------------------
# Alternating-direction implicit solver (polybench ADI)
# --------------------------------------------------------------------------
def adi_coefficients(TSTEPS, N):
    """adi_numpy.py:11-25: (a, b, c, d, e, f) as the reference forms them.

    Note b = 1 + mul2 (not mul1), as the reference has it."""
    dx = 1.0 / N
    dy = 1.0 / N
    dt = 1.0 / TSTEPS
    mul1 = 2.0 * dt / (dx * dx)
    mul2 = 1.0 * dt / (dy * dy)
    a = -mul1 / 2.0
    d = -mul2 / 2.0
    return a, 1.0 + mul2, a, d, 1.0 + mul2, d


def _adi_p(lo, mid, hi, steps):
    """p[0] = 0, p[j] = -hi / (lo * p[j-1] + mid) and the denominators, as
    Python floats: the reference computes the same values once per line
    (adi_numpy.py:28,30 and :38,40); every line's p is the same."""
    p, den = [0.0], [None]
    for _ in range(steps):
        den.append(lo * p[-1] + mid)
        p.append(-hi / den[-1])
    return p, den


def adi_column_sweep(u, v, coef):
    """adi_numpy.py:27-34 with coef = (a, b, c, d, f): for every interior
    column i, solve along the rows into v[:, i] (v's rows 0 and N-1 set to 1);
    right-hand side ((-d u[j,i-1] + (1+2d) u[j,i]) - f u[j,i+1])."""
    a, b, c, d, f = coef
    n = u.shape[0]
    p, den = _adi_p(a, b, c, n - 2)
    v[0, 1:n - 1] = 1.0
    v[n - 1, 1:n - 1] = 1.0
    q = np.empty((n - 1, n - 2))
    q[0] = v[0, 1:n - 1]
    for j in range(1, n - 1):
        rhs = -d * u[j, 0:n - 2] + (1.0 + 2.0 * d) * u[j, 1:n - 1] - f * u[j, 2:n]
        q[j] = (rhs - a * q[j - 1]) / den[j]
    for j in range(n - 2, 0, -1):
        v[j, 1:n - 1] = p[j] * v[j + 1, 1:n - 1] + q[j]


def adi_row_sweep(v, u, coef):
    """adi_numpy.py:36-46 with coef = (a, c, d, e, f): for every interior row
    i, solve along the columns into u[i, :] (u's columns 0 and N-1 set to 1);
    right-hand side ((-a v[i-1,j] + (1+2a) v[i,j]) - c v[i+1,j])."""
    a, c, d, e, f = coef
    n = u.shape[0]
    p, den = _adi_p(d, e, f, n - 2)
    u[1:n - 1, 0] = 1.0
    u[1:n - 1, n - 1] = 1.0
    q = np.empty((n - 1, n - 2))
    q[0] = u[1:n - 1, 0]
    for j in range(1, n - 1):
        rhs = -a * v[0:n - 2, j] + (1.0 + 2.0 * a) * v[1:n - 1, j] - c * v[2:n, j]
        q[j] = (rhs - d * q[j - 1]) / den[j]
    for j in range(n - 2, 0, -1):
        u[1:n - 1, j] = p[j] * u[1:n - 1, j + 1] + q[j]


def adi_kernel(TSTEPS, N, u):
    """adi_numpy.kernel(TSTEPS, N, u): TSTEPS column and row sweeps, u updated
    in place and returned."""
    _f64(u)
    if u.shape != (N, N):
        raise ValueError("adi: u must be (N, N)")
    a, b, c, d, e, f = adi_coefficients(TSTEPS, N)
    v = np.empty_like(u)
    for _ in range(1, TSTEPS + 1):
        adi_column_sweep(u, v, (a, b, c, d, f))
        adi_row_sweep(v, u, (a, c, d, e, f))
    return u


# --------------------------------------------------------------------------
# 1-D three-point Jacobi
# --------------------------------------------------------------------------
def _jacobi1_sweep(src, dst):
    """dst[1:-1] = 0.33333 * ((left + centre) + right) (jacobi_1d_numpy.py:7-8);
    dst's two ends are never written."""
    acc = src[:-2] + src[1:-1]
    acc = acc + src[2:]
    dst[1:-1] = 0.33333 * acc


def jacobi_1d_kernel(TSTEPS, A, B):
    """jacobi_1d_numpy.kernel: TSTEPS - 1 ping-pong iterations in place."""
    _f64(A, B)
    for _ in range(1, TSTEPS):
        _jacobi1_sweep(A, B)
        _jacobi1_sweep(B, A)


# --------------------------------------------------------------------------
# 2-D finite-difference time domain
# --------------------------------------------------------------------------
def fdtd_2d_kernel(TMAX, ex, ey, hz, _fict_):
    """fdtd_2d_numpy.kernel: TMAX time steps, ex, ey, hz updated in place.

    Per step (fdtd_2d_numpy.py:7-11): ey's first row takes fict[t]; ey (rows
    1..) and ex (columns 1..) lose 0.5 times hz's backward difference along
    their axis; then hz (all but the last row and column) loses 0.7 times
    ((ex[i, j+1] - ex[i, j]) + ey[i+1, j]) - ey[i, j], with the new e-fields.
    """
    _f64(ex, ey, hz, _fict_)
    for t in range(TMAX):
        ey[0, :] = _fict_[t]
        ey[1:, :] = ey[1:, :] - 0.5 * (hz[1:, :] - hz[:-1, :])
        ex[:, 1:] = ex[:, 1:] - 0.5 * (hz[:, 1:] - hz[:, :-1])
        curl = ex[:-1, 1:] - ex[:-1, :-1]
        curl = curl + ey[1:, :-1]
        curl = curl - ey[:-1, :-1]
        hz[:-1, :-1] = hz[:-1, :-1] - 0.7 * curl


# --------------------------------------------------------------------------
# Lid-driven cavity flow
# --------------------------------------------------------------------------
def _cavity_b(b, rho, dt, u, v, dx, dy):
    """build_up_b (cavity_flow_numpy.py:13-20): the pressure equation's source
    from the velocity field's central differences, interior only."""
    dudx = (u[1:-1, 2:] - u[1:-1, :-2]) / (2 * dx)
    dvdy = (v[2:, 1:-1] - v[:-2, 1:-1]) / (2 * dy)
    cross = (u[2:, 1:-1] - u[:-2, 1:-1]) / (2 * dy) * (v[1:-1, 2:] - v[1:-1, :-2]) / (2 * dx)
    b[1:-1, 1:-1] = rho * (1 / dt * (dudx + dvdy) - dudx ** 2 - 2 * cross - dvdy ** 2)


def _cavity_pressure(nit, p, dx, dy, b):
    """pressure_poisson (cavity_flow_numpy.py:23-33): nit Jacobi iterations,
    each followed by the four boundary statements in the reference's order."""
    den = 2 * (dx ** 2 + dy ** 2)
    for _ in range(nit):
        pn = p.copy()
        avg = ((pn[1:-1, 2:] + pn[1:-1, :-2]) * dy ** 2 + (pn[2:, 1:-1] + pn[:-2, 1:-1]) * dx ** 2) / den
        p[1:-1, 1:-1] = avg - dx ** 2 * dy ** 2 / den * b[1:-1, 1:-1]
        p[:, -1] = p[:, -2]
        p[0, :] = p[1, :]
        p[:, 0] = p[:, 1]
        p[-1, :] = 0


def cavity_flow(nx, ny, nt, nit, u, v, dt, dx, dy, p, rho, nu):
    """cavity_flow_numpy.cavity_flow: nt time steps, u, v, p in place."""
    _f64(u, v, p)
    b = np.zeros((ny, nx))
    for _ in range(nt):
        un = u.copy()
        vn = v.copy()
        _cavity_b(b, rho, dt, u, v, dx, dy)
        _cavity_pressure(nit, p, dx, dy, b)
        c, w, e, s_, n_ = (slice(1, -1), slice(0, -2), slice(2, None), slice(0, -2), slice(2, None))
        uc, vc = un[c, c], vn[c, c]
        u[c, c] = (uc - uc * dt / dx * (uc - un[c, w]) - vc * dt / dy * (uc - un[s_, c])
                   - dt / (2 * rho * dx) * (p[c, e] - p[c, w])
                   + nu * (dt / dx ** 2 * (un[c, e] - 2 * uc + un[c, w])
                           + dt / dy ** 2 * (un[n_, c] - 2 * uc + un[s_, c])))
        v[c, c] = (vc - uc * dt / dx * (vc - vn[c, w]) - vc * dt / dy * (vc - vn[s_, c])
                   - dt / (2 * rho * dy) * (p[n_, c] - p[s_, c])
                   + nu * (dt / dx ** 2 * (vn[c, e] - 2 * vc + vn[c, w])
                           + dt / dy ** 2 * (vn[n_, c] - 2 * vc + vn[s_, c])))
        u[0, :] = 0
        u[:, 0] = 0
        u[:, -1] = 0
        u[-1, :] = 1
        v[0, :] = 0
        v[-1, :] = 0
        v[:, 0] = 0
        v[:, -1] = 0


# --------------------------------------------------------------------------
# Channel flow (periodic in x, walls at rows 0 and ny-1)
# --------------------------------------------------------------------------
def _channel_b(rho, dt, dx, dy, u, v):
    """build_up_b (channel_flow_numpy.py:13-39): the interior as in the
    cavity, plus the periodic columns 0 and nx-1, whose x-neighbours wrap."""
    ny, nx = u.shape
    b = np.zeros_like(u)
    for j in range(nx):
        e, w = (j + 1) % nx, (j - 1) % nx
        dudx = (u[1:-1, e] - u[1:-1, w]) / (2 * dx)
        dvdy = (v[2:, j] - v[:-2, j]) / (2 * dy)
        cross = (u[2:, j] - u[:-2, j]) / (2 * dy) * (v[1:-1, e] - v[1:-1, w]) / (2 * dx)
        b[1:-1, j] = rho * (1 / dt * (dudx + dvdy) - dudx ** 2 - 2 * cross - dvdy ** 2)
    return b


def _channel_pressure(nit, p, dx, dy, b):
    """pressure_poisson_periodic (channel_flow_numpy.py:42-66)."""
    ny, nx = p.shape
    den = 2 * (dx ** 2 + dy ** 2)
    e = np.r_[1:nx, 0]
    w = np.r_[nx - 1, 0:nx - 1]
    for _ in range(nit):
        pn = p.copy()
        p[1:-1, :] = (((pn[1:-1, e] + pn[1:-1, w]) * dy ** 2 + (pn[2:, :] + pn[:-2, :]) * dx ** 2)
                      / den - dx ** 2 * dy ** 2 / den * b[1:-1, :])
        p[-1, :] = p[-2, :]
        p[0, :] = p[1, :]


def channel_flow(nit, u, v, dt, dx, dy, p, rho, nu, F):
    """channel_flow_numpy.channel_flow: time steps until the relative change
    of sum(u) is at most 0.001; u, v, p in place; returns the step count."""
    _f64(u, v, p)
    ny, nx = u.shape
    e = np.r_[1:nx, 0]
    w = np.r_[nx - 1, 0:nx - 1]
    udiff, stepcount = 1, 0
    while udiff > .001:
        un = u.copy()
        vn = v.copy()
        b = _channel_b(rho, dt, dx, dy, u, v)
        _channel_pressure(nit, p, dx, dy, b)
        uc, vc = un[1:-1, :], vn[1:-1, :]
        u[1:-1, :] = (uc - uc * dt / dx * (uc - un[1:-1, w]) - vc * dt / dy * (uc - un[:-2, :])
                      - dt / (2 * rho * dx) * (p[1:-1, e] - p[1:-1, w])
                      + nu * (dt / dx ** 2 * (un[1:-1, e] - 2 * uc + un[1:-1, w])
                              + dt / dy ** 2 * (un[2:, :] - 2 * uc + un[:-2, :])) + F * dt)
        v[1:-1, :] = (vc - uc * dt / dx * (vc - vn[1:-1, w]) - vc * dt / dy * (vc - vn[:-2, :])
                      - dt / (2 * rho * dy) * (p[2:, :] - p[:-2, :])
                      + nu * (dt / dx ** 2 * (vn[1:-1, e] - 2 * vc + vn[1:-1, w])
                              + dt / dy ** 2 * (vn[2:, :] - 2 * vc + vn[:-2, :])))
        u[0, :] = 0
        u[-1, :] = 0
        v[0, :] = 0
        v[-1, :] = 0
        udiff = (np.sum(u) - np.sum(un)) / np.sum(u)
        stepcount += 1
    return stepcount


# --------------------------------------------------------------------------
# Gauss-Seidel 2-D (polybench seidel_2d)
# --------------------------------------------------------------------------
def seidel_2d_kernel(TSTEPS, N, A):
    """seidel_2d_numpy.kernel: TSTEPS - 1 sweeps in place; each row adds the
    sum of its 7 neighbours (the row above already swept, the row below not)
    and then runs A[i, j] = (A[i, j] + A[i, j-1]) / 9 along j."""
    _f64(A)
    for _ in range(0, TSTEPS - 1):
        for i in range(1, N - 1):
            A[i, 1:-1] += (A[i - 1, :-2] + A[i - 1, 1:-1] + A[i - 1, 2:] + A[i, 2:]
                           + A[i + 1, :-2] + A[i + 1, 1:-1] + A[i + 1, 2:])
            row = A[i]
            for j in range(1, N - 1):
                row[j] = (row[j] + row[j - 1]) / 9.0
