"""Pure-Python scalar-loop restatement of the four kernels (tiny shapes only).

Test infrastructure only (see ``oracle/__init__.py``).  This is a second,
structurally different statement of the reference arithmetic: it works point
by point on Python floats (IEEE binary64, round-to-nearest) with explicit
index arithmetic instead of NumPy slices, so an off-by-one slice or a swapped
operand in ``oracle.stencils`` shows up as a bit difference.  Each function
returns new arrays and leaves its inputs alone.

References: jacobi_2d_numpy.py:4-10, heat_3d_numpy.py:4-20,
hdiff_numpy.py:5-29, vadv_numpy.py:4-78, adi_numpy.py:6-52,
jacobi_1d_numpy.py:4-8, fdtd_2d_numpy.py:4-11, cavity_flow_numpy.py:13-83
(npbench/benchmarks/...).
"""

import numpy as np


def jacobi_sweeps(A, B, nsweeps):
    n = A.shape[0]
    a = A.tolist()
    b = B.tolist()
    for s in range(nsweeps):
        src, dst = (a, b) if s % 2 == 0 else (b, a)
        new = [row[:] for row in dst]
        for i in range(1, n - 1):
            for j in range(1, n - 1):
                t = src[i][j]
                t = t + src[i][j - 1]
                t = t + src[i][j + 1]
                t = t + src[i + 1][j]
                t = t + src[i - 1][j]
                new[i][j] = 0.2 * t
        if s % 2 == 0:
            b = new
        else:
            a = new
    return np.array(a), np.array(b)


def jacobi_2d_kernel(TSTEPS, A, B):
    return jacobi_sweeps(A, B, 2 * max(TSTEPS - 1, 0))


def heat_sweeps(A, B, nsweeps):
    n = A.shape[0]
    a = A.tolist()
    b = B.tolist()

    def axis(up, centre, down):
        return 0.125 * ((up - 2.0 * centre) + down)

    for s in range(nsweeps):
        src, dst = (a, b) if s % 2 == 0 else (b, a)
        new = [[row[:] for row in plane] for plane in dst]
        for i in range(1, n - 1):
            for j in range(1, n - 1):
                for k in range(1, n - 1):
                    c = src[i][j][k]
                    ti = axis(src[i + 1][j][k], c, src[i - 1][j][k])
                    tj = axis(src[i][j + 1][k], c, src[i][j - 1][k])
                    tk = axis(src[i][j][k + 1], c, src[i][j][k - 1])
                    new[i][j][k] = ((ti + tj) + tk) + c
        if s % 2 == 0:
            b = new
        else:
            a = new
    return np.array(a), np.array(b)


def heat_3d_kernel(TSTEPS, A, B):
    return heat_sweeps(A, B, 2 * max(TSTEPS - 1, 0))


def hdiff(in_field, coeff):
    """Return out (I, J, K).  Coordinates: output (x, y) sits at in[x+2, y+2]."""
    I, J, K = coeff.shape
    f = in_field.tolist()
    cf = coeff.tolist()

    def lap(p, q, k):
        # Laplacian centred on in-field point (p, q)
        nb = f[p + 1][q][k] + f[p - 1][q][k]
        nb = nb + f[p][q + 1][k]
        nb = nb + f[p][q - 1][k]
        return 4.0 * f[p][q][k] - nb

    def limit(r, grad):
        return 0.0 if r * grad > 0 else r

    def flux_i(p, q, k):
        # flux across the face between in-field points (p, q) and (p+1, q)
        r = lap(p + 1, q, k) - lap(p, q, k)
        return limit(r, f[p + 1][q][k] - f[p][q][k])

    def flux_j(p, q, k):
        # flux across the face between in-field points (p, q) and (p, q+1)
        r = lap(p, q + 1, k) - lap(p, q, k)
        return limit(r, f[p][q + 1][k] - f[p][q][k])

    out = np.empty((I, J, K))
    for x in range(I):
        for y in range(J):
            for k in range(K):
                p, q = x + 2, y + 2
                d = flux_i(p, q, k) - flux_i(p - 1, q, k)
                d = d + flux_j(p, q, k)
                d = d - flux_j(p, q - 1, k)
                out[x, y, k] = f[p][q][k] - cf[x][y][k] * d
    return out


def hdiff_iterated(in_field, coeff, nsteps):
    I, J, _ = coeff.shape
    work = np.array(in_field, dtype=np.float64, copy=True)
    out = None
    for _ in range(nsteps):
        out = hdiff(work, coeff)
        work[2:I + 2, 2:J + 2] = out
    return out


def _thomas(a, b, c, d):
    """Thomas algorithm with the reference's reciprocal-then-multiply pivots."""
    n = len(b)
    cp = [0.0] * n
    dp = [0.0] * n
    inv = 1.0 / b[0]
    cp[0] = c[0] * inv
    dp[0] = d[0] * inv
    for k in range(1, n):
        inv = 1.0 / (b[k] - cp[k - 1] * a[k])
        cp[k] = c[k] * inv
        dp[k] = (d[k] - dp[k - 1] * a[k]) * inv
    x = [0.0] * n
    x[n - 1] = dp[n - 1]
    for k in range(n - 2, -1, -1):
        x[k] = dp[k] - cp[k] * x[k + 1]
    return x


def vadv(utens_stage, u_stage, wcon, u_pos, utens, dtr_stage):
    """Return the new utens_stage (I, J, K)."""
    I, J, K = utens_stage.shape
    out = np.empty((I, J, K))
    for i in range(I):
        for j in range(J):
            u = u_stage[i, j].tolist()
            s = [wcon[i + 1, j, k] + wcon[i, j, k] for k in range(K)]
            a = [0.0] * K
            b = [0.0] * K
            c = [0.0] * K
            d = [0.0] * K
            for k in range(K):
                base = ((dtr_stage * u_pos[i, j, k] + utens[i, j, k]) +
                        utens_stage[i, j, k])
                has_lo = k > 0
                has_hi = k < K - 1
                gav = -0.25 * s[k] if has_lo else 0.0
                gcv = 0.25 * s[k + 1] if has_hi else 0.0
                a[k] = gav * 0.5
                c[k] = gcv * 0.5
                if has_lo and has_hi:
                    b[k] = dtr_stage - a[k] - c[k]
                    corr = (-(gav * 0.5) * (u[k - 1] - u[k]) -
                            (gcv * 0.5) * (u[k + 1] - u[k]))
                elif has_hi:
                    b[k] = dtr_stage - c[k]
                    corr = -(gcv * 0.5) * (u[k + 1] - u[k])
                else:
                    b[k] = dtr_stage - a[k]
                    corr = -(gav * 0.5) * (u[k - 1] - u[k])
                d[k] = base + corr
            x = _thomas(a, b, c, d)
            for k in range(K):
                out[i, j, k] = dtr_stage * (x[k] - u_pos[i, j, k])
    return out


def adi(TSTEPS, N, u):
    """adi_numpy.kernel point by point: each line's p, q and back substitution
    with its own scalar recurrence (p recomputed per line, as the reference
    does), Python floats throughout.  Returns the new u."""
    n = N
    u = [[float(x) for x in row] for row in np.asarray(u)]
    dx = dy = 1.0 / n
    dt = 1.0 / TSTEPS
    mul1 = 2.0 * dt / (dx * dx)
    mul2 = 1.0 * dt / (dy * dy)
    a = c = -mul1 / 2.0
    b = e = 1.0 + mul2
    d = f = -mul2 / 2.0
    v = [[0.0] * n for _ in range(n)]
    for _ in range(TSTEPS):
        for i in range(1, n - 1):  # column i of u -> column i of v
            v[0][i] = 1.0
            p, q = [0.0], [v[0][i]]
            for j in range(1, n - 1):
                den = a * p[j - 1] + b
                p.append(-c / den)
                rhs = -d * u[j][i - 1] + (1.0 + 2.0 * d) * u[j][i] - f * u[j][i + 1]
                q.append((rhs - a * q[j - 1]) / den)
            v[n - 1][i] = 1.0
            for j in range(n - 2, 0, -1):
                v[j][i] = p[j] * v[j + 1][i] + q[j]
        for i in range(1, n - 1):  # row i of v -> row i of u
            u[i][0] = 1.0
            p, q = [0.0], [u[i][0]]
            for j in range(1, n - 1):
                den = d * p[j - 1] + e
                p.append(-f / den)
                rhs = -a * v[i - 1][j] + (1.0 + 2.0 * a) * v[i][j] - c * v[i + 1][j]
                q.append((rhs - d * q[j - 1]) / den)
            u[i][n - 1] = 1.0
            for j in range(n - 2, 0, -1):
                u[i][j] = p[j] * u[i][j + 1] + q[j]
    return np.array(u)


def jacobi_1d(TSTEPS, A, B):
    """jacobi_1d_numpy.kernel point by point; returns new (A, B)."""
    a = [float(x) for x in np.asarray(A)]
    b = [float(x) for x in np.asarray(B)]
    n = len(a)
    for _ in range(1, TSTEPS):
        for src, dst in ((a, b), (b, a)):
            new = [src[i - 1] + src[i] + src[i + 1] for i in range(1, n - 1)]
            for i in range(1, n - 1):
                dst[i] = 0.33333 * new[i - 1]
    return np.array(a), np.array(b)


def fdtd_2d(TMAX, ex, ey, hz, fict):
    """fdtd_2d_numpy.kernel point by point; returns new (ex, ey, hz)."""
    ex = [[float(x) for x in row] for row in np.asarray(ex)]
    ey = [[float(x) for x in row] for row in np.asarray(ey)]
    hz = [[float(x) for x in row] for row in np.asarray(hz)]
    nx, ny = len(ex), len(ex[0])
    for t in range(TMAX):
        for j in range(ny):
            ey[0][j] = float(fict[t])
        for i in range(1, nx):
            for j in range(ny):
                ey[i][j] = ey[i][j] - 0.5 * (hz[i][j] - hz[i - 1][j])
        for i in range(nx):
            for j in range(1, ny):
                ex[i][j] = ex[i][j] - 0.5 * (hz[i][j] - hz[i][j - 1])
        for i in range(nx - 1):
            for j in range(ny - 1):
                hz[i][j] = hz[i][j] - 0.7 * (ex[i][j + 1] - ex[i][j] + ey[i + 1][j] - ey[i][j])
    return np.array(ex), np.array(ey), np.array(hz)


def cavity_flow(nx, ny, nt, nit, u, v, dt, dx, dy, p, rho, nu):
    """cavity_flow_numpy.py:13-83, one point at a time; every statement of
    the reference, boundary copies included, runs in its own order.  Returns
    new (u, v, p)."""
    U = u.tolist()
    V = v.tolist()
    P = p.tolist()
    B = [[0.0] * nx for _ in range(ny)]
    for _ in range(nt):
        un = [r[:] for r in U]
        vn = [r[:] for r in V]
        for i in range(1, ny - 1):  # build_up_b
            for j in range(1, nx - 1):
                dudx = (U[i][j + 1] - U[i][j - 1]) / (2 * dx)
                dvdy = (V[i + 1][j] - V[i - 1][j]) / (2 * dy)
                dudy = (U[i + 1][j] - U[i - 1][j]) / (2 * dy)
                cross = dudy * (V[i][j + 1] - V[i][j - 1]) / (2 * dx)
                t = (1 / dt) * (dudx + dvdy)
                t = t - dudx * dudx
                t = t - 2 * cross
                t = t - dvdy * dvdy
                B[i][j] = rho * t
        for _q in range(nit):  # pressure_poisson
            pn = [r[:] for r in P]
            d2x = dx ** 2
            d2y = dy ** 2
            den = 2 * (d2x + d2y)
            for i in range(1, ny - 1):
                for j in range(1, nx - 1):
                    num = (pn[i][j + 1] + pn[i][j - 1]) * d2y + (pn[i + 1][j] + pn[i - 1][j]) * d2x
                    P[i][j] = num / den - (d2x * d2y / den) * B[i][j]
            for i in range(ny):
                P[i][nx - 1] = P[i][nx - 2]
            for j in range(nx):
                P[0][j] = P[1][j]
            for i in range(ny):
                P[i][0] = P[i][1]
            for j in range(nx):
                P[ny - 1][j] = 0.0
        for i in range(1, ny - 1):
            for j in range(1, nx - 1):
                uc, vc = un[i][j], vn[i][j]
                x = uc - ((uc * dt) / dx) * (uc - un[i][j - 1])
                x = x - ((vc * dt) / dy) * (uc - un[i - 1][j])
                x = x - (dt / (2 * rho * dx)) * (P[i][j + 1] - P[i][j - 1])
                lap = (dt / dx ** 2) * ((un[i][j + 1] - 2 * uc) + un[i][j - 1])
                lap = lap + (dt / dy ** 2) * ((un[i + 1][j] - 2 * uc) + un[i - 1][j])
                U[i][j] = x + nu * lap
                y = vc - ((uc * dt) / dx) * (vc - vn[i][j - 1])
                y = y - ((vc * dt) / dy) * (vc - vn[i - 1][j])
                y = y - (dt / (2 * rho * dy)) * (P[i + 1][j] - P[i - 1][j])
                lap = (dt / dx ** 2) * ((vn[i][j + 1] - 2 * vc) + vn[i][j - 1])
                lap = lap + (dt / dy ** 2) * ((vn[i + 1][j] - 2 * vc) + vn[i - 1][j])
                V[i][j] = y + nu * lap
        for j in range(nx):
            U[0][j] = 0.0
        for i in range(ny):
            U[i][0] = 0.0
            U[i][nx - 1] = 0.0
        for j in range(nx):
            U[ny - 1][j] = 1.0
        for j in range(nx):
            V[0][j] = 0.0
            V[ny - 1][j] = 0.0
        for i in range(ny):
            V[i][0] = 0.0
            V[i][nx - 1] = 0.0
    return np.array(U), np.array(V), np.array(P)


def channel_flow(nit, u, v, dt, dx, dy, p, rho, nu, F):
    """channel_flow_numpy.py:13-152, one point at a time (x-neighbours wrap
    around); the stopping test uses np.sum as the reference does.  Returns
    new (u, v, p, stepcount)."""
    ny, nx = u.shape
    U = u.tolist()
    V = v.tolist()
    P = p.tolist()
    E = [(j + 1) % nx for j in range(nx)]
    W = [(j - 1) % nx for j in range(nx)]
    udiff, steps = 1, 0
    while udiff > .001:
        un = [r[:] for r in U]
        vn = [r[:] for r in V]
        B = [[0.0] * nx for _ in range(ny)]
        for i in range(1, ny - 1):
            for j in range(nx):
                e, w = E[j], W[j]
                dudx = (U[i][e] - U[i][w]) / (2 * dx)
                dvdy = (V[i + 1][j] - V[i - 1][j]) / (2 * dy)
                cross = (U[i + 1][j] - U[i - 1][j]) / (2 * dy) * (V[i][e] - V[i][w]) / (2 * dx)
                B[i][j] = rho * ((((1 / dt) * (dudx + dvdy) - dudx * dudx) - 2 * cross)
                                 - dvdy * dvdy)
        d2x, d2y = dx ** 2, dy ** 2
        den = 2 * (d2x + d2y)
        for _q in range(nit):
            pn = [r[:] for r in P]
            for i in range(1, ny - 1):
                for j in range(nx):
                    num = (pn[i][E[j]] + pn[i][W[j]]) * d2y + (pn[i + 1][j] + pn[i - 1][j]) * d2x
                    P[i][j] = num / den - (d2x * d2y / den) * B[i][j]
            P[ny - 1] = P[ny - 2][:]
            P[0] = P[1][:]
        for i in range(1, ny - 1):
            for j in range(nx):
                e, w = E[j], W[j]
                uc, vc = un[i][j], vn[i][j]
                x = uc - ((uc * dt) / dx) * (uc - un[i][w])
                x = x - ((vc * dt) / dy) * (uc - un[i - 1][j])
                x = x - (dt / (2 * rho * dx)) * (P[i][e] - P[i][w])
                lap = (dt / dx ** 2) * ((un[i][e] - 2 * uc) + un[i][w])
                lap = lap + (dt / dy ** 2) * ((un[i + 1][j] - 2 * uc) + un[i - 1][j])
                U[i][j] = (x + nu * lap) + F * dt
                y = vc - ((uc * dt) / dx) * (vc - vn[i][w])
                y = y - ((vc * dt) / dy) * (vc - vn[i - 1][j])
                y = y - (dt / (2 * rho * dy)) * (P[i + 1][j] - P[i - 1][j])
                lap = (dt / dx ** 2) * ((vn[i][e] - 2 * vc) + vn[i][w])
                lap = lap + (dt / dy ** 2) * ((vn[i + 1][j] - 2 * vc) + vn[i - 1][j])
                V[i][j] = y + nu * lap
        for j in range(nx):
            U[0][j] = U[ny - 1][j] = 0.0
            V[0][j] = V[ny - 1][j] = 0.0
        su, sun = np.sum(np.array(U)), np.sum(np.array(un))
        udiff = (su - sun) / su
        steps += 1
    return np.array(U), np.array(V), np.array(P), steps


def seidel_2d(TSTEPS, N, A):
    """seidel_2d_numpy.py:4-13, one point at a time: the 7-neighbour sum of a
    row is formed from the row's values before its own sweep (NumPy builds
    the whole right-hand side first).  Returns a new array."""
    a = A.tolist()
    for _ in range(TSTEPS - 1):
        for i in range(1, N - 1):
            up, mid, dn = a[i - 1], a[i][:], a[i + 1]
            for j in range(1, N - 1):
                s = up[j - 1] + up[j]
                s = s + up[j + 1]
                s = s + mid[j + 1]
                s = s + dn[j - 1]
                s = s + dn[j]
                s = s + dn[j + 1]
                a[i][j] = mid[j] + s
            for j in range(1, N - 1):
                a[i][j] = (a[i][j] + a[i][j - 1]) / 9.0
    return np.array(a)
