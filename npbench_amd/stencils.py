"""Host API of the MI355X stencil kernels.

Two layers:

* resident operations (``jacobi2d_sweeps``, ``heat3d_sweeps``, ``hdiff``,
  ``vadv``, ``adi``, ``jacobi1d``, ``fdtd2d``, ``cavity_flow``) take device
  tensors already in their storage dtype (float64,
  float32 or bfloat16) and enqueue HIP kernels on a stream without waiting;
* drop-in operations (``jacobi_2d``, ``heat_3d``, ``hdiff_dropin``,
  ``vadv_dropin``, ``adi_dropin``, ``jacobi_1d``, ``fdtd_2d``,
  ``cavity_flow_dropin``) take the reference's float64 arrays -- device tensors
  produced by the framework's copy function, or host NumPy arrays, which are
  copied to the device, processed and copied back in place -- and a
  ``precision`` that picks the storage form.  FP32/BF16 forms convert to the
  storage type on the device, run, and convert back to float64.

Every call goes to libnpbx.so; there is no CPU path.
"""

import ctypes

import numpy as np
import torch

from . import _native as nat
from . import device as dev

F64, F32, BF16 = nat.F64, nat.F32, nat.BF16


# --------------------------------------------------------------------------
# resident operations
# --------------------------------------------------------------------------
def _pair(name, A, B, ndim):
    dev.check_device_array(name + ".A", A)
    dev.check_device_array(name + ".B", B, shape=A.shape, dtype=A.dtype)
    if A.dim() != ndim or len(set(A.shape)) != 1:
        raise ValueError(f"{name}: expected a cubic {ndim}-D grid, got {tuple(A.shape)}")
    if A.data_ptr() == B.data_ptr():
        raise ValueError(f"{name}: A and B must be distinct arrays")
    return dev.dtype_code(A), A.shape[0]


def jacobi2d_sweeps(A, B, nsweeps, first=0, tblock=1, stream=None):
    """Runs sweeps ``first .. first+nsweeps-1`` (even: A->B, odd: B->A)."""
    code, n = _pair("jacobi2d", A, B, 2)
    nat.call("npbx_jacobi2d_sweeps", code, dev.ptr(A), dev.ptr(B), n, int(first), int(nsweeps),
             int(tblock), dev.stream_handle(stream))


def heat3d_sweeps(A, B, nsweeps, first=0, tblock=1, stream=None):
    code, n = _pair("heat3d", A, B, 3)
    nat.call("npbx_heat3d_sweeps", code, dev.ptr(A), dev.ptr(B), n, int(first), int(nsweeps),
             int(tblock), dev.stream_handle(stream))


def hdiff(in_field, out_field, coeff, nsteps=1, stream=None):
    dev.check_device_array("out_field", out_field)
    if out_field.dim() != 3:
        raise ValueError("hdiff: out_field must be 3-D (I, J, K)")
    I, J, K = out_field.shape
    dev.check_device_array("in_field", in_field, shape=(I + 4, J + 4, K), dtype=out_field.dtype)
    dev.check_device_array("coeff", coeff, shape=(I, J, K), dtype=out_field.dtype)
    nat.call("npbx_hdiff", dev.dtype_code(out_field), dev.ptr(in_field), dev.ptr(out_field),
             dev.ptr(coeff), I, J, K, int(nsteps), dev.stream_handle(stream))


def vadv(utens_stage, u_stage, wcon, u_pos, utens, dtr_stage, stream=None):
    dev.check_device_array("utens_stage", utens_stage)
    if utens_stage.dim() != 3:
        raise ValueError("vadv: fields must be 3-D (I, J, K)")
    I, J, K = utens_stage.shape
    dt = utens_stage.dtype
    for name, t in (("u_stage", u_stage), ("u_pos", u_pos), ("utens", utens)):
        dev.check_device_array(name, t, shape=(I, J, K), dtype=dt)
    dev.check_device_array("wcon", wcon, shape=(I + 1, J, K), dtype=dt)
    nat.call("npbx_vadv", dev.dtype_code(utens_stage), dev.ptr(utens_stage), dev.ptr(u_stage),
             dev.ptr(wcon), dev.ptr(u_pos), dev.ptr(utens), float(dtr_stage), I, J, K,
             dev.stream_handle(stream))


def adi(u, tsteps, stream=None):
    """``tsteps`` ADI time steps (adi_numpy.py:6-52) of a float64 (N, N)
    device tensor, in place."""
    dev.check_device_array("u", u, dtype=torch.float64)
    if u.dim() != 2 or u.shape[0] != u.shape[1]:
        raise ValueError("adi: u must be a square (N, N) array")
    nat.call("npbx_adi", dev.dtype_code(u), dev.ptr(u), u.shape[0], int(tsteps),
             dev.stream_handle(stream))


def jacobi1d(A, B, tsteps, stream=None):
    """jacobi_1d_numpy.kernel's TSTEPS - 1 iterations on float64 (N,) device
    tensors, in place."""
    dev.check_device_array("A", A, dtype=torch.float64)
    dev.check_device_array("B", B, shape=tuple(A.shape), dtype=torch.float64)
    if A.dim() != 1:
        raise ValueError("jacobi1d: A and B must be 1-D")
    nat.call("npbx_jacobi1d", dev.dtype_code(A), dev.ptr(A), dev.ptr(B), A.shape[0], int(tsteps),
             dev.stream_handle(stream))


def fdtd2d(ex, ey, hz, fict, tmax, stream=None):
    """fdtd_2d_numpy.kernel's TMAX time steps on float64 (NX, NY) device
    tensors ex, ey, hz (in place) with the float64 source fict."""
    dev.check_device_array("ex", ex, dtype=torch.float64)
    if ex.dim() != 2:
        raise ValueError("fdtd2d: fields must be 2-D (NX, NY)")
    for name, t in (("ey", ey), ("hz", hz)):
        dev.check_device_array(name, t, shape=tuple(ex.shape), dtype=torch.float64)
    dev.check_device_array("fict", fict, dtype=torch.float64)
    if fict.dim() != 1 or fict.shape[0] < int(tmax):
        raise ValueError(f"fdtd2d: fict needs at least TMAX={int(tmax)} entries")
    nat.call("npbx_fdtd2d", dev.dtype_code(ex), dev.ptr(ex), dev.ptr(ey), dev.ptr(hz),
             dev.ptr(fict), ex.shape[0], ex.shape[1], int(tmax), dev.stream_handle(stream))


def cavity_flow(nx, ny, nt, nit, u, v, dt, dx, dy, p, rho, nu, stream=None):
    """cavity_flow_numpy.cavity_flow on float64 (ny, nx) device tensors u, v,
    p (in place); the reference's argument order."""
    for name, t in (("u", u), ("v", v), ("p", p)):
        dev.check_device_array(name, t, shape=(int(ny), int(nx)), dtype=torch.float64)
    nat.call("npbx_cavity_flow", dev.dtype_code(u), int(nx), int(ny), int(nt), int(nit),
             dev.ptr(u), dev.ptr(v), float(dt), float(dx), float(dy), dev.ptr(p), float(rho),
             float(nu), dev.stream_handle(stream))


def channel_flow(nit, u, v, dt, dx, dy, p, rho, nu, F, stream=None, max_steps=None):
    """channel_flow_numpy.channel_flow on float64 (ny, nx) device tensors u,
    v, p (in place); returns the step count.  Runs launches of at most
    max_steps steps (by default about 50 ms of work each), synchronising
    after each, until the reference's criterion stops one; control returns
    to Python between launches, so a harness timeout (an interrupt of the
    main thread) stops an input that never converges, as it stops the
    reference's unbounded loop."""
    ny, nx = (int(d) for d in u.shape)
    for name, t in (("u", u), ("v", v), ("p", p)):
        dev.check_device_array(name, t, shape=(ny, nx), dtype=torch.float64)
    if max_steps is None:
        max_steps = max(1, min(4096, 150_000_000 // (nx * ny * (int(nit) + 3))))
    steps = ctypes.c_int64(0)
    done = ctypes.c_int32(0)
    total = 0
    while not done.value:
        nat.call("npbx_channel_flow_launch", dev.dtype_code(u), int(nit), dev.ptr(u), dev.ptr(v),
                 float(dt), float(dx), float(dy), dev.ptr(p), float(rho), float(nu), float(F),
                 nx, ny, int(max_steps), ctypes.byref(steps), ctypes.byref(done),
                 dev.stream_handle(stream))
        total += steps.value
    return total


def seidel2d(A, TSTEPS, stream=None):
    """seidel_2d_numpy.kernel's TSTEPS - 1 sweeps on a float64 (N, N) device
    tensor, in place."""
    dev.check_device_array("A", A, dtype=torch.float64)
    if A.dim() != 2 or A.shape[0] != A.shape[1]:
        raise ValueError(f"A must be square (N, N), got {tuple(A.shape)}")
    nat.call("npbx_seidel2d", dev.dtype_code(A), dev.ptr(A), int(A.shape[0]), int(TSTEPS),
             dev.stream_handle(stream))


def cast_into(src, dst, stream=None):
    """dst[...] = src converted to dst's dtype (RNE), on the device."""
    dev.check_device_array("src", src)
    dev.check_device_array("dst", dst, shape=src.shape)
    nat.call("npbx_cast", dev.dtype_code(src), dev.ptr(src), dev.dtype_code(dst), dev.ptr(dst),
             src.numel(), dev.stream_handle(stream))


def cast(src, dtype, stream=None):
    out = torch.empty(src.shape, dtype=dtype, device=src.device)
    cast_into(src, out, stream)
    return out


# --------------------------------------------------------------------------
# drop-in operations (reference signatures, float64 in and out)
# --------------------------------------------------------------------------
class _Staged:
    """Maps the reference's float64 operands to device storage tensors.

    Host NumPy arrays are copied to the device and written back in place on
    exit; float64 device tensors are used directly (FP64) or through
    FP32/BF16 storage copies that are converted back on exit.
    """

    def __init__(self, precision, arrays, outputs):
        if precision not in dev.PRECISION_CODE:
            raise ValueError(f"unknown precision {precision!r}; use fp64, fp32 or bf16")
        self.code = dev.PRECISION_CODE[precision]
        self.arrays = arrays
        self.outputs = outputs

    def __enter__(self):
        device = dev.require_device()
        self.host = []
        self.f64 = []
        for a in self.arrays:
            if isinstance(a, np.ndarray):
                if a.dtype != np.float64:
                    raise TypeError("drop-in kernels take float64 arrays")
                self.host.append(a)
                self.f64.append(torch.from_numpy(np.ascontiguousarray(a)).to(device))
            else:
                dev.check_device_array("array", a, dtype=torch.float64)
                self.host.append(None)
                self.f64.append(a)
        if self.code == F64:
            self.work = list(self.f64)
        else:
            self.work = [cast(t, dev.TORCH_DTYPE[self.code]) for t in self.f64]
        return self.work

    def __exit__(self, exc_type, exc, tb):
        if exc_type is not None:
            return False
        for idx in self.outputs:
            if self.code != F64:
                cast_into(self.work[idx], self.f64[idx])
            if self.host[idx] is not None:
                self.host[idx][...] = self.f64[idx].cpu().numpy()
        return False


def jacobi_2d(TSTEPS, A, B, precision="fp64", tblock=1):
    """jacobi_2d_numpy.kernel(TSTEPS, A, B) on the GPU; A and B updated in place."""
    with _Staged(precision, [A, B], outputs=(0, 1)) as (a, b):
        nsweeps = 2 * max(int(TSTEPS) - 1, 0)
        jacobi2d_sweeps(a, b, nsweeps, tblock=tblock)


def heat_3d(TSTEPS, A, B, precision="fp64", tblock=1):
    """heat_3d_numpy.kernel(TSTEPS, A, B) on the GPU; A and B updated in place."""
    with _Staged(precision, [A, B], outputs=(0, 1)) as (a, b):
        nsweeps = 2 * max(int(TSTEPS) - 1, 0)
        heat3d_sweeps(a, b, nsweeps, tblock=tblock)


def hdiff_dropin(in_field, out_field, coeff, precision="fp64", nsteps=1):
    """hdiff_numpy.hdiff(in_field, out_field, coeff); out_field written in place."""
    with _Staged(precision, [in_field, out_field, coeff], outputs=(1,)) as (i, o, c):
        hdiff(i, o, c, nsteps=nsteps)


def vadv_dropin(utens_stage, u_stage, wcon, u_pos, utens, dtr_stage, precision="fp64"):
    """vadv_numpy.vadv(...); utens_stage updated in place (FP64 elimination)."""
    if precision == "bf16":
        raise ValueError("vadv supports fp64 and fp32 field storage")
    with _Staged(precision, [utens_stage, u_stage, wcon, u_pos, utens], outputs=(0,)) as w:
        vadv(w[0], w[1], w[2], w[3], w[4], dtr_stage)


def adi_dropin(TSTEPS, N, u):
    """adi_numpy.kernel(TSTEPS, N, u): u updated in place and returned."""
    if tuple(u.shape) != (N, N):
        raise ValueError(f"adi: u has shape {tuple(u.shape)}, expected ({N}, {N})")
    with _Staged("fp64", [u], outputs=(0,)) as (w,):
        adi(w, TSTEPS)
    return u


def jacobi_1d(TSTEPS, A, B):
    """jacobi_1d_numpy.kernel(TSTEPS, A, B): A and B updated in place."""
    with _Staged("fp64", [A, B], outputs=(0, 1)) as (a, b):
        jacobi1d(a, b, TSTEPS)


def fdtd_2d(TMAX, ex, ey, hz, _fict_):
    """fdtd_2d_numpy.kernel(TMAX, ex, ey, hz, _fict_): the fields updated in
    place."""
    with _Staged("fp64", [ex, ey, hz, _fict_], outputs=(0, 1, 2)) as (a, b, c, f):
        fdtd2d(a, b, c, f, TMAX)


def seidel_2d(TSTEPS, N, A):
    """seidel_2d_numpy.kernel(TSTEPS, N, A): A updated in place."""
    if tuple(A.shape) != (int(N), int(N)):
        raise ValueError(f"A has shape {tuple(A.shape)}, expected ({N}, {N})")
    with _Staged("fp64", [A], outputs=(0,)) as (a,):
        seidel2d(a, TSTEPS)


def channel_flow_dropin(nit, u, v, dt, dx, dy, p, rho, nu, F):
    """channel_flow_numpy.channel_flow(...): u, v and p updated in place; the
    step count returned."""
    with _Staged("fp64", [u, v, p], outputs=(0, 1, 2)) as (a, b, c):
        return channel_flow(nit, a, b, dt, dx, dy, c, rho, nu, F)


def cavity_flow_dropin(nx, ny, nt, nit, u, v, dt, dx, dy, p, rho, nu):
    """cavity_flow_numpy.cavity_flow(...): u, v and p updated in place."""
    with _Staged("fp64", [u, v, p], outputs=(0, 1, 2)) as (a, b, c):
        cavity_flow(nx, ny, nt, nit, a, b, dt, dx, dy, c, rho, nu)
