"""Device plumbing: torch ROCm tensors as HBM buffers, stream handles, and the
deadline-aware synchronisation used by the framework adapter.

PyTorch only provides memory and streams here; every kernel is a HIP kernel
of libnpbx.so called through the C ABI.
"""

import time

import numpy as np
import torch

from . import _native as nat

# storage dtype code <-> torch dtype
TORCH_DTYPE = {nat.F64: torch.float64, nat.F32: torch.float32, nat.BF16: torch.bfloat16}
DTYPE_CODE = {v: k for k, v in TORCH_DTYPE.items()}
PRECISION_CODE = {"fp64": nat.F64, "fp32": nat.F32, "bf16": nat.BF16}
ITEMSIZE = {nat.F64: 8, nat.F32: 4, nat.BF16: 2}

_initialised = set()


def require_device(index: int = 0) -> torch.device:
    """Returns the HIP device, checking the library and the device once."""
    if not torch.cuda.is_available():
        raise RuntimeError("no HIP device is visible; the mi355x framework needs an MI355X")
    if index not in _initialised:
        torch.cuda.set_device(index)
        nat.call("npbx_init", index)
        _initialised.add(index)
    return torch.device("cuda", index)


def stream_handle(stream=None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


def ptr(t: torch.Tensor) -> int:
    return int(t.data_ptr())


def dtype_code(t: torch.Tensor) -> int:
    try:
        return DTYPE_CODE[t.dtype]
    except KeyError:
        raise TypeError(f"unsupported storage dtype {t.dtype}; use float64, float32 or "
                        "bfloat16") from None


def check_device_array(name: str, t, shape=None, dtype=None):
    """Validates a device operand: a contiguous tensor on the HIP device."""
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor on the HIP device, got {type(t)}")
    if t.device.type != "cuda":
        raise ValueError(f"{name} is on {t.device}, expected the HIP device")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be C-contiguous")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")
    if dtype is not None and t.dtype != dtype:
        raise TypeError(f"{name} has dtype {t.dtype}, expected {dtype}")
    return t


def to_device(a: np.ndarray, dtype=None) -> torch.Tensor:
    """Host array -> new contiguous device tensor (optionally converted)."""
    dev = require_device()
    t = torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    if dtype is not None and t.dtype != dtype:
        t = t.to(dtype)
    return t


def to_host(t) -> np.ndarray:
    if isinstance(t, torch.Tensor):
        if t.dtype == torch.bfloat16:
            t = t.float()
        return t.detach().cpu().numpy()
    return np.asarray(t)


def sync(stream=None, timeout_s: float = None, poll_s: float = 0.05):
    """Waits for ``stream`` with an optional timeout.

    The wait is split into short native polls so that a KeyboardInterrupt
    raised by a host timer (the harness's timeout) is delivered between them.
    Raises ``NpbxTimeout`` when ``timeout_s`` passes first.
    """
    h = stream_handle(stream)
    # time.monotonic_ns() is CLOCK_MONOTONIC, the clock npbx_sync compares to.
    end = None if timeout_s is None else time.monotonic_ns() + int(timeout_s * 1e9)
    while True:
        step_end = time.monotonic_ns() + int(poll_s * 1e9)
        if end is not None:
            step_end = min(step_end, end)
        try:
            nat.call("npbx_sync", h, step_end)
            return
        except nat.NpbxTimeout:
            if end is not None and time.monotonic_ns() >= end:
                raise


class Timer:
    """HIP-event timer recorded on a given stream (the stream the kernels run on)."""

    def __init__(self, stream=None):
        self.stream = stream_handle(stream)
        import ctypes
        self._start = ctypes.c_void_p()
        self._stop = ctypes.c_void_p()
        nat.call("npbx_event_create", ctypes.byref(self._start))
        nat.call("npbx_event_create", ctypes.byref(self._stop))

    def start(self):
        nat.call("npbx_event_record", self._start, self.stream)

    def stop(self):
        nat.call("npbx_event_record", self._stop, self.stream)

    def elapsed_ms(self) -> float:
        import ctypes
        ms = ctypes.c_float()
        nat.call("npbx_event_elapsed_ms", self._start, self._stop, ctypes.byref(ms))
        return float(ms.value)

    def __del__(self):
        try:
            nat.lib().npbx_event_destroy(self._start)
            nat.lib().npbx_event_destroy(self._stop)
        except Exception:
            pass
