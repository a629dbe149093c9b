"""The C ABI library loads and exports what include/npbx.h declares (CPU only).

No compute runs here: the calls below are rejected by argument checks that
happen before any HIP call.
"""

import ctypes
import os
import re

import pytest

from npbench_amd import _native as nat

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "npbx.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for m in re.finditer(r"\bint\s+(npbx_\w+)\s*\(([^)]*)\)\s*;", text):
        params = [p for p in m.group(2).split(",") if p.strip() and p.strip() != "void"]
        out[m.group(1)] = len(params)
    return out


def test_header_declares_entry_points():
    decl = declared()
    for name in ("npbx_jacobi2d", "npbx_heat3d", "npbx_hdiff", "npbx_vadv", "npbx_sync",
                 "npbx_cast", "npbx_last_error"):
        assert name in decl


def test_every_declared_symbol_is_exported_and_bound():
    lib = nat.lib()
    decl = declared()
    for name, nparams in decl.items():
        assert hasattr(lib, name), f"{name} not exported by libnpbx.so"
        assert name in nat.SIGNATURES, f"{name} has no ctypes prototype"
        assert len(nat.SIGNATURES[name]) == nparams, f"{name}: arity mismatch"
    assert set(nat.SIGNATURES) == set(decl), "prototypes for undeclared symbols"


def test_version_string():
    assert "gfx950" in nat.version()


@pytest.mark.parametrize("call,needle", [
    (lambda: nat.call("npbx_jacobi2d", 0, None, None, -5, 3, 1, None), "N=-5"),
    (lambda: nat.call("npbx_heat3d_sweeps", 0, None, None, 10, -1, 2, 1, None), "negative"),
    (lambda: nat.call("npbx_hdiff", 0, None, None, None, 4, 4, 4, 0, None), "nsteps"),
    (lambda: nat.call("npbx_vadv", 0, None, None, None, None, None, 0.15, 2, 2, 1, None), "K >= 2"),
    (lambda: nat.call("npbx_cast", 0, None, 7, None, 10, None), "null"),
    (lambda: nat.call("npbx_jacobi2d", 9, 1, 2, 10, 3, 1, None), "dtype"),
    (lambda: nat.call("npbx_adi", 0, None, 10, 0, None), "TSTEPS=0"),
    (lambda: nat.call("npbx_adi", 0, None, 0, 5, None), "N=0"),
    (lambda: nat.call("npbx_adi", 1, None, 10, 5, None), "dtype"),
    (lambda: nat.call("npbx_adi", 0, None, 10, 5, None), "null"),
    (lambda: nat.call("npbx_jacobi1d", 1, None, None, 10, 5, None), "dtype"),
    (lambda: nat.call("npbx_jacobi1d", 0, None, None, 10, 5, None), "distinct"),
    (lambda: nat.call("npbx_fdtd2d", 1, None, None, None, None, 4, 4, 2, None), "dtype"),
    (lambda: nat.call("npbx_fdtd2d", 0, None, None, None, None, 4, 4, 2, None), "null"),
    (lambda: nat.call("npbx_cavity_flow", 1, 5, 5, 2, 2, None, None, 0.1, 0.5, 0.5, None, 1.0,
                      0.1, None), "dtype"),
    (lambda: nat.call("npbx_cavity_flow", 0, 2, 5, 2, 2, None, None, 0.1, 0.5, 0.5, None, 1.0,
                      0.1, None), "out of range"),
    (lambda: nat.call("npbx_cavity_flow", 0, 5, 5, -1, 2, None, None, 0.1, 0.5, 0.5, None, 1.0,
                      0.1, None), "negative"),
    (lambda: nat.call("npbx_cavity_flow", 0, 5, 5, 2, 2, None, None, 0.1, 0.5, 0.5, None, 1.0,
                      0.1, None), "distinct"),
    (lambda: nat.call("npbx_channel_flow", 1, 2, None, None, 0.1, 0.5, 0.5, None, 1.0, 0.1, 1.0,
                      5, 5, 10, ctypes.byref(ctypes.c_int64()), None), "dtype"),
    (lambda: nat.call("npbx_channel_flow", 0, 2, None, None, 0.1, 0.5, 0.5, None, 1.0, 0.1, 1.0,
                      2, 5, 10, ctypes.byref(ctypes.c_int64()), None), "out of range"),
    (lambda: nat.call("npbx_channel_flow", 0, 2, None, None, 0.1, 0.5, 0.5, None, 1.0, 0.1, 1.0,
                      5, 5, 0, ctypes.byref(ctypes.c_int64()), None), "max_steps"),
    (lambda: nat.call("npbx_channel_flow", 0, 2, None, None, 0.1, 0.5, 0.5, None, 1.0, 0.1, 1.0,
                      5, 5, 10, None, None), "stepcount"),
    (lambda: nat.call("npbx_channel_flow", 0, 2, None, None, 0.1, 0.5, 0.5, None, 1.0, 0.1, 1.0,
                      5, 5, 10, ctypes.byref(ctypes.c_int64()), None), "distinct"),
    (lambda: nat.call("npbx_seidel2d", 1, None, 10, 3, None), "dtype"),
    (lambda: nat.call("npbx_seidel2d", 0, None, 20000, 3, None), "N=20000"),
    (lambda: nat.call("npbx_seidel2d", 0, None, 10, 3, None), "null"),
    (lambda: nat.call("npbx_set_tuning", b"seidel.blocks", 0), "seidel.blocks"),
    (lambda: nat.call("npbx_set_tuning", b"jacobi.deep", 9), "jacobi.deep"),
    (lambda: nat.call("npbx_set_tuning", b"hdiff.fuse_rows", 3), "hdiff.fuse_rows"),
    (lambda: nat.call("npbx_set_tuning", b"hdiff.shape", 14), "hdiff.shape"),
    (lambda: nat.call("npbx_set_tuning", b"jacobi.deep_waves", 2), "jacobi.deep_waves"),
    (lambda: nat.call("npbx_adi_noalloc", 0, None, 10, 5, None), "null"),
    (lambda: nat.call("npbx_adi_noalloc", 1, None, 10, 5, None), "dtype"),
    # rejected by the size check before any HIP call (the pointer is never used)
    (lambda: nat.call("npbx_adi_noalloc", 0, ctypes.c_void_p(256), 1001, 3, None), "N=1001"),
])
def test_argument_errors_raise_with_message(call, needle):
    with pytest.raises(nat.NpbxError) as ei:
        call()
    assert ei.value.code == nat.EINVAL
    assert needle in str(ei.value)


def test_empty_work_is_a_no_op():
    # nothing to update: no launch, no device needed
    assert nat.call("npbx_jacobi2d", 0, None, None, 2, 10, 1, None) == nat.OK
    assert nat.call("npbx_heat3d", 0, None, None, 10, 1, 1, None) == nat.OK
    assert nat.call("npbx_cast", 0, None, 1, None, 0, None) == nat.OK
    assert nat.call("npbx_adi", 0, None, 2, 5, None) == nat.OK    # no interior line
    assert nat.call("npbx_adi", 0, None, 10, -3, None) == nat.OK   # no time step
    assert nat.call("npbx_jacobi1d", 0, None, None, 2, 5, None) == nat.OK  # no interior
    assert nat.call("npbx_jacobi1d", 0, None, None, 10, 1, None) == nat.OK  # no sweep
    assert nat.call("npbx_fdtd2d", 0, None, None, None, None, 0, 5, 3, None) == nat.OK
    assert nat.call("npbx_fdtd2d", 0, None, None, None, None, 4, 5, 0, None) == nat.OK
    assert nat.call("npbx_cavity_flow", 0, 5, 5, 0, 3, None, None, 0.1, 0.5, 0.5, None, 1.0, 0.1,
                    None) == nat.OK  # no time step
    assert nat.call("npbx_seidel2d", 0, None, 2, 5, None) == nat.OK   # no interior
    assert nat.call("npbx_seidel2d", 0, None, 10, 1, None) == nat.OK  # no sweep


def test_last_error_buffer_is_terminated():
    with pytest.raises(nat.NpbxError):
        nat.call("npbx_hdiff", 0, None, None, None, -1, 1, 1, 1, None)
    buf = ctypes.create_string_buffer(8)
    assert nat.lib().npbx_last_error(buf, len(buf)) == nat.OK
    assert len(buf.value) == 7


def test_missing_library_is_loud(monkeypatch, tmp_path):
    monkeypatch.setattr(nat, "_lib", None)
    monkeypatch.setattr(nat, "LIB_PATH", tmp_path / "libnpbx.so")
    with pytest.raises(nat.NativeLibraryMissing):
        nat.lib()
