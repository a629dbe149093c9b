"""End to end through the harness counterpart on the GPU: the mi355x
framework entry runs each benchmark's S preset with validation against
NumPy, as npbench's run_benchmark.py would (test.py:53-163)."""

import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bench", ["jacobi_2d", "heat_3d", "hdiff", "vadv"])
def test_framework_entry_validates(bench, tmp_path):
    from npbench_amd.infrastructure import Benchmark, Test, generate_framework
    print(f"harness {bench} S", flush=True)
    fw = generate_framework("mi355x")
    t = Test(Benchmark(bench), fw, generate_framework("numpy"))
    rows = t.run("S", validate=True, repeat=2, timeout=120.0, ignore_errors=True,
                 database=str(tmp_path / "npbench.db"), sidecar=str(tmp_path / "rows.jsonl"))
    by = {}
    for r in rows:
        by.setdefault(r["details"], []).append(r["validated"])
    # the FP64 drop-in and the FP32 storage form validate at rtol 1e-5
    assert by["default"] and all(by["default"])
    assert by["fp32"] and all(by["fp32"])
    # BF16 storage is recorded (validated or not) like any other variant
    if bench != "vadv":
        assert "bf16" in by


def test_run_benchmark_cli_on_the_gpu(tmp_path):
    """python -m npbench_amd.run_benchmark -b heat_3d -f mi355x -p S: every
    row of the FP64 drop-in validates against NumPy."""
    import sqlite3
    from npbench_amd import run_benchmark
    print("run_benchmark heat_3d mi355x S", flush=True)
    db = tmp_path / "r.db"
    assert run_benchmark.main(["-b", "heat_3d", "-f", "mi355x", "-r", "2", "--database",
                               str(db)]) == 0
    con = sqlite3.connect(db)
    rows = con.execute("select details, validated from results where framework='mi355x'")\
        .fetchall()
    con.close()
    assert rows and all(v for d, v in rows if d == "default")


def test_finalize_releases_and_recovers():
    """npbx_finalize frees the workspace; a later call that needs it (vadv)
    allocates it again and still matches the oracle bitwise."""
    import numpy as np
    import torch
    from npbench_amd import _native, device, stencils
    from oracle import stencils as ref
    device.require_device()
    rng = np.random.default_rng(4)
    I, J, K = 8, 12, 16
    f = [rng.random((I, J, K)), rng.random((I, J, K)), 2 * rng.random((I + 1, J, K)) - 1,
         rng.random((I, J, K)), rng.random((I, J, K))]
    _native.call("npbx_finalize")
    d = [device.to_device(a) for a in f]
    stencils.vadv(*d, 0.15)
    r = [a.copy() for a in f]
    ref.vadv(*r, 0.15)
    torch.cuda.synchronize()
    assert np.array_equal(d[0].cpu().numpy(), r[0])
    _native.call("npbx_finalize")
