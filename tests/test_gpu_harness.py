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
