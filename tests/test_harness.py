"""The harness counterpart (CPU only): specs, adapter strings, the runner's
validation / timeout / error semantics, and the result store.

Reference behaviour followed: npbench/infrastructure/benchmark.py:31-66,
framework.py:89-162, test.py:16-163, utilities.py:135-180,
timeout_decorator.py:46-72.
"""

import json
import sqlite3
import time

import numpy as np
import pytest

from npbench_amd.infrastructure import (Benchmark, Framework, MI355XFramework, Test,
                                        generate_framework, timeout, utilities as util)
from npbench_amd.benchmarks.polybench.jacobi_2d import jacobi_2d
from oracle import stencils as ref


# ------------------------------------------------------------------ specs
def test_get_data_reference_preset_and_cache():
    b = Benchmark("jacobi_2d")
    d = b.get_data("S")
    assert d["TSTEPS"] == 50 and d["N"] == 150
    A, B = jacobi_2d.initialize(150)
    assert np.array_equal(d["A"], A) and np.array_equal(d["B"], B)
    assert b.get_data("S") is d  # cached per preset


def test_get_data_config_preset_uses_seeded_init():
    b = Benchmark("heat_3d")
    spec = b.init_spec("C2")  # (256^3 itself is too large to build in a unit test)
    assert spec["func_name"] == "initialize_uniform" and spec["input_args"] == ["N", "seed"]
    b2 = Benchmark("vadv")
    b2.info["parameters"]["T"] = {"I": 3, "J": 4, "K": 5, "seed": 42}
    b2.info["preset_init"]["T"] = b2.info["preset_init"]["C4"]
    d = b2.get_data("T")
    assert d["dtr_stage"] == 0.15 and d["wcon"].shape == (4, 4, 5)
    assert d["wcon"].min() < 0.0  # uniform[-1, 1)


def test_unknown_preset_raises():
    with pytest.raises(NotImplementedError):
        Benchmark("hdiff").get_data("XL")


def test_lups():
    assert Benchmark("jacobi_2d").lups("C1") == 2 * 100 * 998 ** 2
    assert Benchmark("heat_3d").lups("C2") == 200 * 254 ** 3
    assert Benchmark("vadv").lups("paper") == 256 * 256 * 160


# ------------------------------------------------------------------ adapters
def test_numpy_framework_strings():
    fw = generate_framework("numpy")
    b = Benchmark("jacobi_2d")
    assert fw.setup_str(b) == "__npb_np_A, __npb_np_B = __npb_copy(A), __npb_copy(B)"
    assert fw.exec_str(b) == "__npb_result = __npb_impl(TSTEPS, __npb_np_A, __npb_np_B)"
    assert fw.inout_args(b) == ["__npb_np_A", "__npb_np_B"]


def test_mi355x_framework_is_resolved_and_syncs():
    fw = generate_framework("mi355x")
    assert isinstance(fw, MI355XFramework)
    b = Benchmark("vadv")
    assert fw.info["prefix"] == "mx" and fw.info["arch"] == "gpu"
    assert fw.exec_str(b) == ("__npb_result = __npb_impl(__npb_mx_utens_stage, __npb_mx_u_stage, "
                              "__npb_mx_wcon, __npb_mx_u_pos, __npb_mx_utens, dtr_stage); "
                              "__npb_sync()")
    assert fw.setup_str(b).endswith("; __npb_sync()")


def test_mi355x_implementations_list_variants():
    fw = generate_framework("mi355x")
    names = [n for _, n in fw.implementations(Benchmark("hdiff"))]
    assert names == ["default", "fp32", "bf16"]
    names = [n for _, n in fw.implementations(Benchmark("vadv"))]
    assert names == ["default", "fp32"]


def test_numpy_framework_uses_the_oracle_modules():
    fw = generate_framework("numpy")
    (impl, name), = fw.implementations(Benchmark("vadv"))
    assert impl is ref.vadv and name == "default"


# ------------------------------------------------------------------ runner
class FakeFramework(Framework):
    """A CPU framework whose implementations are supplied by the test."""

    def __init__(self, impls):
        self.fname = "fake"
        self.info = {"simple_name": "fake", "full_name": "Fake", "prefix": "fk",
                     "postfix": "fake", "class": "FakeFramework", "arch": "cpu"}
        self._impls = impls

    def version(self):
        return "0"

    def implementations(self, bench):
        return self._impls


def tiny_jacobi():
    b = Benchmark("jacobi_2d")
    b.info = json.loads(json.dumps(b.info))
    b.info["parameters"]["T"] = {"TSTEPS": 4, "N": 12}
    return b


def good(TSTEPS, A, B):
    ref.jacobi_2d_kernel(TSTEPS, A, B)


def wrong(TSTEPS, A, B):
    ref.jacobi_2d_kernel(TSTEPS, A, B)
    A[5, 5] += 1.0


def raises(TSTEPS, A, B):
    raise RuntimeError("boom")


def slow(TSTEPS, A, B):
    for _ in range(200):  # sleeps in small slices so the interrupt lands
        time.sleep(0.01)
    ref.jacobi_2d_kernel(TSTEPS, A, B)


def run(impls, tmp_path, **kw):
    t = Test(tiny_jacobi(), FakeFramework(impls), generate_framework("numpy"))
    db = tmp_path / "npbench.db"
    side = tmp_path / "rows.jsonl"
    rows = t.run("T", validate=True, repeat=3, database=str(db), sidecar=str(side), **kw)
    return rows, db, side


def test_runner_validates_and_records(tmp_path):
    rows, db, side = run([(good, "good"), (wrong, "wrong")], tmp_path)
    by = {}
    for r in rows:
        by.setdefault(r["details"], []).append(r)
    assert len(by["good"]) == 3 and all(r["validated"] for r in by["good"])
    assert len(by["wrong"]) == 3 and not any(r["validated"] for r in by["wrong"])
    with sqlite3.connect(db) as conn:
        got = conn.execute("select benchmark, preset, mode, framework, details, validated "
                           "from results").fetchall()
        cols = [c[1] for c in conn.execute("pragma table_info(results)")]
    assert cols == ["id", "timestamp", "benchmark", "kind", "domain", "dwarf", "preset", "mode",
                    "framework", "version", "details", "validated", "time"]
    assert ("jacobi2d", "T", "main", "fake", "good", 1) in got
    assert ("jacobi2d", "T", "main", "fake", "wrong", 0) in got
    lines = [json.loads(x) for x in side.read_text().splitlines()]
    assert all(x["lups"] == 2 * 3 * 10 ** 2 and x["glups"] > 0 for x in lines)


def test_runner_skips_a_raising_implementation(tmp_path):
    rows, _, _ = run([(raises, "bad"), (good, "good")], tmp_path)
    assert {r["details"] for r in rows} == {"good"}


def test_runner_raises_when_errors_are_fatal(tmp_path):
    with pytest.raises(RuntimeError):
        run([(raises, "bad")], tmp_path, ignore_errors=False)


def test_runner_times_out_a_slow_implementation(tmp_path):
    rows, _, _ = run([(slow, "slow"), (good, "good")], tmp_path, timeout=0.3)
    assert {r["details"] for r in rows} == {"good"}


def test_validation_exception_marks_invalid(tmp_path):
    class BadBack(FakeFramework):
        def copy_back_func(self):
            def back(x):
                raise TypeError("no copy back")
            return back

    t = Test(tiny_jacobi(), BadBack([(good, "good")]), generate_framework("numpy"))
    rows = t.run("T", validate=True, repeat=1, database=str(tmp_path / "x.db"))
    assert rows and not any(r["validated"] for r in rows)


def test_benchmark_runs_repeat_plus_one(monkeypatch):
    calls = []
    ctx = {"f": lambda: calls.append(1)}
    out, times = util.benchmark("__npb_result = f()", "pass", "x", 4, ctx, "__npb_result",
                                verbose=False)
    assert len(times) == 4 and len(calls) == 5


def test_validate_semantics():
    r = np.linspace(1.0, 2.0, 100)
    assert util.validate(r, r * (1 + 1e-7))
    assert util.validate([r, r], [r, r + 1e-9])
    v = r.copy()
    v[0] += 1e-3  # fails allclose and the 2-norm criterion (||d||/||r|| ~ 6.5e-5)
    assert not util.validate(r, v)
    with pytest.raises(ValueError):
        util.validate([r, r], [r])


def test_error_report():
    r = np.array([1.0, -4.0, 0.0])
    v = np.array([1.0, -4.0 + 4e-6, 1e-6])
    rep = util.error_report(r, v)
    assert rep["max_rel_inf"] == pytest.approx(1e-6)
    assert rep["max_rel_elem_floored"] == pytest.approx(1e-6 / 4e-3)


def test_exit_after_interrupts_main_thread():
    @timeout.exit_after(0.2)
    def spin():
        end = time.time() + 5
        while time.time() < end:
            time.sleep(0.01)

    with pytest.raises(KeyboardInterrupt):
        spin()

    @timeout.exit_after(1.0)
    def quick():
        return 7

    assert quick() == 7


# ------------------------------------------------------------------ autotuner
def test_tuner_search_space_and_cache(tmp_path, monkeypatch):
    torch = pytest.importorskip("torch")
    from npbench_amd import tune
    heat = tune.candidates("heat_3d", torch.float32)
    assert {"T": 1, "rows": 0, "seg": 0, "shape": 3} in heat
    assert {(c["T"], c["rows"]) for c in heat if c["T"] > 1} == set(tune.SHAPES[torch.float32])
    jac = tune.candidates("jacobi_2d", torch.float64)
    assert [c["T"] for c in jac] == [1, 2]
    # the cache round-trips without a device (measurement monkeypatched)
    calls = []

    def fake_measure(kind, dtype, n, count, cfg, sweeps=16, reps=2, seed=0):
        calls.append(cfg)
        return 100.0 + cfg["T"] * 10 + cfg["rows"] - cfg["seg"] / 100

    monkeypatch.setattr(tune, "measure", fake_measure)
    monkeypatch.setattr(tune, "_memo", {})
    path = tmp_path / "tune.json"
    cfg = tune.best("heat_3d", torch.float32, 256, 16, cache=path)
    assert cfg["T"] == 4 and cfg["rows"] == 4 and cfg["seg"] == 32
    n_calls = len(calls)
    monkeypatch.setattr(tune, "_memo", {})
    assert tune.best("heat_3d", torch.float32, 256, 16, cache=path) == cfg
    assert len(calls) == n_calls  # served from the file


# ------------------------------------------------------------------ report
def test_speedup_report_from_the_result_store(tmp_path):
    from npbench_amd import report
    db = tmp_path / "npbench.db"
    conn = util.open_results(str(db))
    rows = []
    for fw, det, t, ok in (("numpy", "default", 1.0, True), ("mi355x", "default", 0.01, True),
                           ("mi355x", "bf16", 0.005, False)):
        for rep in range(5):
            rows.append({"timestamp": 0, "benchmark": "heat3d", "kind": "", "domain": "",
                         "dwarf": "", "preset": "S", "mode": "main", "framework": fw,
                         "version": "1", "details": det, "validated": ok,
                         "time": t * (1 + 0.01 * rep)})
    util.insert_results(conn, rows)
    conn.close()
    sp = report.speedups(report.load(str(db)))
    r = sp[("heat3d", "S")]["mi355x"]
    assert r["details"] == "bf16" and not r["validated"]  # fastest implementation kept
    assert 150 < r["speedup"] < 250 and r["ci"][0] <= r["speedup"] <= r["ci"][1]
    md = report.markdown(sp)
    assert "not validated" in md and "heat3d" in md
    pytest.importorskip("matplotlib")
    report.heatmap(sp, str(tmp_path / "s.png"))
    assert (tmp_path / "s.png").stat().st_size > 0


def test_run_benchmark_cli(tmp_path):
    """The run_benchmark.py counterpart: options, a NumPy run of the S
    preset, result rows in the database and the JSON-lines file."""
    import json
    import sqlite3
    from npbench_amd import run_benchmark
    a = run_benchmark.parse(["-b", "jacobi_2d"])
    assert (a["framework"], a["preset"], a["validate"], a["repeat"], a["timeout"]) == \
        ("numpy", "S", True, 10, 200.0)
    with pytest.raises(SystemExit):
        run_benchmark.parse(["-b", "gemm"])  # only the stencil path is built
    db, jl = tmp_path / "r.db", tmp_path / "r.jsonl"
    rc = run_benchmark.main(["-b", "jacobi_2d", "-f", "numpy", "-r", "2", "-v", "false",
                             "--database", str(db), "--jsonl", str(jl)])
    assert rc == 0
    con = sqlite3.connect(db)
    n = con.execute("select count(*) from results where benchmark='jacobi2d' "
                    "and framework='numpy'").fetchone()[0]
    con.close()
    assert n == 2
    rows = [json.loads(x) for x in jl.read_text().splitlines()]
    assert len(rows) == 2 and all(r["glups"] > 0 for r in rows)
