"""All-benchmarks driver of npbench's run_framework.py (run_framework.py:1-83):
every benchmark in bench_info/ for one framework and preset, each in a
process of its own, so that a benchmark that crashes, aborts or raises
cannot take the others down; those whose process ends with a non-zero exit
code are listed as failed at the end.

    python -m npbench_amd.run_framework -f mi355x -p paper -r 10 -t 200

Options follow run_framework.py:22-58 (the DaCe SDFG options are accepted
and ignored) plus ``--database``, ``--jsonl`` and ``--benchmarks`` (a subset,
comma-separated).  Processes are started with the "spawn" method: each child
initialises HIP on its own, and the parent never touches the GPU.
"""

import argparse
import multiprocessing as mp
import sys

from .infrastructure import utilities as util
from .infrastructure.benchmark import BENCH_INFO


def benchmark_names():
    """Every benchmark with a spec in bench_info/, sorted (run_framework.py:64-67)."""
    return sorted(p.stem for p in BENCH_INFO.glob("*.json"))


def run_benchmark(benchname, fname, preset, validate, repeat, timeout, ignore_errors,
                  database="npbench.db", sidecar=None):
    """The child's work (run_framework.py:12-20): one benchmark, one framework,
    NumPy for validation.  Line counting (pygount) is not provided."""
    from .infrastructure import Benchmark, Test, generate_framework
    frmwrk = generate_framework(fname)
    numpy = generate_framework("numpy")
    test = Test(Benchmark(benchname), frmwrk, numpy)
    test.run(preset, validate, repeat, timeout, ignore_errors, database=database, sidecar=sidecar)


def run_all(benchnames, fname, preset, validate=True, repeat=10, timeout=200.0,
            ignore_errors=True, database="npbench.db", sidecar=None, runner=run_benchmark):
    """Runs ``runner`` for each benchmark in its own spawned process, one after
    another; returns the names whose process exited non-zero
    (run_framework.py:68-78)."""
    ctx = mp.get_context("spawn")
    failed = []
    for name in benchnames:
        p = ctx.Process(target=runner, args=(name, fname, preset, validate, repeat, timeout,
                                             ignore_errors, database, sidecar))
        p.start()
        p.join()
        if p.exitcode != 0:
            failed.append(name)
    return failed


def parse(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("-f", "--framework", type=str, nargs="?", default="numpy")
    ap.add_argument("-p", "--preset", type=str, nargs="?", default="S",
                    help="S, M, L, paper, or a configuration preset of the spec (e.g. C2)")
    ap.add_argument("-m", "--mode", type=str, nargs="?", default="main")
    ap.add_argument("-v", "--validate", type=util.str2bool, nargs="?", default=True)
    ap.add_argument("-r", "--repeat", type=int, nargs="?", default=10)
    ap.add_argument("-t", "--timeout", type=float, nargs="?", default=200.0)
    ap.add_argument("--ignore-errors", type=util.str2bool, nargs="?", default=True)
    ap.add_argument("-s", "--save-strict-sdfg", type=util.str2bool, nargs="?", default=False)
    ap.add_argument("-l", "--load-strict-sdfg", type=util.str2bool, nargs="?", default=False)
    ap.add_argument("--database", type=str, default="npbench.db")
    ap.add_argument("--jsonl", type=str, default=None)
    ap.add_argument("--benchmarks", type=str, default=None,
                    help="comma-separated subset of the benchmarks (default: all)")
    return vars(ap.parse_args(argv))


def main(argv=None):
    args = parse(argv)
    names = benchmark_names()
    if args["benchmarks"]:
        wanted = [b.strip() for b in args["benchmarks"].split(",") if b.strip()]
        unknown = sorted(set(wanted) - set(names))
        if unknown:
            raise SystemExit(f"unknown benchmarks: {', '.join(unknown)}")
        names = [b for b in names if b in wanted]
    failed = run_all(names, args["framework"], args["preset"], args["validate"], args["repeat"],
                     args["timeout"], args["ignore_errors"], args["database"], args["jsonl"])
    if failed:
        print(f"Failed: {len(failed)} out of {len(names)}")
        for name in failed:
            print(name)
    return 0


if __name__ == "__main__":
    sys.exit(main())
