"""Command-line driver with npbench's run_benchmark.py options
(run_benchmark.py:6-60): one benchmark, one framework, one preset.

    python -m npbench_amd.run_benchmark -b heat_3d -f mi355x -p paper -r 10 -t 200

NumPy (framework "numpy", the oracle modules) runs first for validation, then
every implementation the framework lists is timed ``repeat`` times after a
first execution under the timeout; rows go to the ``results`` table of
``npbench.db`` (the reference's schema) and, with ``--jsonl``, to a JSON-lines
file with GLUP/s.  The reference's line counting (LineCount, which needs
pygount) and its DaCe SDFG options are not provided here.
"""

import argparse
import sys

from .infrastructure import Benchmark, Test, generate_framework
from .infrastructure import utilities as util

# The benchmarks with MI355X implementations (bench_info/*.json).
BENCHMARKS = ("jacobi_2d", "heat_3d", "hdiff", "vadv", "adi", "jacobi_1d", "fdtd_2d",
              "cavity_flow", "channel_flow", "seidel_2d")


def parse(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("-b", "--benchmark", type=str, required=True, choices=BENCHMARKS)
    ap.add_argument("-f", "--framework", type=str, default="numpy")
    ap.add_argument("-p", "--preset", type=str, default="S",
                    help="S, M, L, paper, or a configuration preset of the spec (e.g. C2)")
    ap.add_argument("-v", "--validate", type=util.str2bool, nargs="?", const=True, default=True)
    ap.add_argument("-r", "--repeat", type=int, default=10)
    ap.add_argument("-t", "--timeout", type=float, default=200.0)
    ap.add_argument("--database", type=str, default="npbench.db")
    ap.add_argument("--jsonl", type=str, default=None, help="also append rows with GLUP/s here")
    return vars(ap.parse_args(argv))


def main(argv=None):
    args = parse(argv)
    bench = Benchmark(args["benchmark"])
    frmwrk = generate_framework(args["framework"])
    numpy = generate_framework("numpy")
    test = Test(bench, frmwrk, numpy)
    rows = test.run(args["preset"], args["validate"], args["repeat"], args["timeout"],
                    database=args["database"], sidecar=args["jsonl"])
    return 0 if rows else 1


if __name__ == "__main__":
    sys.exit(main())
