"""Speedup-versus-NumPy report from the harness's result store.

Reads the ``results`` table of ``npbench.db`` (the reference schema,
npbench/infrastructure/utilities.py:75-98) and, like
npbench's plot_results.py:54-168, keeps for each (benchmark, framework) the
implementation (``details``) with the lowest median time, reports its speedup
over NumPy's median on the same preset with a bootstrap confidence interval,
and marks rows that did not validate.  Output: a Markdown table, and a
heatmap PNG when matplotlib is available.

    python -m npbench_amd.report npbench.db [--preset paper] [--png speedup.png]
"""

import argparse
import sqlite3
import statistics

import numpy as np


def load(db_path, preset=None):
    con = sqlite3.connect(db_path)
    try:
        q = "select benchmark, preset, framework, details, validated, time from results"
        rows = con.execute(q).fetchall()
    finally:
        con.close()
    out = {}
    for bench, pre, fw, det, valid, t in rows:
        if preset and pre != preset:
            continue
        out.setdefault((bench, pre, fw, det), {"times": [], "valid": True})
        out[(bench, pre, fw, det)]["times"].append(float(t))
        out[(bench, pre, fw, det)]["valid"] &= bool(valid)
    return out


def bootstrap_ratio_ci(num, den, n_boot=1000, seed=0, alpha=0.05):
    """CI of median(num) / median(den) by resampling both samples."""
    rng = np.random.default_rng(seed)
    num = np.asarray(num)
    den = np.asarray(den)
    r = [np.median(rng.choice(num, num.size)) / np.median(rng.choice(den, den.size))
         for _ in range(n_boot)]
    return float(np.quantile(r, alpha / 2)), float(np.quantile(r, 1 - alpha / 2))


def speedups(table):
    """{(bench, preset): {framework: row}} with the best implementation."""
    by = {}
    for (bench, pre, fw, det), rec in table.items():
        by.setdefault((bench, pre), {}).setdefault(fw, []).append((det, rec))
    out = {}
    for key, fws in by.items():
        if "numpy" not in fws:
            continue
        np_times = min(fws["numpy"], key=lambda x: statistics.median(x[1]["times"]))[1]["times"]
        np_med = statistics.median(np_times)
        for fw, impls in fws.items():
            det, rec = min(impls, key=lambda x: statistics.median(x[1]["times"]))
            med = statistics.median(rec["times"])
            lo, hi = bootstrap_ratio_ci(np_times, rec["times"])
            out.setdefault(key, {})[fw] = {"details": det, "median_s": med,
                                           "speedup": np_med / med, "ci": (lo, hi),
                                           "validated": rec["valid"]}
    return out


def markdown(sp):
    fws = sorted({fw for v in sp.values() for fw in v})
    lines = ["| benchmark | preset | " + " | ".join(fws) + " |",
             "|---|---|" + "---|" * len(fws)]
    for (bench, pre), v in sorted(sp.items()):
        cells = []
        for fw in fws:
            if fw not in v:
                cells.append("")
                continue
            r = v[fw]
            mark = "" if r["validated"] else " (not validated)"
            cells.append(f"{r['speedup']:.1f}x [{r['ci'][0]:.1f}, {r['ci'][1]:.1f}] "
                         f"{r['details']}{mark}")
        lines.append(f"| {bench} | {pre} | " + " | ".join(cells) + " |")
    return "\n".join(lines)


def heatmap(sp, path):
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    keys = sorted(sp)
    fws = sorted({fw for v in sp.values() for fw in v})
    data = np.full((len(keys), len(fws)), np.nan)
    for i, k in enumerate(keys):
        for j, fw in enumerate(fws):
            if fw in sp[k]:
                data[i, j] = np.log10(sp[k][fw]["speedup"])
    fig, ax = plt.subplots(figsize=(1.6 * len(fws) + 2, 0.5 * len(keys) + 1.5))
    im = ax.imshow(data, cmap="RdYlGn", vmin=-2, vmax=4, aspect="auto")
    ax.set_xticks(range(len(fws)), fws)
    ax.set_yticks(range(len(keys)), [f"{b} ({p})" for b, p in keys])
    for i, k in enumerate(keys):
        for j, fw in enumerate(fws):
            if fw in sp[k]:
                ax.text(j, i, f"{sp[k][fw]['speedup']:.1f}x", ha="center", va="center",
                        fontsize=8)
    fig.colorbar(im, ax=ax, label="log10 speedup over NumPy")
    fig.tight_layout()
    fig.savefig(path, dpi=120)
    plt.close(fig)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("db", nargs="?", default="npbench.db")
    ap.add_argument("--preset")
    ap.add_argument("--png")
    a = ap.parse_args(argv)
    sp = speedups(load(a.db, a.preset))
    print(markdown(sp))
    if a.png:
        heatmap(sp, a.png)


if __name__ == "__main__":
    main()
