"""Timing, validation and result-store helpers of the harness counterpart.

Semantics follow npbench/infrastructure/utilities.py:
  benchmark        (:135-151) timeit.repeat(stmt, setup, repeat, number=1)
                   over a copy of the context, median printed in ms, then
                   setup+stmt run once more in the caller's dict to capture
                   the output -- so an implementation runs repeat+1 times.
  validate         (:154-180) per output pair: allclose(rtol, atol), else the
                   relative 2-norm error must be below norm_error.
  relative_error   (:27-28)   ||ref - val||_2 / ||ref||_2.
  results schema   (:75-98)   the `results` table, unchanged, so MI355X rows
                   sit beside other frameworks' rows.
Differences, on purpose: ``validate`` raises when the lists differ in length
(the reference's zip truncates silently) and torch tensors and bfloat16 are
converted to float64 NumPy before comparison.
"""

import argparse
import json
import sqlite3
import statistics
import timeit

import numpy as np


def str2bool(v):
    if isinstance(v, bool):
        return v
    s = str(v).strip().lower()
    if s in ("yes", "true", "t", "y", "1"):
        return True
    if s in ("no", "false", "f", "n", "0"):
        return False
    raise argparse.ArgumentTypeError("boolean value expected")


def time_to_ms(seconds):
    return int(round(seconds * 1000))


def as_numpy(x):
    """Host float64-compatible view of a NumPy array, scalar or torch tensor."""
    if type(x).__module__.startswith("torch"):
        t = x.detach()
        if t.dtype.is_floating_point and t.dtype.itemsize < 4:
            t = t.float()
        return t.cpu().numpy()
    return np.asarray(x)


def relative_error(ref, val):
    ref = as_numpy(ref)
    val = as_numpy(val)
    return float(np.linalg.norm(ref - val) / np.linalg.norm(ref))


def benchmark(stmt, setup="pass", out_text="", repeat=1, context=None, output=None,
              verbose=True):
    """Times ``stmt`` (median of ``repeat`` runs, number=1) and captures ``output``."""
    context = {} if context is None else context
    scratch = dict(context)
    times = timeit.repeat(stmt, setup=setup, repeat=repeat, number=1, globals=scratch)
    if verbose:
        print(f"{out_text}: {time_to_ms(statistics.median(times))}ms", flush=True)
    result = None
    if output is not None:
        exec(setup, context)
        exec(stmt, context)
        result = context[output]
    return result, times


def validate(ref, val, framework="Unknown", rtol=1e-5, atol=1e-8, norm_error=1e-5):
    refs = list(ref) if isinstance(ref, (list, tuple)) else [ref]
    vals = list(val) if isinstance(val, (list, tuple)) else [val]
    if len(refs) != len(vals):
        raise ValueError(f"{framework}: {len(vals)} outputs to validate against {len(refs)}")
    valid = True
    for r, v in zip(refs, vals):
        r = as_numpy(r)
        v = as_numpy(v)
        if np.allclose(r, v, rtol=rtol, atol=atol):
            continue
        err = relative_error(r, v)
        if err < norm_error:
            continue
        valid = False
        print(f"Relative error: {err}")
    if not valid:
        print(f"{framework} did not validate!")
    return valid


def error_report(ref, val):
    """Max inf-norm relative error, elementwise-floored error and 2-norm ratio."""
    r = as_numpy(ref).astype(np.float64)
    v = as_numpy(val).astype(np.float64)
    scale = np.max(np.abs(r)) if r.size else 0.0
    diff = np.abs(v - r)
    inf_rel = float(np.max(diff) / scale) if scale > 0 else float(np.max(diff, initial=0.0))
    floor = np.maximum(np.abs(r), 1e-3 * scale)
    elem = float(np.max(diff / floor)) if scale > 0 else inf_rel
    l2 = float(np.linalg.norm(r - v) / np.linalg.norm(r)) if scale > 0 else 0.0
    return {"max_rel_inf": inf_rel, "max_rel_elem_floored": elem, "rel_l2": l2}


# ---- result store (schema of npbench/infrastructure/utilities.py:75-98) ----
SQL_CREATE_RESULTS = """
CREATE TABLE IF NOT EXISTS results (
    id integer PRIMARY KEY,
    timestamp integer NOT NULL,
    benchmark text NOT NULL,
    kind text,
    domain text,
    dwarf text,
    preset text NOT NULL,
    mode text NOT NULL,
    framework text NOT NULL,
    version text NOT NULL,
    details text,
    validated integer,
    time real
);
"""

SQL_INSERT_RESULT = """
INSERT INTO results(
    timestamp, benchmark, kind, domain, dwarf, preset, mode,
    framework, version, details, validated, time
) VALUES (?, ?, ?, ?, ?, ?, ?, ?, ?, ?, ?, ?);
"""

RESULT_FIELDS = ("timestamp", "benchmark", "kind", "domain", "dwarf", "preset", "mode",
                 "framework", "version", "details", "validated", "time")


def open_results(db_path):
    conn = sqlite3.connect(db_path)
    conn.execute(SQL_CREATE_RESULTS)
    conn.commit()
    return conn


def insert_results(conn, rows):
    conn.executemany(SQL_INSERT_RESULT, [tuple(r[f] for f in RESULT_FIELDS) for r in rows])
    conn.commit()


def append_jsonl(path, records):
    with open(path, "a") as f:
        for rec in records:
            f.write(json.dumps(rec) + "\n")
