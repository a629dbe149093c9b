"""Runs one framework on one benchmark: first run under a deadline,
validation against NumPy, timed repetitions, and result rows.

Follows npbench/infrastructure/test.py:11-163:
  * NumPy runs once on the preset's data to produce the reference outputs;
  * for each implementation of the framework, the first execution runs under
    ``timeout.exit_after`` (a timeout skips the implementation), its outputs
    are copied back and validated with rtol/atol/norm_error (defaults 1e-5,
    1e-8, 1e-5, overridable per benchmark spec), then ``repeat`` timed
    executions are recorded, each as one row of the ``results`` table;
  * outputs are the returned values followed by the spec's output_args.
Deliberate difference: an exception during validation marks the
implementation as not validated (the reference leaves it validated).
Each row also goes to a JSON-lines sidecar with the preset's grid-point
updates so GLUP/s can be derived.
"""

import time

from . import timeout as tout
from . import utilities as util


class Test:
    __test__ = False  # a benchmark runner, not a pytest test class

    def __init__(self, bench, frmwrk, npfrmwrk=None):
        self.bench = bench
        self.frmwrk = frmwrk
        self.numpy = npfrmwrk

    def _execute(self, frmwrk, impl, impl_name, mode, bdata, repeat, ignore_errors):
        label = f"{frmwrk.info['full_name']} - {impl_name}"
        try:
            copy = frmwrk.copy_func()
            setup = frmwrk.setup_str(self.bench, impl)
            stmt = frmwrk.exec_str(self.bench, impl)
        except Exception as e:
            print(f"Failed to load the {label} implementation.")
            print(e)
            if not ignore_errors:
                raise
            return None, None
        ldict = {"__npb_impl": impl, "__npb_copy": copy, **bdata}
        try:
            out, times = util.benchmark(stmt, setup, f"{label} - {mode}", repeat, ldict,
                                        "__npb_result")
        except Exception as e:
            print(f"Failed to execute the {label} implementation.")
            print(e)
            if not ignore_errors:
                raise
            return None, None
        if out is None:
            out = []
        elif isinstance(out, (tuple, list)):
            out = list(out)
        else:
            out = [out]
        if "output_args" in self.bench.info:
            n_ret = len(out)
            out += [ldict[a] for a in frmwrk.inout_args(self.bench)]
            if len(out) != n_ret + len(self.bench.info["output_args"]):
                raise AssertionError("Number of output arguments does not match.")
        return out, times

    def run(self, preset, validate, repeat, timeout=200.0, ignore_errors=True,
            database="npbench.db", sidecar=None):
        """Returns the list of result rows (also written to ``database``)."""
        info = self.bench.info
        print(f"***** Testing {self.frmwrk.info['full_name']} with {self.bench.bname} on the "
              f"{preset} dataset *****", flush=True)
        bdata = self.bench.get_data(preset)

        np_out = None
        if validate and self.frmwrk.fname != "numpy" and self.numpy is not None:
            np_impl, np_name = self.numpy.implementations(self.bench)[0]
            np_out, _ = self._execute(self.numpy, np_impl, np_name, "validation", bdata, 1,
                                      ignore_errors)
        else:
            validate = False

        version = str(self.frmwrk.version())
        context = {**bdata, **self.frmwrk.imports()}

        @tout.exit_after(timeout)
        def first_execution(impl, impl_name):
            return self._execute(self.frmwrk, impl, impl_name, "first/validation", context, 1,
                                 ignore_errors)

        rows = []
        for impl, impl_name in self.frmwrk.implementations(self.bench):
            try:
                fw_out, _ = first_execution(impl, impl_name)
            except KeyboardInterrupt:
                print(f'Implementation "{impl_name}" timed out.', flush=True)
                continue
            except Exception:
                if not ignore_errors:
                    raise
                continue
            if fw_out is None:
                continue  # failed to load or execute, already reported

            valid = True
            if validate and np_out is not None:
                name = f"{self.frmwrk.info['full_name']} - {impl_name}"
                try:
                    back = self.frmwrk.copy_back_func()
                    fw_host = [back(a) for a in fw_out]
                    valid = util.validate(np_out, fw_host, name,
                                          rtol=info.get("rtol", 1e-5),
                                          atol=info.get("atol", 1e-8),
                                          norm_error=info.get("norm_error", 1e-5))
                    if valid:
                        print(f"{name} - validation: SUCCESS")
                    elif not ignore_errors:
                        raise ValueError(f"{name} did not validate!")
                except ValueError:
                    raise
                except Exception as e:
                    valid = False
                    print(f"Failed to run {self.frmwrk.info['full_name']} validation: {e}")
                    if not ignore_errors:
                        raise

            _, times = self._execute(self.frmwrk, impl, impl_name, "median", context, repeat,
                                     ignore_errors)
            for t in times or []:
                rows.append({"details": impl_name, "validated": bool(valid), "time": t})

        stamp = int(time.time())
        records = [{
            "timestamp": stamp,
            "benchmark": info["short_name"],
            "kind": info.get("kind", ""),
            "domain": info.get("domain", ""),
            "dwarf": info.get("dwarf", ""),
            "preset": preset,
            "mode": "main",
            "framework": self.frmwrk.info["simple_name"],
            "version": version,
            "details": r["details"],
            "validated": r["validated"],
            "time": r["time"],
        } for r in rows]
        if database:
            conn = util.open_results(database)
            try:
                util.insert_results(conn, records)
            finally:
                conn.close()
        if sidecar:
            lups = self.bench.lups(preset)
            util.append_jsonl(sidecar, [
                {**rec, "lups": lups,
                 "glups": (lups / rec["time"] / 1e9) if lups and rec["time"] > 0 else None}
                for rec in records])
        return records
