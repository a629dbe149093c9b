"""Serving: continuous batching keeps every result bitwise the result of the
same request run alone (tolerance: exact), admission respects the HBM
budget, and latencies are recorded."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def small_mix(count, seed):
    from npbench_amd import serve
    return serve.poisson_mix(count, qps=2000.0, seed=seed, jacobi_sides=(16, 300),
                             heat_sides=(8, 48), steps=(1, 9))


def alone(req, dtype):
    from npbench_amd import serve, stencils
    A, B = serve.make_inputs(req, dtype)
    f = stencils.jacobi2d_sweeps if req.kind == "jacobi_2d" else stencils.heat3d_sweeps
    f(A, B, req.sweeps)
    torch.cuda.synchronize()
    return A, B


@pytest.mark.parametrize("tblock", [2, 4])
def test_each_result_same_alone_and_in_the_mix(tblock):
    from npbench_amd import serve
    print(f"serve mix tblock={tblock}", flush=True)
    reqs = small_mix(24, seed=tblock)
    srv = serve.Server(dtype=torch.float32, tblock=tblock, keep_results=True, max_active=10)
    done = srv.run(reqs, realtime=True)
    assert sorted(c.req.rid for c in done) == list(range(len(reqs)))
    assert srv.peak_active >= 2, "requests never overlapped on the device"
    for c in done:
        A, B = alone(c.req, torch.float32)
        assert torch.equal(c.A, A) and torch.equal(c.B, B), f"request {c.req.rid}"
    srv.close()


def test_mix_matches_the_fp64_oracle_for_fp64_requests():
    from npbench_amd import serve
    from oracle import stencils as ref
    reqs = small_mix(6, seed=11)
    srv = serve.Server(dtype=torch.float64, tblock=2, keep_results=True)
    done = srv.run(reqs, realtime=False)
    for c in done:
        A, B = serve.make_inputs(c.req, torch.float64)
        a, b = A.cpu().numpy(), B.cpu().numpy()
        (ref.jacobi_sweeps if c.req.kind == "jacobi_2d" else ref.heat_sweeps)(a, b, c.req.sweeps)
        assert np.array_equal(c.A.cpu().numpy(), a) and np.array_equal(c.B.cpu().numpy(), b)
    srv.close()


def test_admission_respects_the_budget():
    from npbench_amd import serve
    reqs = small_mix(16, seed=3)
    biggest = max(r.footprint(torch.float32) for r in reqs)
    budget = 2 * biggest
    srv = serve.Server(dtype=torch.float32, tblock=2, hbm_budget=budget)
    done = srv.run(reqs, realtime=False)
    assert len(done) == len(reqs)
    assert 0 < srv.peak_used <= budget
    s = serve.summarize(done, max(c.finished for c in done))
    assert s["p50_ms"] <= s["p99_ms"] and s["glups"] > 0
    srv.close()


def test_oversized_request_is_rejected():
    from npbench_amd import serve
    req = serve.Request(0, "heat_3d", 64, 2, 0.0, 1)
    srv = serve.Server(dtype=torch.float32, tblock=2, hbm_budget=1024)
    with pytest.raises(ValueError):
        srv.run([req], realtime=False)
    srv.close()
