"""Serving engine: continuous batching of cluster-solve requests.

Each request's result is the same when it is served alone and when it is
served in a mix with other requests that arrive while it is on the GPU
(bit for bit: a system's arithmetic does not depend on the batch), requests
wait for memory under a tight HBM budget and still complete, and a system
larger than the budget is rejected instead of blocking the queue."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def serve():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    import miepy_amd.serve as S
    return S


def _requests(seed, count, Nmax=120):
    from miepy_amd import workloads as W
    gen = W.serving_requests(seed=seed, rate=50.0, duration=1e9, Nmax=Nmax)
    out = []
    for _ in range(count):
        t, ws = next(gen)
        out.append((t, ws[:3]))  # at most 3 wavelengths per request keeps the test short
    return out


def _results(req):
    return [(r["cross_sections"], r["force"]) for r in req.results]


def test_alone_equals_mixed(serve):
    reqs = _requests(11, 6)
    alone = []
    for t, ws in reqs:
        eng = serve.ServingEngine(hbm_budget_bytes=50e9, iterations_per_step=3)
        r = serve.Request(0, ws, 0.0)
        eng.submit(r)
        eng.drain()
        assert r.status == "done"
        alone.append(_results(r))
    eng = serve.ServingEngine(hbm_budget_bytes=50e9, iterations_per_step=1)
    mixed = []
    for i, (t, ws) in enumerate(reqs):
        r = serve.Request(i, ws, 0.0)
        eng.submit(r)
        mixed.append(r)
        eng.step()  # later requests join while earlier ones are mid-solve
    eng.drain()
    for a, r in zip(alone, mixed):
        assert r.status == "done"
        for (ca, fa), (cm, fm) in zip(a, _results(r)):
            assert np.array_equal(ca, cm) and np.array_equal(fa, fm)
            np.testing.assert_allclose(cm, ca, rtol=1e-12)  # the stated tolerance holds a fortiori


def test_admission_by_footprint_and_rejection(serve):
    from miepy_amd import workloads as W
    reqs = _requests(12, 5)
    fp = max(serve.system_footprint(w) for _, ws in reqs for w in ws)
    eng = serve.ServingEngine(hbm_budget_bytes=1.5 * fp, iterations_per_step=2)
    rs = []
    for i, (t, ws) in enumerate(reqs):
        r = serve.Request(i, ws, 0.0)
        eng.submit(r)
        rs.append(r)
    peak = 0
    while eng.queue or eng.active:
        eng.step()
        peak = max(peak, eng.used)
        assert eng.used <= eng.budget
    assert all(r.status == "done" for r in rs)
    big = W.scaled_random(400, 5, seed=1)  # n = 28000: 12.5 GB matrix > 1.5 x the small footprint
    rb = serve.Request(99, [big], 0.0)
    eng.submit(rb)
    assert rb.status == "too_large" and not eng.queue
