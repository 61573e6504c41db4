"""The benchmark configurations at their full sizes, on the GPU.

- rand1000 (1000 random metal spheres, lmax 3, n = 30000, tol 1e-10): the
  FP64 stored path against the CPU oracle's own path (assembly + Eigen-style
  BiCGSTAB, solved to 1e-12 on the host): p_scat and cross-sections within
  1e-9, forces within 1e-8; the FP64 matrix-free path and the complex64 +
  FP64-refinement matrix-free path against the same oracle result
  (1e-9 / 1e-8 / 1e-8 and 1e-8 respectively).
- rand2500 (2500 spheres, lmax 3, n = 75000, 16 sources, refinement to
  1e-10): the stored mixed path and the matrix-free mixed path against the
  FP64 stored path (cross-sections and p_scat within 1e-8 of it), and the
  FP64 matrix-free path against the FP64 stored path.
- batch256 (256 uneven clusters x 8 sources, FP32 + refinement to 1e-8):
  every cluster of the mixed and the mixed matrix-free batches against the
  FP64 batch solved to 1e-12 (cross-sections within 1e-8 of the largest of
  each source, forces within 1e-7 of the largest); the FP64 batch's smallest
  clusters of each lmax against the CPU oracle (1e-9).
- serving: the opening requests of the serving stream (seed 5, 200 req/s,
  N up to 3000, lmax 2-5, 1-32 wavelengths, 1-4 sources) served together by
  the mixed engine; every request converges, and sampled requests give the
  same results alone as in the mix (bit for bit).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    import miepy_amd.device as D
    D.require_gpu()
    return D


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _cs_err(a, b):
    """Largest cross-section difference relative to the largest cross-section of each source."""
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b) / np.abs(b).max(axis=-1, keepdims=True)))


def _solve_paths(dev, w, paths, tol=None):
    """{path: observables on the host} for one workload through bench's step."""
    import torch
    from bench import Problem, Solver
    prob = Problem(w)
    if tol is not None:
        w.tol = tol
    out = {}
    for p in paths:
        s = Solver(prob, p)
        s.step()
        torch.cuda.synchronize()
        assert all(v == 0 for v in getattr(s.report, "status", [0])), (p, s.report)
        out[p] = {k: (v.cpu().numpy() if v is not None else None) for k, v in s.obs.items()}
        out[p]["report"] = s.report
        del s
        torch.cuda.empty_cache()
    return prob, out


def test_rand1000_fp64_against_oracle(dev):
    import oracle.host as H
    from miepy_amd import workloads as W
    w = W.rand1000()
    src = w.sources[0]
    print(f"[configs] rand1000: GPU paths, n = {w.n}", flush=True)
    prob, res = _solve_paths(dev, w, ["fp64", "matfree", "mixed_matfree"])
    print("[configs] rand1000: oracle assembly + BiCGSTAB to 1e-12 on the host", flush=True)
    ref = H.solve_cluster(w.position, w.radius, w.eps, w.wavelength, w.lmax, polarization=src.polarization,
                          theta=src.theta, phi=src.phi, eps_b=w.medium_eps, tol=1e-12, maxiter=5000)
    assert ref["err"] <= 1e-12
    for p, bar_p, bar_c, bar_f in (("fp64", 1e-9, 1e-9, 1e-8), ("matfree", 1e-9, 1e-9, 1e-8),
                                   ("mixed_matfree", 1e-8, 1e-8, 1e-8)):
        o = res[p]
        e_p = _rel(o["p_scat"][0], ref["p_scat"])
        e_c = _cs_err(o["C_total"][0], ref["cross_sections"])
        e_f = float(np.max(np.abs(o["F"][0] - ref["force"])) / np.abs(ref["force"]).max())
        print(f"[configs] rand1000 {p}: p_scat {e_p:.2e}, cross-sections {e_c:.2e}, forces {e_f:.2e}", flush=True)
        assert e_p <= bar_p and e_c <= bar_c and e_f <= bar_f, (p, e_p, e_c, e_f)


def test_rand2500_sixteen_sources(dev):
    from miepy_amd import workloads as W
    w = W.rand2500()
    print(f"[configs] rand2500: n = {w.n}, {len(w.sources)} sources", flush=True)
    _, res = _solve_paths(dev, w, ["fp64", "mixed", "mixed_matfree", "matfree"])
    ref = res["fp64"]
    for p, bar_p, bar_c in (("mixed", 1e-8, 1e-8), ("mixed_matfree", 1e-8, 1e-8), ("matfree", 1e-8, 1e-9)):
        o = res[p]
        e_p = max(_rel(o["p_scat"][s], ref["p_scat"][s]) for s in range(16))
        e_c = _cs_err(o["C_total"], ref["C_total"])
        e_f = float(np.max(np.abs(o["F"] - ref["F"])) / np.abs(ref["F"]).max())
        print(f"[configs] rand2500 {p}: p_scat {e_p:.2e}, cross-sections {e_c:.2e}, forces {e_f:.2e}", flush=True)
        assert e_p <= bar_p and e_c <= bar_c and e_f <= 1e-7, (p, e_p, e_c, e_f)


def test_batch256_every_cluster(dev, oracle):
    import torch
    import oracle.host as H
    from miepy_amd import workloads as W
    from miepy_amd.batch import ClusterBatch
    ws = W.batch256()
    print(f"[configs] batch256: {len(ws)} clusters, FP64 reference batch to 1e-12", flush=True)
    R = ClusterBatch(ws, precision="fp64", tol_scale=1e-4)
    R.solve()
    assert R.converged()
    ref = R.host_results()
    del R
    torch.cuda.empty_cache()
    for prec in ("mixed", "mixed_matfree"):
        B = ClusterBatch(ws, precision=prec)
        B.solve()
        assert B.converged(), prec
        got = B.host_results()
        del B
        torch.cuda.empty_cache()
        worst_c = max(_cs_err(a["C_total"], b["C_total"]) for a, b in zip(got, ref))
        worst_f = max(float(np.max(np.abs(a["F"] - b["F"])) / np.abs(b["F"]).max()) for a, b in zip(got, ref))
        print(f"[configs] batch256 {prec}: worst cross-section error {worst_c:.2e}, forces {worst_f:.2e}", flush=True)
        assert worst_c <= 1e-8 and worst_f <= 1e-7, (prec, worst_c, worst_f)
    # the FP64 reference's smallest cluster of each lmax against the CPU oracle (LU), every source
    for lmax in (2, 3, 4):
        i = min((j for j, w in enumerate(ws) if w.lmax == lmax), key=lambda j: ws[j].n)
        w = ws[i]
        for s, src in enumerate(w.sources):
            o = H.solve_cluster(w.position, w.radius, w.eps, w.wavelength, w.lmax, polarization=src.polarization,
                                theta=src.theta, phi=src.phi, eps_b=w.medium_eps, method="exact")
            e_c = _cs_err(ref[i]["C_total"][s], o["cross_sections"])
            e_p = _rel(ref[i]["p_inc"][s].reshape(-1), o["p_inc"].reshape(-1))
            assert e_c <= 1e-9 and e_p <= 1e-9, (i, s, e_c, e_p)


def _serve_results(r):
    return [(x["cross_sections"], x["force"]) for x in r.results]


def test_serving_stream_opening(dev):
    import torch
    from miepy_amd import serve
    from miepy_amd import workloads as W
    stream = []
    for t, ws in W.serving_requests(seed=5, rate=200.0, duration=60.0, Nmax=3000):
        stream.append((t, ws))
        if len(stream) == 10:
            break
    print(f"[configs] serving: {len(stream)} requests, {sum(len(ws) for _, ws in stream)} systems", flush=True)
    eng = serve.ServingEngine(hbm_budget_bytes=200e9, iterations_per_step=2, precision="mixed")
    reqs = [serve.Request(i, ws, 0.0) for i, (_, ws) in enumerate(stream)]
    for r in reqs:
        eng.submit(r)
    eng.drain()
    print(f"[configs] serving: {eng.solves_done} systems done, {eng.failed_systems} failed, "
          f"{eng.fp64_fallbacks} FP64 safety nets", flush=True)
    assert all(r.status == "done" for r in reqs), [r.status for r in reqs]
    del eng
    torch.cuda.empty_cache()
    for i in (1, 5, 7):  # sampled requests, alone
        one = serve.ServingEngine(hbm_budget_bytes=200e9, iterations_per_step=2, precision="mixed")
        r = serve.Request(0, stream[i][1], 0.0)
        one.submit(r)
        one.drain()
        assert r.status == "done"
        for (ca, fa), (cm, fm) in zip(_serve_results(r), _serve_results(reqs[i])):
            assert np.array_equal(ca, cm) and np.array_equal(fa, fm), i
        del one
        torch.cuda.empty_cache()
