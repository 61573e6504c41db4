"""Batched solves (many independent cluster systems per launch) through the
C ABI: every system's result equals its own single-system solve bit for bit
(the per-system arithmetic and reduction order do not depend on the batch),
and agrees with the CPU oracle's Eigen-style BiCGSTAB iteration counts."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

nm = 1e-9


@pytest.fixture(scope="module")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    import miepy_amd.device as D
    D.require_gpu()
    return D


def _systems(dev, seed=0, count=6):
    from miepy_amd.workloads import random_cluster, random_plane_waves
    from miepy_amd.mie import mie_coefficients
    rng = np.random.default_rng(seed)
    out = []
    for c in range(count):
        N = int(rng.integers(3, 40))
        lmax = int(rng.choice([1, 2, 3, 4]))
        S = int(rng.choice([1, 2, 3, 5]))
        a = 80 * nm
        pos = random_cluster(N, 600 * nm * N ** (1 / 3), a, rng)
        k = 2 * np.pi / (rng.uniform(450, 900) * nm)
        eps = complex(rng.uniform(2, 6), rng.uniform(0, 0.2)) if c % 2 else -15 + 1j
        ms, _ = mie_coefficients(np.full(N, a), np.full(N, eps), np.ones(N), 1.0, 1.0, k, lmax)
        srcs = random_plane_waves(S, rng, random_pol=True)
        b = np.stack([s.structure(pos, k, lmax).reshape(-1) for s in srcs])
        out.append((pos, ms, k, b))
    return out


def test_batched_bicgstab_equals_single(dev, oracle):
    systems = _systems(dev)
    As, bs, As32, bs32 = [], [], [], []
    for pos, ms, k, b in systems:
        As.append(dev.assemble(pos, ms, k, add_identity=True))
        bs.append(dev.to_device(b, dev.C128))
    xs, reps = dev.bicgstab_batched(As, bs, tol=1e-10, maxiter=2000)
    for i, (A, b) in enumerate(zip(As, bs)):
        x1, rep1 = dev.bicgstab(A, b, tol=1e-10, maxiter=2000)
        assert reps[i].iterations == rep1.iterations, (i, reps[i], rep1)
        assert reps[i].converged
        assert np.array_equal(xs[i].cpu().numpy(), x1.cpu().numpy())
        An = A.cpu().numpy()
        for s in range(b.shape[0]):
            bn = b[s].cpu().numpy()
            _, it_o, _ = oracle.bicgstab_eigen(An, bn, tol=1e-10, maxiter=2000)
            assert abs(reps[i].iterations[s] - it_o) <= 1
            r = np.linalg.norm(An @ xs[i][s].cpu().numpy() - bn) / np.linalg.norm(bn)
            assert r <= 1.01e-10


def test_batched_refined_equals_single(dev):
    systems = _systems(dev, seed=1, count=5)
    A32s, A64s, bs = [], [], []
    for pos, ms, k, b in systems:
        A64s.append(dev.assemble(pos, ms, k, add_identity=True))
        A32s.append(dev.assemble(pos, ms, k, dtype=dev.C64, add_identity=True))
        bs.append(dev.to_device(b, dev.C128))
    xs, reps = dev.solve_refined_batched(A32s, A64s, bs, tol=1e-10)
    for i in range(len(bs)):
        assert reps[i].converged, reps[i]
        x1, rep1 = dev.solve_refined(A32s[i], A64s[i], bs[i], tol=1e-10)
        assert np.array_equal(xs[i].cpu().numpy(), x1.cpu().numpy())
        xf, _ = dev.bicgstab(A64s[i], bs[i], tol=1e-12, maxiter=5000)
        err = np.linalg.norm((xs[i] - xf).cpu().numpy()) / np.linalg.norm(xf.cpu().numpy())
        assert err < 1e-8, err


def test_batched_mixed_status(dev):
    """A system forced to stop at maxiter reports it; the others converge."""
    systems = _systems(dev, seed=2, count=3)
    As = [dev.assemble(p, m, k, add_identity=True) for p, m, k, _ in systems]
    bs = [dev.to_device(b, dev.C128) for *_, b in systems]
    xs, reps = dev.bicgstab_batched(As, bs, tol=[1e-10, 1e-16, 1e-10], maxiter=[2000, 3, 2000])
    assert reps[0].converged and reps[2].converged
    assert all(s == 3 for s in reps[1].status) and all(i == 3 for i in reps[1].iterations)


def test_cluster_batch_matches_single_clusters(dev):
    """ClusterBatch (all clusters in one batched solve) vs one sphere_cluster
    per cluster: cross-sections and forces per source within 1e-9 (FP64 and
    matrix-free FP64) and 1e-8 (complex64 + FP64 refinement, stored or with
    the matrix-free operators for lmax <= 3 beside stored lmax-4 systems)."""
    import miepy_amd as M
    from miepy_amd.batch import ClusterBatch
    from miepy_amd.workloads import batch256
    ws = batch256(seed=7, count=5)
    for w in ws:
        w.position = w.position[: min(w.N, 60)]
        w.N = w.position.shape[0]
        w.radius, w.eps = w.radius[: w.N], w.eps[: w.N]
        w.tol = 1e-11
    assert {w.lmax for w in ws} >= {4} and min(w.lmax for w in ws) <= 3
    for precision, rtol in (("fp64", 1e-9), ("mixed", 1e-8), ("matfree", 1e-9), ("mixed_matfree", 1e-8)):
        B = ClusterBatch(ws, precision=precision, matfree_min_n=1000)  # the larger clusters run matrix-free
        if precision in ("fp64", "mixed"):
            tuned = B.autotune(reps=1)  # the batched autotuner's pick serves the solve below
            assert tuned and all(name and ms > 0 for name, ms in tuned.values())
        B.solve()
        assert B.converged()
        res = B.host_results()
        for w, r in zip(ws, res):
            one = M.sphere_cluster(position=w.position, radius=w.radius, material=w.materials(), source=w.sources,
                                   wavelength=w.wavelength, lmax=w.lmax, medium=w.medium(), tol=1e-12)
            C = np.array(one.cross_sections()).T  # [S, 3]
            np.testing.assert_allclose(r["C_total"], C, rtol=rtol, atol=rtol * np.abs(C).max())
            F = one.force()
            assert np.max(np.abs(r["F"] - F)) <= rtol * 10 * np.abs(F).max()


def test_graph_replay_equals_eager(dev):
    """Iterations replayed from captured HIP graphs (the default) give the
    eager launches' results bit for bit, for single and batched solves,
    stored and matrix-free operators, c128 and c64, one or several
    iterations per graph; each solve replays its graph many times with host
    synchronization between replays."""
    import miepy_amd._lib as L
    lib = L.load()
    systems = _systems(dev, seed=3, count=4)
    As = [dev.assemble(pos, ms, k, add_identity=True) for pos, ms, k, _ in systems]
    bs = [dev.to_device(b, dev.C128) for *_, b in systems]
    ops = [dev.ClusterOperator(pos, ms, k, max_rhs=b.shape[0]) for pos, ms, k, b in systems]
    assert lib.mpa_get_option(L.MPA_OPT_GRAPHS) == 1

    def run():
        out = []
        for sync in (1, 3):
            xs, reps = dev.bicgstab_batched(As, bs, tol=1e-10, maxiter=2000, sync_every=sync)
            out += [(x.cpu().numpy(), r.iterations) for x, r in zip(xs, reps)]
            xs, reps = dev.bicgstab_batched(ops, bs, tol=1e-10, maxiter=2000, add_identity=True,
                                            sync_every=sync)
            out += [(x.cpu().numpy(), r.iterations) for x, r in zip(xs, reps)]
        A32 = [A.to(dev.C64) for A in As]
        xs, reps = dev.bicgstab_batched(A32, [b.to(dev.C64) for b in bs], tol=1e-5, maxiter=2000)
        out += [(x.cpu().numpy(), r.iterations) for x, r in zip(xs, reps)]
        x, rep = dev.bicgstab(As[0], bs[0], tol=1e-10)
        out.append((x.cpu().numpy(), rep.iterations))
        return out

    graphed = run()
    try:
        assert lib.mpa_set_option(L.MPA_OPT_GRAPHS, 0) == 0
        eager = run()
    finally:
        lib.mpa_set_option(L.MPA_OPT_GRAPHS, 1)
    assert len(graphed) == len(eager)
    for (xg, ig), (xe, ie) in zip(graphed, eager):
        assert ig == ie
        assert np.array_equal(xg, xe)


def test_observables_batched_equals_single(dev):
    """mpa_cluster_observables_batched over clusters of mixed N, lmax and
    source counts (more than one launch's worth of descriptors) gives each
    cluster exactly the observables of its own mpa_cluster_observables call."""
    import torch
    rng = np.random.default_rng(3)
    clusters = []
    for q in range(37):  # > 32 descriptors: two launches
        N = int(rng.integers(1, 70))
        lmax = int(rng.choice([1, 2, 3, 4, 5]))
        S = int(rng.choice([1, 2, 4]))
        rmax = lmax * (lmax + 2)
        cplx = lambda *shape: dev.to_device(rng.normal(size=shape) + 1j * rng.normal(size=shape), dev.C128)
        clusters.append((cplx(S, N, 2, rmax), cplx(S, N, 2, rmax), cplx(N, 2, lmax), cplx(N, 2, lmax),
                         float(rng.uniform(5e6, 2e7)), float(rng.uniform(1.0, 2.0))))
    got = dev.observables_batched(clusters)
    torch.cuda.synchronize()
    for q, (c, g) in enumerate(zip(clusters, got)):
        one = dev.observables(*c)
        for key in ("p_scat", "p_int", "C_part", "C_total", "F", "T"):
            assert np.array_equal(g[key].cpu().numpy(), one[key].cpu().numpy()), (q, key)


def test_overlapped_product_phases_equal_sequential(dev):
    """MPA_OPT_OVERLAP: the batched GEMM over the stored matrices on the side
    stream beside the matrix-free products gives the same solve, bit for bit,
    as the two one after the other, with the iterations replayed from HIP
    graphs (the fork and join captured as parallel branches) and eagerly."""
    import miepy_amd._lib as L
    from miepy_amd.batch import ClusterBatch
    from miepy_amd.workloads import batch256
    lib = L.load()
    ws = batch256(seed=9, count=6)
    for w in ws:
        w.position = w.position[: min(w.N, 80)]
        w.N = w.position.shape[0]
        w.radius, w.eps = w.radius[: w.N], w.eps[: w.N]
    saved = (lib.mpa_get_option(L.MPA_OPT_OVERLAP), lib.mpa_get_option(L.MPA_OPT_GRAPHS))
    out = {}
    try:
        for overlap, graphs in ((0, 1), (1, 1), (1, 0)):
            lib.mpa_set_option(L.MPA_OPT_OVERLAP, overlap)
            lib.mpa_set_option(L.MPA_OPT_GRAPHS, graphs)
            B = ClusterBatch(ws, precision="mixed_matfree", matfree_min_n=1500)
            assert any(c.op is not None for c in B.clusters) and any(c.op is None for c in B.clusters)
            for _ in range(2):  # a second solve replays what the first recorded
                B.solve()
            assert B.converged()
            out[(overlap, graphs)] = B.host_results()
    finally:
        lib.mpa_set_option(L.MPA_OPT_OVERLAP, saved[0])
        lib.mpa_set_option(L.MPA_OPT_GRAPHS, saved[1])
    ref = out[(0, 1)]
    for key in ((1, 1), (1, 0)):
        for a, b in zip(out[key], ref):
            for k in ("p_inc", "C_total", "F"):
                assert np.array_equal(a[k], b[k]), (key, k)


def test_batched_assembly_equals_single(dev):
    """mpa_sphere_aggregate_tmatrix_batched over clusters of mixed N and lmax
    (1-6) and all three output kinds (c128, c64, c128 + c64 in one pass),
    more than one launch's worth of each kind: every matrix equals its own
    single-cluster assembly bit for bit."""
    import torch
    from miepy_amd.workloads import random_cluster
    from miepy_amd.mie import mie_coefficients
    rng = np.random.default_rng(12)
    items, refs = [], []
    for q in range(70):
        N = int(rng.integers(1, 24))
        lmax = int(rng.integers(1, 7))
        k = 2 * np.pi / (rng.uniform(450, 900) * nm)
        pos = random_cluster(N, 500 * nm * N ** (1 / 3) + 200 * nm, 60 * nm, rng)
        ms, _ = mie_coefficients(np.full(N, 60 * nm), np.full(N, 3.0 + 0.1j), np.ones(N), 1.0, 1.0, k, lmax)
        pos_d, ms_d = dev.to_device(pos, torch.float64), dev.to_device(ms, dev.C128)
        n = 2 * lmax * (lmax + 2) * N
        kind = q % 3
        if kind == 0:
            out, out32 = torch.empty((n, n), dtype=dev.C128, device="cuda"), None
            ref = (dev.assemble(pos_d, ms_d, k, dtype=dev.C128, add_identity=True),)
        elif kind == 1:
            out, out32 = torch.empty((n, n), dtype=dev.C64, device="cuda"), None
            ref = (dev.assemble(pos_d, ms_d, k, dtype=dev.C64, add_identity=True),)
        else:
            out = torch.empty((n, n), dtype=dev.C128, device="cuda")
            out32 = torch.empty((n, n), dtype=dev.C64, device="cuda")
            ref = dev.assemble_dual(pos_d, ms_d, k, add_identity=True)
        items.append((pos_d, ms_d, k, out, out32))
        refs.append(ref)
    dev.assemble_batched(items, add_identity=True)
    torch.cuda.synchronize()
    for q, ((_, _, _, out, out32), ref) in enumerate(zip(items, refs)):
        assert np.array_equal(out.cpu().numpy(), ref[0].cpu().numpy()), q
        if out32 is not None:
            assert np.array_equal(out32.cpu().numpy(), ref[1].cpu().numpy()), q
