"""Matrix-free operator: (I + T) x recomputed inside every product.

Parity targets: the oracle's assembled T times x (FP64, 1e-12 relative to
|T||x| per row block: the kernel sums the same products in another order),
the device's own assembled matvec, and the oracle's solution of the cluster
system (BiCGSTAB to 1e-10 and LU), for the shapes the dense path covers
(lmax 1-6, several right-hand sides, aligned pairs, single particles).
"""

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


@pytest.fixture(scope="module")
def H(oracle):
    import oracle.host as H
    return H


def _cluster(N, lmax, seed, spacing=350 * nm):
    rng = np.random.default_rng(seed)
    from miepy_amd.workloads import random_cluster
    pos = random_cluster(N, spacing * N ** (1 / 3) * 1.5, 75 * nm, rng)
    mie = rng.normal(size=(N, 2, lmax)) * 0.3 + 1j * rng.normal(size=(N, 2, lmax)) * 0.3
    return pos, mie


def _rel_err(got, ref):
    return np.max(np.abs(got - ref)) / np.max(np.abs(ref))


@pytest.mark.parametrize("N,lmax,S", [(2, 1, 1), (5, 2, 2), (9, 3, 1), (7, 4, 3), (5, 5, 1), (4, 6, 2),
                                      (40, 2, 5), (33, 3, 4)])
def test_matfree_matches_oracle_gemv(dev, oracle, N, lmax, S):
    pos, mie = _cluster(N, lmax, seed=100 + N * 10 + lmax)
    k = 2 * np.pi / (600 * nm)
    T = oracle.sphere_aggregate_tmatrix(pos, mie, k)
    n = T.shape[0]
    rng = np.random.default_rng(N + S)
    X = rng.normal(size=(S, n)) + 1j * rng.normal(size=(S, n))
    op = dev.ClusterOperator(pos, mie, k, max_rhs=S)
    Xd = dev.to_device(X, dev.C128)
    got = dev.matfree_matvec(op, Xd).cpu().numpy()
    got_i = dev.matfree_matvec(op, Xd, add_identity=True).cpu().numpy()
    for s in range(S):
        ref = oracle.gemv(T, X[s])
        assert _rel_err(got[s], ref) <= 1e-12
        assert _rel_err(got_i[s], ref + X[s]) <= 1e-12
    # against the device's assembled matrix and GEMV
    A = dev.assemble(pos, mie, k, add_identity=True)
    dense = dev.matvec(A, Xd).cpu().numpy()
    assert _rel_err(got_i, dense) <= 1e-12


def test_matfree_edge_geometries(dev, oracle):
    k = 2 * np.pi / (700 * nm)
    # single particle: T = 0
    op = dev.ClusterOperator(np.zeros((1, 3)), np.ones((1, 2, 3)), k)
    x = dev.to_device(np.arange(30, dtype=complex).reshape(1, 30), dev.C128)
    assert np.all(dev.matfree_matvec(op, x).cpu().numpy() == 0)
    assert np.all(dev.matfree_matvec(op, x, add_identity=True).cpu().numpy() == x.cpu().numpy())
    # pairs along z (theta = 0, pi) and in the xy plane
    pos = np.array([[0, 0, 0], [0, 0, 300 * nm], [0, 0, -300 * nm], [300 * nm, 0, 0], [-300 * nm, 0, 0],
                    [0, 300 * nm, 0]], float)
    mie = np.full((6, 2, 3), 0.2 + 0.1j)
    T = oracle.sphere_aggregate_tmatrix(pos, mie, k)
    X = np.random.default_rng(1).normal(size=T.shape[0]) + 0j
    got = dev.matfree_matvec(dev.ClusterOperator(pos, mie, k), dev.to_device(X.reshape(1, -1), dev.C128))
    assert _rel_err(got.cpu().numpy()[0], oracle.gemv(T, X)) <= 1e-12


def test_matfree_large_cluster_matches_dense(dev):
    """rand1000-sized chunking (several column chunks per row)."""
    from miepy_amd.workloads import rand1000
    from miepy_amd.mie import mie_coefficients
    w = rand1000()
    ms, _ = mie_coefficients(w.radius, w.eps, np.ones(w.N), w.medium_eps, 1.0, w.k, w.lmax)
    import torch
    g = torch.Generator(device="cuda").manual_seed(0)
    x = torch.randn((2, w.n), dtype=dev.C128, device="cuda", generator=g)
    op = dev.ClusterOperator(w.position, ms, w.k, max_rhs=2)
    got = dev.matfree_matvec(op, x, add_identity=True)
    A = dev.assemble(w.position, ms, w.k, add_identity=True)
    ref = dev.matvec(A, x)
    del A
    err = (got - ref).abs().max().item() / ref.abs().max().item()
    assert err <= 1e-12, err


def test_matfree_bicgstab_matches_oracle(dev, oracle, H):
    """Solve with the matrix-free operator: same iterations as the oracle's
    Eigen BiCGSTAB on the assembled matrix, solution equal to LU."""
    N, lmax = 12, 2
    pos, _ = _cluster(N, lmax, seed=5)
    k = 2 * np.pi / (600 * nm)
    ms, _ = H.mie_coefficients(np.full(N, 75 * nm), np.full(N, -15 + 1j), np.ones(N), 1.0, 1.0, k, lmax)
    T = oracle.sphere_aggregate_tmatrix(pos, ms, k)
    n = T.shape[0]
    rng = np.random.default_rng(9)
    b = rng.normal(size=n) + 1j * rng.normal(size=n)
    A = np.eye(n) + T
    x_lu = oracle.lu_solve(A, b)
    op = dev.ClusterOperator(pos, ms, k)
    for tol in (1e-5, 1e-10):
        _, ito, _ = oracle.bicgstab_eigen(A, b, tol=tol)
        x, rep = dev.bicgstab_matfree(op, dev.to_device(b.reshape(1, n), dev.C128), tol=tol)
        x = x.cpu().numpy()[0]
        assert abs(rep.iterations[0] - ito) <= 1, (rep.iterations, ito)
        assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) <= 1.01 * tol
        if tol <= 1e-10:
            assert np.linalg.norm(x - x_lu) / np.linalg.norm(x_lu) < 100 * tol


def test_matfree_and_dense_in_one_batch(dev, H):
    """A batch mixing stored matrices and matrix-free systems gives each
    system the result it gets alone."""
    k = 2 * np.pi / (600 * nm)
    systems = []
    for seed, (N, lmax, S) in enumerate([(10, 2, 1), (20, 3, 2), (8, 1, 3)]):
        pos, _ = _cluster(N, lmax, seed=seed + 50)
        ms, _ = H.mie_coefficients(np.full(N, 75 * nm), np.full(N, -15 + 1j), np.ones(N), 1.0, 1.0, k, lmax)
        rng = np.random.default_rng(seed)
        n = 2 * lmax * (lmax + 2) * N
        b = dev.to_device(rng.normal(size=(S, n)) + 1j * rng.normal(size=(S, n)), dev.C128)
        systems.append((pos, ms, b))
    ops = [dev.ClusterOperator(p, m, k, max_rhs=b.shape[0]) for p, m, b in systems]
    dense = [dev.assemble(p, m, k) for p, m, _ in systems]  # T only: add_identity adds I
    rhs = [b for _, _, b in systems]
    xs, reps = dev.bicgstab_batched([ops[0], dense[1], ops[2]], rhs, tol=1e-10, add_identity=True)
    assert all(r.converged for r in reps)
    x_dense, _ = dev.bicgstab(dense[1], rhs[1], tol=1e-10, add_identity=True)
    assert np.array_equal(xs[1].cpu().numpy(), x_dense.cpu().numpy())
    for i in (0, 2):
        x_alone, _ = dev.bicgstab_matfree(ops[i], rhs[i], tol=1e-10)
        assert np.array_equal(xs[i].cpu().numpy(), x_alone.reshape(xs[i].shape).cpu().numpy())
        ref = dev.bicgstab(dense[i], rhs[i], tol=1e-10, add_identity=True)[0].cpu().numpy()
        got = xs[i].cpu().numpy()
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-8


def test_matfree_bad_arguments(dev):
    import ctypes
    import miepy_amd._lib as L
    lib = L.load()
    assert lib.mpa_matfree_matvec(None, 1, 0, None, None, None) == L.MPA_ERR_ARG
    assert lib.mpa_matfree_workspace_bytes(10, 7, 1) == 0
    op = L.MpaClusterOperator(4, 9, None, None, 1.0, None)
    assert lib.mpa_matfree_matvec(ctypes.byref(op), 1, 0, None, None, None) == L.MPA_ERR_ARG


@pytest.mark.parametrize("kernel", [0, 1, 2])
@pytest.mark.parametrize("N,lmax,S", [(300, 1, 1), (200, 2, 1), (140, 3, 1), (150, 3, 2), (33, 3, 3)])
def test_factorized_operator_matches_dense(dev, N, lmax, S, kernel):
    """Every FP64 kernel for lmax <= 3 (0: matrix cores, 1: LDS kernel, 2:
    factorized, one lane per row particle, 128+ particles): equal to the
    stored matrix's product to 1e-12, spheres and general T-matrices."""
    import miepy_amd._lib as L
    lib = L.load()
    lib.mpa_set_option(L.MPA_OPT_MATFREE_FP64_KERNEL, kernel)
    try:
        _factorized_vs_dense(dev, N, lmax, S)
    finally:
        lib.mpa_set_option(L.MPA_OPT_MATFREE_FP64_KERNEL, 0)


def _factorized_vs_dense(dev, N, lmax, S):
    pos, mie = _cluster(N, lmax, seed=200 + N + lmax, spacing=300 * nm)
    k = 2 * np.pi / (600 * nm)
    n = 2 * lmax * (lmax + 2) * N
    rng = np.random.default_rng(N)
    X = dev.to_device(rng.normal(size=(S, n)) + 1j * rng.normal(size=(S, n)), dev.C128)
    A = dev.assemble(pos, mie, k, add_identity=True)
    ref = dev.matvec(A, X)
    got = dev.matfree_matvec(dev.ClusterOperator(pos, mie, k, max_rhs=S), X, add_identity=True)
    err = (got - ref).abs().max().item() / ref.abs().max().item()
    assert err <= 1e-12, err
    rmax = lmax * (lmax + 2)
    tm = 0.05 * (rng.normal(size=(N, 2 * rmax, 2 * rmax)) + 1j * rng.normal(size=(N, 2 * rmax, 2 * rmax)))
    At = dev.assemble_tmatrix(pos, tm.reshape(N, 2, rmax, 2, rmax), k, add_identity=True)
    ref = dev.matvec(At, X)
    got = dev.matfree_matvec(dev.ClusterOperator(pos, None, k, max_rhs=S, tmatrix=tm), X, add_identity=True)
    err = (got - ref).abs().max().item() / ref.abs().max().item()
    assert err <= 1e-12, err


@pytest.mark.parametrize("kernel", [0, 1, 2])
def test_factorized_operator_matches_oracle(dev, oracle, kernel):
    import miepy_amd._lib as L
    lib = L.load()
    lib.mpa_set_option(L.MPA_OPT_MATFREE_FP64_KERNEL, kernel)
    try:
        _factorized_vs_oracle(dev, oracle)
    finally:
        lib.mpa_set_option(L.MPA_OPT_MATFREE_FP64_KERNEL, 0)


def _factorized_vs_oracle(dev, oracle):
    N, lmax = 130, 2
    pos, mie = _cluster(N, lmax, seed=9, spacing=300 * nm)
    k = 2 * np.pi / (600 * nm)
    T = oracle.sphere_aggregate_tmatrix(pos, mie, k)
    x = np.random.default_rng(4).normal(size=T.shape[0]) + 0j
    got = dev.matfree_matvec(dev.ClusterOperator(pos, mie, k), dev.to_device(x.reshape(1, -1), dev.C128))
    assert _rel_err(got.cpu().numpy()[0], oracle.gemv(T, x)) <= 1e-12


@pytest.fixture(params=[(0, 4), (0, 6), (1, None)], ids=["fp16split", "fp16shared", "fp32"])
def c64_kernel(request, dev):
    """Selects the complex64 matrix-free kernel (FP16 matrix cores on hi/lo
    split operands: the column-pair kernel, or the kernel whose sources share
    each pair's special functions, with jobs batched per launch; or FP32
    matrix cores) and restores the defaults after."""
    import miepy_amd._lib as L
    lib = L.load()
    option, variant = request.param
    old = lib.mpa_get_option(L.MPA_OPT_MATFREE_C64_KERNEL)
    old_v = lib.mpa_matfree_variant(L.MPA_C64, -1, None)
    lib.mpa_set_option(L.MPA_OPT_MATFREE_C64_KERNEL, option)
    if variant is not None:
        assert lib.mpa_matfree_variant(L.MPA_C64, variant, None) == variant
    yield request.param
    lib.mpa_set_option(L.MPA_OPT_MATFREE_C64_KERNEL, old)
    lib.mpa_matfree_variant(L.MPA_C64, old_v, None)


def _c64_vs_c128(dev, pos, mie, k, S, seed, identity):
    n = 2 * mie.shape[2] * (mie.shape[2] + 2) * pos.shape[0]
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(S, n)) + 1j * rng.normal(size=(S, n))
    op = dev.ClusterOperator(pos, mie, k, max_rhs=S)
    ref = dev.matfree_matvec(op, dev.to_device(X, dev.C128), add_identity=identity).cpu().numpy()
    got = dev.matfree_matvec(op, dev.to_device(X, dev.C64), add_identity=identity).cpu().numpy()
    assert np.all(np.isfinite(got))
    return np.max(np.abs(got - ref)) / np.max(np.abs(ref))


@pytest.mark.parametrize("N,lmax,S", [(7, 1, 1), (20, 2, 2), (45, 3, 1), (300, 3, 1), (160, 2, 3), (130, 1, 2),
                                     (9, 4, 1), (150, 4, 2)])
def test_c64_matrix_core_operator(dev, c64_kernel, N, lmax, S):
    """The complex64 operator (special functions in FP64/FP32, FP16-split or
    FP32 matrix cores) equals the FP64 operator to FP32 accuracy: 2e-6
    relative to the largest entry of (I + T) x, 5e-6 for T x alone."""
    pos, mie = _cluster(N, lmax, seed=300 + N, spacing=300 * nm)
    k = 2 * np.pi / (600 * nm)
    assert _c64_vs_c128(dev, pos, mie, k, S, N + 1, True) <= 2e-6
    assert _c64_vs_c128(dev, pos, mie, k, S, N + 2, False) <= 5e-6


@pytest.mark.parametrize("lmax", [3, 4])
@pytest.mark.parametrize("kd_near,kd_far", [(0.05, 5.0), (2.0, 2000.0)])
def test_c64_operator_extreme_separations(dev, c64_kernel, kd_near, kd_far, lmax):
    """Pairs far inside the near field (|h_6(kd)| ~ 1e15 at kd = 0.05, far
    beyond the FP16 range before scaling) next to well separated pairs: the
    per-pair and per-column scaling keeps the FP16-split operator at FP32
    accuracy (lmax 3 and 4; lmax 4 runs the FP16-split kernel whatever the
    kernel option)."""
    k = 2 * np.pi / (600 * nm)
    rng = np.random.default_rng(11)
    centers = rng.normal(size=(6, 3))
    centers *= kd_far / k / np.max(np.linalg.norm(centers[:, None] - centers[None], axis=-1))
    offs = np.array([[0, 0, 0], [kd_near / k, 0, 0], [0, kd_near / k, kd_near / k]])
    pos = (centers[:, None, :] + offs[None]).reshape(-1, 3)
    mie = (rng.normal(size=(len(pos), 2, lmax)) + 1j * rng.normal(size=(len(pos), 2, lmax))) * 1e-3
    assert _c64_vs_c128(dev, pos, mie, k, 1, 5, False) <= 5e-6


@pytest.mark.parametrize("lmax", [1, 2, 3, 4])
def test_matfree_every_variant(dev, lmax):
    """Each matrix-core kernel variant of the matrix-free product (accumulator
    chains, interleaved special functions; FP64: the slot-mapped and the
    per-lane-select kernels) against the variant-0 product: c64 to 2e-6 of
    the largest entry, c128 to 1e-13."""
    import ctypes
    import miepy_amd._lib as L
    lib = L.load()
    pos, mie = _cluster(170, lmax, seed=23, spacing=300 * nm)
    k = 2 * np.pi / (600 * nm)
    op = dev.ClusterOperator(pos, mie, k, max_rhs=2)
    rng = np.random.default_rng(8)
    X = dev.to_device(rng.normal(size=(2, op.n)) + 1j * rng.normal(size=(2, op.n)), dev.C128)
    for code, dt, tol in ((L.MPA_C64, dev.C64, 2e-6), (L.MPA_C128, dev.C128, 1e-13)):
        cnt = ctypes.c_int(0)
        saved = lib.mpa_matfree_variant(code, -1, ctypes.byref(cnt))
        assert cnt.value >= 2
        try:
            lib.mpa_matfree_variant(code, 0, None)
            ref = dev.matfree_matvec(op, X.to(dt)).cpu().numpy()
            for v in range(1, cnt.value):
                assert lib.mpa_matfree_variant(code, v, None) == v
                got = dev.matfree_matvec(op, X.to(dt)).cpu().numpy()
                assert np.max(np.abs(got - ref)) <= tol * np.max(np.abs(ref)), (code, v)
        finally:
            lib.mpa_matfree_variant(code, saved, None)
        assert lib.mpa_matfree_variant(code, cnt.value, None) == -1


def test_matfree_autotune_keeps_results(dev):
    """Every accumulator-chain variant of the matrix-core kernels gives the
    same product to rounding; the autotuner keeps one of them."""
    N, lmax = 200, 3
    pos, mie = _cluster(N, lmax, seed=21, spacing=300 * nm)
    k = 2 * np.pi / (600 * nm)
    op = dev.ClusterOperator(pos, mie, k, max_rhs=2)
    n = op.n
    rng = np.random.default_rng(3)
    X = dev.to_device(rng.normal(size=(2, n)) + 1j * rng.normal(size=(2, n)), dev.C128)
    ref64 = dev.matfree_matvec(op, X).cpu().numpy()
    ref32 = dev.matfree_matvec(op, X.to(dev.C64)).cpu().numpy()
    for dt, ref, tol in ((dev.C64, ref32, 1e-5), (dev.C128, ref64, 1e-13)):
        v, ms = dev.autotune_matfree(op, 2, dt, reps=2)
        assert v >= 0 and ms > 0
        got = dev.matfree_matvec(op, X.to(dt)).cpu().numpy()
        assert np.max(np.abs(got - ref)) <= tol * np.max(np.abs(ref))


def test_matrix_free_mixed_precision_solve(dev, H):
    """FP32 matrix-core corrections + FP64 matrix-free refinement reach the
    FP64 solution (tol 1e-10 -> 1e-8 relative), alone and batched with a
    stored-matrix system."""
    from miepy_amd.workloads import rand1000
    w = rand1000()
    ms, _ = H.mie_coefficients(w.radius, w.eps, np.ones(w.N), w.medium_eps, 1.0, w.k, w.lmax)
    b = dev.to_device(np.stack([s.structure(w.position, w.k, w.lmax) for s in w.sources]).reshape(1, -1), dev.C128)
    op = dev.ClusterOperator(w.position, ms, w.k)
    x64, rep64 = dev.bicgstab_matfree(op, b, tol=1e-10)
    x, rep = dev.solve_refined_matfree(op, b, tol=1e-10)
    assert rep.converged, rep
    x64, x = x64.cpu().numpy(), x.cpu().numpy()
    assert np.linalg.norm(x - x64) / np.linalg.norm(x64) <= 1e-8
    # batched with a stored system
    N2, lmax2 = 12, 2
    pos2, _ = _cluster(N2, lmax2, seed=3)
    ms2, _ = H.mie_coefficients(np.full(N2, 75 * nm), np.full(N2, 2.5 + 0.1j), np.ones(N2), 1.0, 1.0, w.k, lmax2)
    A64, A32 = dev.assemble_dual(pos2, ms2, w.k, add_identity=True)
    b2 = dev.to_device(np.random.default_rng(1).normal(size=(1, A64.shape[0])) + 0j, dev.C128)
    xs, reps = dev.solve_refined_batched([op, A32], [None, A64], [b, b2], tol=1e-10)
    assert all(r.converged for r in reps), reps
    assert np.array_equal(xs[0].cpu().numpy(), x)
    ref2 = np.linalg.solve(A64.cpu().numpy(), b2.cpu().numpy()[0])
    assert np.linalg.norm(xs[1].cpu().numpy()[0] - ref2) / np.linalg.norm(ref2) <= 1e-8


@pytest.mark.parametrize("N,lmax", [(1, 1), (1, 3), (2, 2), (2, 4), (3, 3)])
def test_c64_operator_tiny_clusters(dev, c64_kernel, N, lmax):
    """One particle (T = 0: the product is exactly 0, (I + T) x = x in
    complex64), two and three particles (every chunk holds a self pair)."""
    pos, mie = _cluster(N, lmax, seed=40 + N, spacing=300 * nm)
    k = 2 * np.pi / (600 * nm)
    n = 2 * lmax * (lmax + 2) * N
    rng = np.random.default_rng(N)
    X = rng.normal(size=(2, n)) + 1j * rng.normal(size=(2, n))
    op = dev.ClusterOperator(pos, mie, k, max_rhs=2)
    got = dev.matfree_matvec(op, dev.to_device(X, dev.C64), add_identity=False).cpu().numpy()
    ref = dev.matfree_matvec(op, dev.to_device(X, dev.C128), add_identity=False).cpu().numpy()
    assert np.all(np.isfinite(got))
    if N == 1:
        assert np.all(got == 0) and np.all(ref == 0)
        gi = dev.matfree_matvec(op, dev.to_device(X, dev.C64), add_identity=True).cpu().numpy()
        assert np.array_equal(gi, X.astype(np.complex64))
    else:
        assert np.max(np.abs(got - ref)) <= 5e-6 * np.max(np.abs(ref))


def _physical_cluster(N, lmax, seed):
    """Spheres of radius 75 nm, eps 2.5 + 0.1j, in vacuum at 600 nm, with their
    Mie coefficients (random coefficients can make I + T nearly singular)."""
    from miepy_amd.mie import mie_coefficients
    from miepy_amd.workloads import random_cluster
    rng = np.random.default_rng(seed)
    pos = random_cluster(N, 300 * nm * N ** (1 / 3) * 1.5, 75 * nm, rng)
    k = 2 * np.pi / (600 * nm)
    mie, _ = mie_coefficients(np.full(N, 75 * nm), np.full(N, 2.5 + 0.1j), np.ones(N), 1.0, 1.0, k, lmax)
    return pos, mie, k


@pytest.mark.parametrize("N,lmax", [(1, 2), (2, 3), (2, 4), (5, 4)])
def test_mixed_matfree_solve_tiny_clusters(dev, N, lmax):
    """The matrix-free mixed solve (complex64 corrections, FP64 residuals) on
    one to five particles equals the FP64 matrix-free BiCGSTAB (1e-9)."""
    pos, mie, k = _physical_cluster(N, lmax, 50 + N)
    n = 2 * lmax * (lmax + 2) * N
    b = dev.to_device(np.random.default_rng(5).normal(size=(1, n)) + 0j, dev.C128)
    op = dev.ClusterOperator(pos, mie, k)
    x64, r64 = dev.bicgstab_matfree(op, b, tol=1e-12, maxiter=5000)
    assert r64.converged, r64
    x, rep = dev.solve_refined_matfree(op, b, tol=1e-11)
    assert rep.converged and not rep.fp64_fallback, rep
    assert np.max(np.abs(x.cpu().numpy() - x64.cpu().numpy())) <= 1e-9 * np.max(np.abs(x64.cpu().numpy()))


def test_refinement_failure_keeps_best_iterate_and_falls_back(dev):
    """A refinement that cannot reach its tolerance (inner solves cut at 2
    iterations, 2 corrections) reports "not converged" and keeps an iterate no
    worse than x = 0; with the FP64 safety net (the default) the system is
    re-solved with FP64 BiCGSTAB and matches it."""
    pos, mie, k = _physical_cluster(40, 3, 9)
    n = 2 * 15 * 40
    b = dev.to_device(np.random.default_rng(6).normal(size=(1, n)) + 0j, dev.C128)
    op = dev.ClusterOperator(pos, mie, k)
    raw, rep0 = dev.solve_refined_matfree(op, b, tol=1e-11, maxiter=2, max_refine=2, fallback=False)
    assert not rep0.converged and not rep0.fp64_fallback
    y = dev.matfree_matvec(op, raw.reshape(1, -1), add_identity=True).cpu().numpy()
    bn = np.linalg.norm(b.cpu().numpy())
    assert np.linalg.norm(b.cpu().numpy() - y) <= bn
    x, rep = dev.solve_refined_matfree(op, b, tol=1e-11, maxiter=2, max_refine=2)
    assert rep.converged and rep.fp64_fallback, rep
    x64, _ = dev.bicgstab_matfree(op, b, tol=1e-11, maxiter=5000)
    assert np.max(np.abs(x.cpu().numpy() - x64.cpu().numpy())) <= 1e-12 * np.max(np.abs(x64.cpu().numpy()))
