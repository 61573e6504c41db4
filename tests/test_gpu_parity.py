"""GPU parity: the HIP path through the C ABI vs the CPU FP64 oracle.

Tolerances (FP64 path): matrix entries 1e-13 relative to the largest entry
of the block row (the kernel and the oracle evaluate the same FP64
recurrences in a different order); p_scat and cross-sections 1e-9 relative;
forces 1e-8 relative to the largest force.  Mixed precision (complex64
corrections + FP64 refinement to 1e-10): 1e-8 relative against FP64.
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


def _cluster(N, lmax, seed, spacing=350 * nm, eps=-15 + 1j):
    rng = np.random.default_rng(seed)
    from miepy_amd.workloads import random_cluster
    pos = random_cluster(N, spacing * N ** (1 / 3) * 1.5, 75 * nm, rng)
    mie = rng.normal(size=(N, 2, lmax)) * 0.3 + 1j * rng.normal(size=(N, 2, lmax)) * 0.3
    return pos, mie


@pytest.mark.parametrize("N,lmax", [(2, 1), (5, 2), (9, 3), (7, 4), (5, 5), (4, 6), (37, 2)])
def test_assembly_matches_oracle(dev, oracle, N, lmax):
    pos, mie = _cluster(N, lmax, seed=N * 10 + lmax)
    k = 2 * np.pi / (600 * nm)
    ref = oracle.sphere_aggregate_tmatrix(pos, mie, k)
    got = dev.assemble(pos, mie, k).cpu().numpy()
    scale = np.abs(ref).max()
    assert np.max(np.abs(got - ref)) <= 1e-13 * scale
    got_i = dev.assemble(pos, mie, k, add_identity=True).cpu().numpy()
    assert np.max(np.abs(got_i - ref - np.eye(ref.shape[0]))) <= 1e-13 * scale
    got32 = dev.assemble(pos, mie, k, dtype=dev.C64).cpu().numpy()
    assert np.max(np.abs(got32 - ref)) <= 2e-7 * scale


def test_assembly_hex37(dev, oracle, H):
    from miepy_amd.workloads import hex37
    w = hex37()
    ms, _ = H.mie_coefficients(w.radius, w.eps, np.ones(w.N), w.medium_eps, 1.0, w.k, w.lmax)
    ref = oracle.sphere_aggregate_tmatrix(w.position, ms, w.k)
    got = dev.assemble(w.position, ms, w.k).cpu().numpy()
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.abs(ref).max()


def test_single_particle_and_aligned_pairs(dev, oracle):
    k = 2 * np.pi / (700 * nm)
    mie = np.ones((1, 2, 3), complex)
    assert np.all(dev.assemble(np.zeros((1, 3)), mie, k).cpu().numpy() == 0)
    # pairs along z (theta = 0 / pi) and in the xy plane (phi edge cases)
    pos = np.array([[0, 0, 0], [0, 0, 300 * nm], [0, 0, -300 * nm], [300 * nm, 0, 0], [-300 * nm, 0, 0],
                    [0, 300 * nm, 0]], float)
    mie = np.full((6, 2, 3), 0.2 + 0.1j)
    ref = oracle.sphere_aggregate_tmatrix(pos, mie, k)
    got = dev.assemble(pos, mie, k).cpu().numpy()
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("S", [1, 2, 3, 5, 8, 11, 16, 21])
def test_matvec(dev, S):
    """Every tile class: vector-ALU tiles (S <= 4) and the matrix-core GEMM
    (S >= 5, 16 RHS per pass, K steps of 32 columns with a ragged tail,
    partial row blocks)."""
    rng = np.random.default_rng(S)
    for n in (30, 64, 130, 600, 1000):
        A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        x = rng.normal(size=(S, n)) + 1j * rng.normal(size=(S, n))
        y = dev.matvec(dev.to_device(A, dev.C128), dev.to_device(x, dev.C128)).cpu().numpy()
        ref = x @ A.T
        assert np.max(np.abs(y - ref)) <= 1e-12 * np.abs(ref).max()
        yi = dev.matvec(dev.to_device(A, dev.C128), dev.to_device(x, dev.C128), add_identity=True).cpu().numpy()
        assert np.max(np.abs(yi - ref - x)) <= 1e-12 * np.abs(ref).max()
        y32 = dev.matvec(dev.to_device(A, dev.C64), dev.to_device(x, dev.C64)).cpu().numpy()
        assert np.max(np.abs(y32 - ref)) <= 1e-5 * np.abs(ref).max()


def test_bicgstab_matches_oracle_iterations(dev, oracle, H):
    """Same algorithm as the oracle's Eigen restatement: same iteration count,
    solution equal to LU within the tolerance."""
    pos, _ = _cluster(12, 2, seed=5)
    rng = np.random.default_rng(9)
    k = 2 * np.pi / (600 * nm)
    ms, _ = H.mie_coefficients(np.full(12, 75 * nm), np.full(12, -15 + 1j), np.ones(12), 1.0, 1.0, k, 2)
    T = oracle.sphere_aggregate_tmatrix(pos, ms, k)
    n = T.shape[0]
    b = rng.normal(size=n) + 1j * rng.normal(size=n)
    A = np.eye(n) + T
    x_lu = oracle.lu_solve(A, b)
    for tol in (1e-5, 1e-10, 1e-12):
        xo, ito, erro = oracle.bicgstab_eigen(A, b, tol=tol)
        Ad = dev.assemble(pos, ms, k, add_identity=True)
        x, rep = dev.bicgstab(Ad, dev.to_device(b.reshape(1, n), dev.C128), tol=tol)
        x = x.cpu().numpy()[0]
        assert abs(rep.iterations[0] - ito) <= 1, (rep.iterations, ito)
        assert rep.residual[0] <= tol
        assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) <= 1.01 * tol
        if tol <= 1e-10:
            assert np.linalg.norm(x - x_lu) / np.linalg.norm(x_lu) < 100 * tol


def test_bicgstab_multi_rhs_and_edge_cases(dev, oracle):
    rng = np.random.default_rng(3)
    n = 300
    # spectral radius of the perturbation ~ 0.02 sqrt(2n) < 1: well conditioned
    A = np.eye(n) + 0.02 * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    B = rng.normal(size=(5, n)) + 1j * rng.normal(size=(5, n))
    B[2] = 0  # zero right-hand side: x = 0 after 0 iterations
    Ad = dev.to_device(A, dev.C128)
    x, rep = dev.bicgstab(Ad, dev.to_device(B, dev.C128), tol=1e-11)
    assert rep.converged, rep
    x = x.cpu().numpy()
    for s in range(5):
        if s == 2:
            assert rep.iterations[s] == 0 and np.all(x[s] == 0)
            continue
        ref = np.linalg.solve(A, B[s])
        assert np.linalg.norm(x[s] - ref) / np.linalg.norm(ref) < 1e-9
        xs, its, _ = oracle.bicgstab_eigen(A, B[s], tol=1e-11)
        assert abs(rep.iterations[s] - its) <= 1
    # forced non-convergence is reported, not silent
    x, rep = dev.bicgstab(Ad, dev.to_device(B[:1], dev.C128), tol=1e-15, maxiter=2)
    assert rep.iterations[0] == 2 and rep.status[0] == 3 and not rep.converged


def test_solve_refined_matches_fp64(dev):
    rng = np.random.default_rng(4)
    n = 400
    A = np.eye(n) + 0.015 * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    B = rng.normal(size=(3, n)) + 1j * rng.normal(size=(3, n))
    x, rep = dev.solve_refined(dev.to_device(A, dev.C64), dev.to_device(A, dev.C128), dev.to_device(B, dev.C128),
                               tol=1e-12)
    x = x.cpu().numpy()
    assert rep.converged, rep
    for s in range(3):
        ref = np.linalg.solve(A, B[s])
        assert np.linalg.norm(x[s] - ref) / np.linalg.norm(ref) < 1e-10


def test_observables_match_oracle(dev, oracle):
    rng = np.random.default_rng(8)
    N, lmax, S = 7, 3, 2
    rmax = lmax * (lmax + 2)
    k = 2 * np.pi / (700 * nm)
    p_inc = rng.normal(size=(S, N, 2, rmax)) + 1j * rng.normal(size=(S, N, 2, rmax))
    p_src = rng.normal(size=(S, N, 2, rmax)) + 1j * rng.normal(size=(S, N, 2, rmax))
    ms = rng.normal(size=(N, 2, lmax)) + 1j * rng.normal(size=(N, 2, lmax))
    mi = rng.normal(size=(N, 2, lmax)) + 1j * rng.normal(size=(N, 2, lmax))
    eps_b = 1.7689
    o = dev.observables(*[dev.to_device(a, dev.C128) for a in (p_inc, p_src, ms, mi)], k, eps_b)
    o = {key: v.cpu().numpy() for key, v in o.items()}
    from oracle.host import n_index
    nidx = n_index(lmax)
    for s in range(S):
        ps = p_inc[s] * ms[:, :, nidx]
        np.testing.assert_allclose(o["p_scat"][s], ps, rtol=1e-15, atol=0)
        np.testing.assert_allclose(o["p_int"][s], p_inc[s] * mi[:, ::-1, nidx], rtol=1e-15, atol=0)
        tot = np.zeros(3)
        for i in range(N):
            C = oracle.particle_cross_sections(ps[i], p_inc[s, i], p_src[s, i], k)
            for c in range(3):
                np.testing.assert_allclose(o["C_part"][s, i, c], C[c], rtol=1e-13, atol=1e-13 * np.abs(C[c]).max())
            tot += [np.sum(c) for c in C]
        np.testing.assert_allclose(o["C_total"][s], tot, rtol=1e-12)
        F = oracle.force_all(ps, p_inc[s], k, eps_b)
        T = oracle.torque_all(ps, p_inc[s], k, eps_b)
        np.testing.assert_allclose(o["F"][s], F, rtol=1e-12, atol=1e-13 * np.abs(F).max())
        np.testing.assert_allclose(o["T"][s], T, rtol=1e-12, atol=1e-13 * np.abs(T).max())


def _oracle_cluster(H, w, method="bicgstab", tol=None, src=0):
    s = w.sources[src]
    return H.solve_cluster(w.position, w.radius, w.eps, w.wavelength, w.lmax, polarization=s.polarization,
                           theta=s.theta, phi=s.phi, eps_b=w.medium_eps, method=method,
                           tol=w.tol if tol is None else tol, maxiter=100000)


def _gpu_cluster(w, precision="fp64", sources=None, **kw):
    import miepy_amd as M
    src = w.sources if sources is None else sources
    return M.sphere_cluster(position=w.position, radius=w.radius, material=w.materials(), source=src,
                            wavelength=w.wavelength, lmax=w.lmax, medium=w.medium(), precision=precision,
                            tol=w.tol, **kw)


def test_end_to_end_hex37(dev, H):
    """The 37-sphere hexagonal cluster (lmax 2, tol 1e-12): p_scat, cross-
    sections within 1e-9 of the oracle's exact (LU) solve, forces 1e-8."""
    from miepy_amd.workloads import hex37
    w = hex37()
    ref = _oracle_cluster(H, w, method="exact")
    c = _gpu_cluster(w, sources=w.sources[0])
    err = np.linalg.norm(c.p_scat - ref["p_scat"]) / np.linalg.norm(ref["p_scat"])
    assert err < 1e-9, err
    C, Cr = np.array(c.cross_sections()), ref["cross_sections"]
    # scattering and extinction to 1e-9; absorption of these lossless spheres is
    # roundoff (~1e-30), compared on the extinction scale
    np.testing.assert_allclose(C[[0, 2]], Cr[[0, 2]], rtol=1e-9)
    assert abs(C[1] - Cr[1]) < 1e-9 * Cr[2]
    Fs = np.abs(ref["force"]).max()
    assert np.max(np.abs(c.force() - ref["force"])) < 1e-8 * Fs
    Ts = np.abs(ref["torque"]).max()
    assert np.max(np.abs(c.torque() - ref["torque"])) < 1e-8 * max(Ts, 1e-300) or Ts < 1e-40
    # per-particle cross sections
    np.testing.assert_allclose([c.cross_sections_of_particle(i) for i in range(w.N)], ref["per_particle"],
                               rtol=1e-8, atol=1e-9 * np.abs(ref["per_particle"]).max())
    # lossless spheres: no absorption
    assert abs(c.cross_sections().absorption) < 1e-9 * c.cross_sections().extinction


@pytest.mark.parametrize("precision", ["mixed", "matfree", "mixed_matfree"])
def test_end_to_end_mixed_matches_fp64(dev, H, precision):
    """sphere_cluster with complex64 + FP64 refinement (stored or matrix-free)
    and the FP64 matrix-free solve against the stored FP64 solve (1e-8) and
    the oracle's LU solve (1e-9 for FP64)."""
    from miepy_amd.workloads import scaled_random
    w = scaled_random(60, 3, seed=1, tol=1e-10)
    f64 = _gpu_cluster(w, sources=w.sources[0])
    mix = _gpu_cluster(w, precision=precision, sources=w.sources[0])
    assert mix.report.converged, mix.report
    err = np.linalg.norm(mix.p_scat - f64.p_scat) / np.linalg.norm(f64.p_scat)
    assert err < 1e-8, err
    np.testing.assert_allclose(mix.cross_sections(), f64.cross_sections(), rtol=1e-8)
    assert np.max(np.abs(mix.force() - f64.force())) <= 1e-8 * np.abs(f64.force()).max()
    ref = _oracle_cluster(H, w, method="exact")
    np.testing.assert_allclose(f64.cross_sections(), ref["cross_sections"], rtol=1e-9)


def test_multi_source_equals_single(dev):
    from miepy_amd.workloads import scaled_random, random_plane_waves
    w = scaled_random(30, 2, seed=2, tol=1e-11)
    srcs = random_plane_waves(3, np.random.default_rng(0), random_pol=True)
    multi = _gpu_cluster(w, sources=srcs)
    for s, src in enumerate(srcs):
        one = _gpu_cluster(w, sources=src)
        np.testing.assert_allclose(multi.p_scat[s], one.p_scat, rtol=1e-9, atol=1e-12 * np.abs(one.p_scat).max())
        np.testing.assert_allclose(np.array(multi.cross_sections())[:, s], one.cross_sections(), rtol=1e-9)


def test_backend_dropin_functions(dev, oracle, H):
    """The reference's dispatch functions, same signatures, same results."""
    import miepy_amd.backend as B
    pos, _ = _cluster(6, 2, seed=11)
    k = 2 * np.pi / (650 * nm)
    ms, _ = H.mie_coefficients(np.full(6, 75 * nm), np.full(6, 4 + 0.1j), np.ones(6), 1.0, 1.0, k, 2)
    T = B.sphere_aggregate_tmatrix(pos, ms, k)
    assert T.shape == (6, 2, 8, 6, 2, 8)
    Tn = np.asarray(T.reshape(96, 96))
    ref = oracle.sphere_aggregate_tmatrix(pos, ms, k)
    assert np.max(np.abs(Tn - ref)) <= 1e-13 * np.abs(ref).max()
    p_src = H.plane_wave_structure(pos, k, 2, [1, 0])
    p_inc = B.solve_linear_system(T, p_src, B.solver.exact)
    xr, _, _ = oracle.solve_linear_system(ref, p_src, method="exact")
    np.testing.assert_allclose(p_inc.reshape(-1), xr, rtol=1e-10, atol=1e-12 * np.abs(xr).max())
    from oracle.host import n_index
    ps = p_inc * ms[:, :, n_index(2)]
    C = B.particle_cross_sections(ps[0], p_inc[0], p_src[0], k)
    Cr = oracle.particle_cross_sections(ps[0], p_inc[0], p_src[0], k)
    for a, b in zip(C, Cr):
        np.testing.assert_allclose(a, b, rtol=1e-13, atol=1e-13 * np.abs(b).max())
    F = B.force_all(ps, p_inc, k, 1.0, 1.0)
    np.testing.assert_allclose(F, oracle.force_all(ps, p_inc, k), rtol=1e-12, atol=1e-14 * np.abs(F).max())
    Tq = B.torque_all(ps, p_inc, k, 1.0, 1.0)
    np.testing.assert_allclose(Tq, oracle.torque_all(ps, p_inc, k), rtol=1e-12, atol=1e-14 * np.abs(Tq).max())


def test_dual_assembly_equals_single(dev, oracle):
    """One pass writing c128 and c64: the c128 output equals the c128-only
    assembly bit for bit, the c64 output is its rounding."""
    pos, mie = _cluster(11, 3, seed=21)
    k = 2 * np.pi / (640 * nm)
    a64 = dev.assemble(pos, mie, k, add_identity=True).cpu().numpy()
    d64, d32 = dev.assemble_dual(pos, mie, k, add_identity=True)
    assert np.array_equal(d64.cpu().numpy(), a64)
    assert np.array_equal(d32.cpu().numpy(), a64.astype(np.complex64))


def test_every_matvec_variant_and_autotune(dev):
    """Each tile variant the autotuner may pick gives the product (c128 1e-12,
    c64 1e-5 relative), and the autotuner leaves a valid variant in place."""
    import miepy_amd._lib as L
    lib = L.load()
    rng = np.random.default_rng(7)
    n = 610
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    for dtype, code, tol in ((dev.C128, L.MPA_C128, 1e-12), (dev.C64, L.MPA_C64, 1e-5)):
        Ad = dev.to_device(A, dtype)
        for S in (1, 2, 4, 8, 9, 16):
            x = rng.normal(size=(S, n)) + 1j * rng.normal(size=(S, n))
            ref = x @ A.T
            xd = dev.to_device(x, dtype)
            saved = lib.mpa_matvec_variant(code, S, -1)
            v = 0
            while lib.mpa_matvec_variant_name(code, S, v):
                assert lib.mpa_matvec_variant(code, S, v) == v
                y = dev.matvec(Ad, xd, add_identity=True).cpu().numpy()
                err = np.max(np.abs(y - ref - x)) / np.abs(ref).max()
                assert err <= tol, (dtype, S, lib.mpa_matvec_variant_name(code, S, v).decode(), err)
                v += 1
            assert v >= 2
            lib.mpa_matvec_variant(code, S, saved)
            name, ms = dev.autotune_matvec(Ad, S, reps=2)
            assert name and ms > 0
            y = dev.matvec(Ad, xd).cpu().numpy()
            assert np.max(np.abs(y - ref)) <= tol * np.abs(ref).max()


def test_batched_matvec_autotune(dev):
    """The batched autotuner times the variants of the class of nrhs[0] on one
    launch over several systems (one of another class left out) and leaves a
    variant in place whose products are right."""
    import ctypes
    import miepy_amd._lib as L
    lib = L.load()
    rng = np.random.default_rng(8)
    for dtype, code, tol in ((dev.C128, L.MPA_C128, 1e-12), (dev.C64, L.MPA_C64, 1e-5)):
        sizes = (64, 250, 610)
        A = [rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) for n in sizes]
        Ad = [dev.to_device(a, dtype) for a in A]
        name, ms = dev.autotune_matvec_batched(Ad, 8, reps=2)
        assert name == lib.mpa_matvec_variant_name(code, 8, lib.mpa_matvec_variant(code, 8, -1)).decode() and ms > 0
        for a, ad in zip(A, Ad):
            x = rng.normal(size=(8, a.shape[0])) + 1j * rng.normal(size=(8, a.shape[0]))
            y = dev.matvec(ad, dev.to_device(x, dtype)).cpu().numpy()
            ref = x @ a.T
            assert np.max(np.abs(y - ref)) <= tol * np.abs(ref).max()
        # a system of another RHS class is left out of the timing
        x1 = [torch_like(ad, 8) for ad in Ad[:2]] + [torch_like(Ad[2], 1)]
        y1 = [t.clone() for t in x1]
        v, t = ctypes.c_int(-1), ctypes.c_double(0)
        arr = lambda ts: (ctypes.c_void_p * 3)(*[q.data_ptr() for q in ts])
        L.check(lib.mpa_autotune_matvec_batched(3, (ctypes.c_int * 3)(*sizes), (ctypes.c_int * 3)(8, 8, 1), code,
                                                arr(Ad), arr(x1), arr(y1), 1, ctypes.byref(v), ctypes.byref(t), None),
                "autotune_matvec_batched")
        assert v.value >= 0 and t.value > 0
        assert np.array_equal(y1[2].cpu().numpy(), x1[2].cpu().numpy())  # untouched


def torch_like(A, S):
    import torch
    return torch.randn((S, A.shape[0]), dtype=A.dtype, device=A.device)


def test_assembly_autotune_variants_match_oracle(dev, oracle):
    """Every assembly tile variant (tile width, table in LDS or not) equals the
    oracle's matrix to 1e-13; the autotuner keeps one of them."""
    import ctypes
    import torch
    import miepy_amd._lib as L
    lib = L.load()
    for N, lmax in ((23, 2), (19, 3), (9, 4)):
        pos, mie = _cluster(N, lmax, seed=40 + lmax)
        k = 2 * np.pi / (600 * nm)
        ref = oracle.sphere_aggregate_tmatrix(pos, mie, k)
        n = ref.shape[0]
        cnt = ctypes.c_int(0)
        saved = lib.mpa_assembly_variant(lmax, -1, ctypes.byref(cnt))
        assert cnt.value >= 2
        for v in range(cnt.value):
            assert lib.mpa_assembly_variant(lmax, v, None) == v
            got = dev.assemble(pos, mie, k).cpu().numpy()
            assert np.max(np.abs(got - ref)) <= 1e-13 * np.abs(ref).max(), (lmax, v)
        lib.mpa_assembly_variant(lmax, saved, None)
        out = torch.empty((n, n), dtype=dev.C128, device="cuda")
        J, ms = dev.autotune_assembly(pos, mie, k, out, reps=1)
        assert J >= 1 and ms > 0
        got = dev.assemble(pos, mie, k).cpu().numpy()
        assert np.max(np.abs(got - ref)) <= 1e-13 * np.abs(ref).max()
