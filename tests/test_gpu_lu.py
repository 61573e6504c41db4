"""Dense LU with partial pivoting on the device (the `exact` method of
solve_linear_system, cpp/src/interactions.cpp:356-357) against the oracle's
LU restatement and numpy.

Bars: residual ||A x - b|| <= 1e-13 ||A|| ||x|| and the oracle's LU solution
to 1e-11 relative on random complex systems (n = 1 ... 1000, sizes across
the 32-column panel and 64-row tile edges, several right-hand sides); a
matrix whose diagonal is zero (pivoting required) where BiCGSTAB stalls;
add_identity = factoring I + T; a singular matrix raises; `exact` through
the backend, `sphere_cluster` and `cluster` equals the oracle's LU path
(1e-12 on p_inc, 1e-10 on cross-sections).
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


def _random(rng, n, S):
    A = (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) / np.sqrt(n) + 1.5 * np.eye(n)
    b = rng.normal(size=(S, n)) + 1j * rng.normal(size=(S, n))
    return A, b


@pytest.mark.parametrize("n,S", [(1, 1), (5, 2), (31, 1), (32, 3), (33, 1), (64, 2), (65, 1), (97, 4), (256, 1),
                                 (1000, 2)])
def test_lu_matches_oracle(dev, oracle, n, S):
    rng = np.random.default_rng(100 + n)
    A, b = _random(rng, n, S)
    x = dev.lu_solve(dev.to_device(A, dev.C128), dev.to_device(b, dev.C128)).cpu().numpy()
    nA = np.linalg.norm(A, 2) if n <= 256 else np.linalg.norm(A)
    for s in range(S):
        assert np.linalg.norm(A @ x[s] - b[s]) <= 1e-13 * nA * np.linalg.norm(x[s])
        ref = oracle.lu_solve(A, b[s])
        assert np.linalg.norm(x[s] - ref) <= 1e-11 * np.linalg.norm(ref)


def test_lu_pivots_where_bicgstab_stalls(dev, oracle):
    """A scaled cyclic shift: zero diagonal, eigenvalues around the origin."""
    rng = np.random.default_rng(7)
    n = 200
    A = np.roll(np.eye(n), 1, axis=1).astype(complex) * (1 + 0.1 * rng.normal(size=n))[:, None]
    b = rng.normal(size=n) + 1j * rng.normal(size=n)
    _, _, err = oracle.bicgstab_eigen(A, b, 1e-10, 1000)
    assert err > 1e-3  # the iterative solver does not get there
    x = dev.lu_solve(dev.to_device(A, dev.C128), dev.to_device(b, dev.C128)).cpu().numpy()
    assert np.linalg.norm(A @ x - b) <= 1e-14 * np.linalg.norm(b)
    assert np.linalg.norm(x - oracle.lu_solve(A, b)) <= 1e-13 * np.linalg.norm(x)


def test_lu_add_identity_and_overwrite(dev):
    rng = np.random.default_rng(8)
    n = 150
    A, b = _random(rng, n, 1)
    T = A - np.eye(n)
    Td = dev.to_device(T, dev.C128)
    x1 = dev.lu_solve(Td, dev.to_device(b, dev.C128), add_identity=True).cpu().numpy()
    assert np.array_equal(Td.cpu().numpy(), T)  # copied, not factored in place
    x2 = dev.lu_solve(dev.to_device(A, dev.C128), dev.to_device(b, dev.C128)).cpu().numpy()
    assert np.array_equal(x1, x2)
    F = dev.to_device(A, dev.C128)
    x3 = dev.lu_solve(F, dev.to_device(b, dev.C128), overwrite=True).cpu().numpy()
    assert np.array_equal(x3, x2)
    LU = F.cpu().numpy()
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    # P A = L U for some row permutation: every row of L U is a row of A
    rows = np.argmin(np.abs((L @ U)[:, None, :] - A[None, :, :]).max(axis=2), axis=1)
    assert sorted(rows) == list(range(n))
    assert np.abs(L @ U - A[rows]).max() <= 1e-13 * np.abs(A).max()
    assert np.all(np.abs(L) <= 1 + 1e-12)  # partial pivoting bounds |L| by 1


def test_lu_singular_raises(dev):
    rng = np.random.default_rng(9)
    n = 70
    A, b = _random(rng, n, 1)
    A[:, 40] = 0.0
    with pytest.raises(np.linalg.LinAlgError, match="column 40"):
        dev.lu_solve(dev.to_device(A, dev.C128), dev.to_device(b, dev.C128))


def test_exact_method_through_backend_and_clusters(dev, oracle):
    import miepy_amd as M
    import oracle.host as H
    from miepy_amd import backend as B
    from miepy_amd.workloads import hex37
    w = hex37()
    c = M.sphere_cluster(position=w.position, radius=w.radius, material=w.materials(), source=w.sources[0],
                         wavelength=w.wavelength, lmax=w.lmax, medium=w.medium(), method=B.solver.exact)
    s = w.sources[0]
    ref = H.solve_cluster(w.position, w.radius, w.eps, w.wavelength, w.lmax, polarization=s.polarization,
                          theta=s.theta, phi=s.phi, eps_b=w.medium_eps, method="exact")
    assert np.linalg.norm(c.p_inc - ref["p_inc"]) <= 1e-12 * np.linalg.norm(ref["p_inc"])
    assert np.max(np.abs(np.array(c.cross_sections()) - ref["cross_sections"])) <= 1e-10 * abs(
        ref["cross_sections"][2])
    # backend dispatch: solve_linear_system(T, p_src, exact)
    k = float(np.real(c.material_data.k_b))
    T = B.sphere_aggregate_tmatrix(w.position, c.mie_scat, k)
    p = B.solve_linear_system(T, c.p_src, B.solver.exact)
    To = oracle.sphere_aggregate_tmatrix(w.position, c.mie_scat, k)
    po, _, _ = oracle.solve_linear_system(To, c.p_src.reshape(-1), method="exact")
    assert np.linalg.norm(p.reshape(-1) - po) <= 1e-12 * np.linalg.norm(po)
    # general T-matrix cluster, stored mode
    pos = w.position[:6]
    rmax = w.lmax * (w.lmax + 2)
    rng = np.random.default_rng(10)
    tm = 0.1 * (rng.normal(size=(6, 2, rmax, 2, rmax)) + 1j * rng.normal(size=(6, 2, rmax, 2, rmax)))
    cl = M.cluster(particles=[M.tmatrix_particle(p_, t) for p_, t in zip(pos, tm)],
                   source=M.plane_wave.from_string(polarization="y"), wavelength=600 * nm, lmax=w.lmax,
                   method=B.solver.exact)
    kk = 2 * np.pi / (600 * nm)
    Tt = oracle.particle_aggregate_tmatrix(pos, tm, kk)
    pi = oracle.lu_solve(np.eye(Tt.shape[0]) + Tt, cl.p_src.reshape(-1))
    assert np.linalg.norm(cl.p_inc.reshape(-1) - pi) <= 1e-12 * np.linalg.norm(pi)


def test_lu_large_system(dev):
    import time
    import torch
    rng = np.random.default_rng(11)
    n = 4000
    A, b = _random(rng, n, 1)
    Ad = dev.to_device(A, dev.C128)
    bd = dev.to_device(b, dev.C128)
    dev.lu_solve(Ad, bd)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x = dev.lu_solve(Ad, bd).cpu().numpy()
    dt = time.perf_counter() - t0
    print(f"LU solve n={n}: {dt * 1e3:.1f} ms")
    assert np.linalg.norm(A @ x[0] - b[0]) <= 1e-13 * np.linalg.norm(A) * np.linalg.norm(x[0])
