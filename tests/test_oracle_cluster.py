"""Pins the oracle's solve, cross-section and force/torque paths end to end.

Reference: cpp/src/interactions.cpp:98-209, 347-368 (solvers);
cpp/src/flux.cpp:5-57; cpp/src/forces.cpp:7-171; src/miepy/sphere_cluster.py.

Pins:
- LU vs numpy.linalg.solve; BiCGSTAB residual <= tol; BiCGSTAB at the
  reference default 1e-5 vs exact to 1e-6 (tests/test_jax_end_to_end.py:315-327);
  custom bicgstab zero-RHS early return (interactions.cpp:107-109);
- single sphere GMT == textbook single-sphere Mie theory (cross-sections,
  radiation force), as tests/test_single_gmt.py:21-71;
- origin independence (tests/test_cluster.py:17-50) and interactions-off =
  N x single (tests/test_cluster.py:53-95);
- lossless spheres absorb nothing;
- the reference's own FDTD fixture for the Au dimer
  (tests/test_fdtd_comparison.py:318-356 tolerances);
- cluster_cross_sections: extinction of a single centred sphere.
"""

import os

import numpy as np
import pytest

nm = 1e-9


@pytest.fixture(scope="module")
def H(oracle):
    import oracle.host as H
    return H


def _rand_system(n, seed=0, scale=0.05):
    rng = np.random.default_rng(seed)
    T = scale * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    b = rng.normal(size=n) + 1j * rng.normal(size=n)
    return T, b


def test_lu_vs_numpy(oracle):
    T, b = _rand_system(60, 1, 0.2)
    A = np.eye(60) + T
    x = oracle.lu_solve(A, b)
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-12, atol=1e-13)


def test_bicgstab_eigen_residual(oracle):
    T, b = _rand_system(80, 2)
    A = np.eye(80) + T
    for tol in (1e-5, 1e-10, 1e-13):
        x, it, err = oracle.bicgstab_eigen(A, b, tol=tol)
        rel = np.linalg.norm(A @ x - b) / np.linalg.norm(b)
        assert rel <= tol * 1.01 and err <= tol
        assert it > 0


def test_bicgstab_zero_rhs(oracle):
    A = np.eye(10, dtype=complex)
    x, it, err = oracle.bicgstab_eigen(A, np.zeros(10, complex))
    assert it == 0 and np.all(x == 0)
    x, it, res = oracle.bicgstab(A, np.zeros(10, complex), 100, 1e-5)
    assert it == 0 and np.all(x == 0)


def test_bicgstab_custom(oracle):
    T, b = _rand_system(50, 3)
    A = np.eye(50) + T
    x, it, res = oracle.bicgstab(A, b, 1000, 1e-12)
    assert np.linalg.norm(A @ x - b) < 1e-11
    x1, it1, _ = oracle.bicgstab(A, b, 1, 0.0)  # forced non-convergence returns silently
    assert it1 == 2


def test_bicgstab_maxiter(oracle):
    T, b = _rand_system(50, 4, 0.3)
    x, it, err = oracle.bicgstab_eigen(np.eye(50) + T, b, tol=1e-14, maxiter=3)
    assert it == 3 and err > 1e-14


def _single_mie_closed_form(radius, eps, wavelength, lmax, eps_b=1.0):
    """Bohren & Huffman sums (mie_single/scattering.py:12-57, mie_sphere.py:188-208)."""
    from scipy import special
    nb = np.sqrt(eps_b)
    k = 2 * np.pi * nb / wavelength
    x = k * radius
    m = np.sqrt(eps / eps_b)
    a = np.zeros(lmax, complex)
    b = np.zeros(lmax, complex)
    for n in range(1, lmax + 1):
        psi = lambda z: z * special.spherical_jn(n, z)
        dpsi = lambda z: special.spherical_jn(n, z) + z * special.spherical_jn(n, z, True)
        xi = x * (special.spherical_jn(n, x) + 1j * special.spherical_yn(n, x))
        dxi = dpsi(x) + 1j * (special.spherical_yn(n, x) + x * special.spherical_yn(n, x, True))
        a[n - 1] = (m * psi(m * x) * dpsi(x) - psi(x) * dpsi(m * x)) / (m * psi(m * x) * dxi - xi * dpsi(m * x))
        b[n - 1] = (psi(m * x) * dpsi(x) - m * psi(x) * dpsi(m * x)) / (psi(m * x) * dxi - m * xi * dpsi(m * x))
    nv = np.arange(1, lmax + 1)
    Cs = 2 * np.pi / k ** 2 * np.sum((2 * nv + 1) * (abs(a) ** 2 + abs(b) ** 2))
    Ce = 2 * np.pi / k ** 2 * np.sum((2 * nv + 1) * np.real(a + b))
    Fz = 0.0
    for i in range(lmax):
        n = i + 1
        Fz += 2 * np.pi * (2 * n + 1) / k ** 2 * np.real(a[i] + b[i])
        Fz -= 4 * np.pi * (2 * n + 1) / (n * (n + 1)) / k ** 2 * np.real(a[i] * np.conj(b[i]))
        if n < lmax:
            Fz -= 4 * np.pi * (n * (n + 2)) / (n + 1) / k ** 2 * np.real(a[i] * np.conj(a[i + 1]) + b[i] * np.conj(b[i + 1]))
    eps0 = 8.854187817e-12
    Fz *= eps0 * eps_b / 2
    return Cs, Ce - Cs, Ce, Fz


def test_single_sphere_vs_mie_theory(H):
    """tests/test_single_gmt.py: dielectric n=3.7+, water, lmax 5."""
    eps = 3.7 ** 2 + 0.1j
    eps_b = 1.33 ** 2
    for wl in np.linspace(400 * nm, 1000 * nm, 4):
        r = H.solve_cluster([0, 0, 0], 100 * nm, eps, wl, 5, eps_b=eps_b)
        Cs, Ca, Ce, Fz = _single_mie_closed_form(100 * nm, eps, wl, 5, eps_b)
        C = r["cross_sections"]
        assert C[0] == pytest.approx(Cs, rel=1e-13)
        assert C[1] == pytest.approx(Ca, rel=1e-12)
        assert C[2] == pytest.approx(Ce, rel=1e-13)
        assert r["force"][0, 2] == pytest.approx(Fz, rel=1e-9)
        assert abs(r["force"][0, 0]) < 1e-12 * abs(Fz)


def test_origin_independence(H):
    """tests/test_cluster.py:17-50"""
    eps = -15 + 1.0j
    for wl in (600 * nm, 800 * nm):
        a = H.solve_cluster([0, 0, 0], 75 * nm, eps, wl, 2, polarization=(0, 1))["cross_sections"]
        b = H.solve_cluster([40 * nm, 50 * nm, 60 * nm], 75 * nm, eps, wl, 2, polarization=(0, 1))["cross_sections"]
        np.testing.assert_allclose(a, b, rtol=1e-13)


def test_interactions_off(H):
    """tests/test_cluster.py:53-95"""
    eps = -15 + 1.0j
    pos = [[100 * nm, 0, 0], [-100 * nm, 0, 0]]
    r = H.solve_cluster(pos, 75 * nm, eps, 700 * nm, 2, polarization=(0, 1), interactions=False)
    s = H.solve_cluster([0, 0, 0], 75 * nm, eps, 700 * nm, 2, polarization=(0, 1))
    assert r["cross_sections"][2] == pytest.approx(2 * s["cross_sections"][2], rel=1e-14)


def test_lossless_absorbs_nothing(H):
    pos = [[0, 0, 0], [300 * nm, 0, 0], [0, 350 * nm, 100 * nm]]
    r = H.solve_cluster(pos, 75 * nm, 2.56, 800 * nm, 3, eps_b=1.33 ** 2, method="exact")
    C = r["cross_sections"]
    assert abs(C[1]) < 1e-12 * C[2]


def test_exact_vs_bicgstab(H):
    """tests/test_jax_end_to_end.py:315-327 (default tol 1e-5 agrees to 1e-6)
    and the tight tolerance agrees to 1e-10."""
    ag600 = -15.8 + 1.07j  # silver-like permittivity at 600 nm (constant here)
    dimer = [[100 * nm, 0, 0], [-100 * nm, 0, 0]]
    e = H.solve_cluster(dimer, 75 * nm, ag600, 600 * nm, 2, method="exact")
    b5 = H.solve_cluster(dimer, 75 * nm, ag600, 600 * nm, 2, method="bicgstab", tol=1e-5)
    np.testing.assert_allclose(b5["p_inc"], e["p_inc"], rtol=1e-6, atol=1e-15)
    pos = [[0, 0, 0], [250 * nm, 0, 0], [-250 * nm, 0, 0], [0, 250 * nm, 0], [0, -250 * nm, 0]]
    kw = dict(polarization=(1, 1j))
    e = H.solve_cluster(pos, 75 * nm, -15 + 1j, 600 * nm, 2, method="exact", **kw)
    b12 = H.solve_cluster(pos, 75 * nm, -15 + 1j, 600 * nm, 2, method="bicgstab", tol=1e-12, **kw)
    np.testing.assert_allclose(b12["p_scat"], e["p_scat"], rtol=1e-10, atol=1e-12 * np.abs(e["p_scat"]).max())
    np.testing.assert_allclose(b12["cross_sections"], e["cross_sections"], rtol=1e-11)


def test_fdtd_fixture(H, golden_dir):
    """The reference's FDTD data for an Au dimer (tests/test_fdtd_comparison.py)."""
    d = np.load(os.path.join(golden_dir, "fdtd_au_dimer.npz"))
    sep, rad = float(d["separation"]), float(d["radius"])
    pos = [[-sep / 2, 0, 0], [sep / 2, 0, 0]]
    C = np.zeros((3, 40))
    F = np.zeros((3, 40))
    for i, (wl, eps) in enumerate(zip(d["wavelengths"], d["eps"])):
        r = H.solve_cluster(pos, rad, eps, wl, 2, polarization=(1, 1j))
        C[:, i] = r["cross_sections"]
        F[:, i] = r["force"][1]
    for ref, got, tol in [(d["scat"], C[0], 8e-15), (d["absorb"], C[1], 1.1e-14), (d["extinct"], C[2], 1.9e-14),
                          (d["Fx"], F[0], 5e-27), (d["Fy"], F[1], 3.4e-28), (d["Fz"], F[2], 3.7e-26)]:
        assert np.linalg.norm(ref - got) / ref.shape[0] < tol


def test_cluster_cross_sections_single_sphere(oracle, H):
    r = H.solve_cluster([0, 0, 0], 90 * nm, 4.0 + 0.2j, 650 * nm, 3)
    # a single sphere at the origin is its own cluster expansion
    Cs, Ca, Ce = oracle.cluster_cross_sections(r["p_scat"][0], r["p_src"][0], r["k"])
    np.testing.assert_allclose([Cs.sum(), Ca.sum(), Ce.sum()], r["cross_sections"], rtol=1e-13)


def test_torque_zero_for_linear_pol_single_sphere(H):
    r = H.solve_cluster([0, 0, 0], 90 * nm, 4.0 + 0.2j, 650 * nm, 3, polarization=(1, 0))
    rc = H.solve_cluster([0, 0, 0], 90 * nm, 4.0 + 0.2j, 650 * nm, 3, polarization=(1, 1j))
    # circular light twists an absorbing sphere about k; linear light does not.
    # Angular momentum balance: T_z = P_abs / omega = I * C_abs / omega with
    # I = eps0 c |E|^2 / 2 for a unit-amplitude wave; (1, i) along +z is
    # x + i y, whose spin Im(E* x E) points along +z.
    c = 299792458.0
    omega = 2 * np.pi * c / (650 * nm)
    expect = 8.854187817e-12 * c / 2 * rc["cross_sections"][1] / omega
    assert rc["torque"][0, 2] == pytest.approx(expect, rel=1e-9)
    assert np.all(np.abs(r["torque"]) < 1e-10 * abs(expect))
