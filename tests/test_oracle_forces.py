"""Pins the oracle's force and torque (cpp/src/forces.cpp:7-171) and its
plane-wave source decomposition at oblique incidence
(src/miepy/sources/plane_waves.py:126-156) without retyping their formulas.

Pins:
- Maxwell stress tensor: the field of a particle is the regular expansion of
  p_inc plus the outgoing expansion of p_scat about its centre; the
  time-averaged stress of that field, evaluated with the independent scipy
  VSH evaluator (tests/vsh_fields.py) and integrated over a sphere around the
  particle (Gauss-Legendre x trapezoid, exact for the band-limited
  integrand), equals every component of the oracle's Barton closed forms for
  force and torque: random coefficients at lmax 1-4, and every particle of a
  solved trimer lit at oblique incidence with elliptical polarization.  The
  integral does not depend on the sphere's radius.  H comes from the curl
  identities of the reference's N and M, themselves checked by finite
  differences;
- the plane-wave p_src at random (theta, phi, polarization, amplitude,
  phase), re-expanded with the incident-mode evaluator, equals the plane
  wave's own field near the particle, the error falling as (k rho)^lmax;
- one sphere lit along +-x, +-y, -z and oblique directions with linear,
  circular and elliptical light: extinction, scattering and absorption do not
  depend on direction or polarization and equal the Bohren-Huffman sums; the
  force along k equals the Bohren-Huffman radiation force and the force
  across k vanishes; the torque lies along k, with the sign of the incident
  wave's spin Im(E* x E).k and size |spin| P_abs / omega (angular momentum
  balance).
"""

import numpy as np
import pytest

import vsh_fields as V

nm = 1e-9
EPS0 = 8.854187817e-12
C_LIGHT = 299792458.0


@pytest.fixture(scope="module")
def H(oracle):
    import oracle.host as H
    return H


def _mst(p_inc, p_scat, lmax, center, k, eps_b, R):
    pts, nrm, wt = V.sphere_quadrature(center, R, 2 * lmax + 10, 4 * lmax + 20)
    Ei, Hi = V.expand_EH(p_inc, lmax, "incident", pts, center, k)
    Es, Hs = V.expand_EH(p_scat, lmax, "outgoing", pts, center, k)
    return V.maxwell_stress_force_torque(Ei + Es, Hi + Hs, pts, nrm, wt, center, EPS0 * eps_b)


@pytest.mark.parametrize("mode", ["incident", "outgoing"])
def test_curl_identities(mode):
    """H' = curl E / (i k) of an expansion = -i factor sum E_mn (p_e M + p_m N)."""
    rng = np.random.default_rng(2)
    lmax, k = 3, 2 * np.pi / 0.8
    p = rng.normal(size=(2, 15)) + 1j * rng.normal(size=(2, 15))
    c = np.array([0.1, -0.2, 0.05])
    pts = c + 0.3 * rng.normal(size=(5, 3))
    _, Hp = V.expand_EH(p, lmax, mode, pts, c, k)
    curl = V.curl_fd(lambda q: V.expand(p, lmax, mode, q, c, k), pts, 1e-4)
    assert np.max(np.abs(curl / (1j * k) - Hp)) < 1e-8 * np.max(np.abs(Hp))


@pytest.mark.parametrize("lmax", [1, 2, 3, 4])
def test_force_torque_maxwell_stress_random(oracle, lmax):
    rng = np.random.default_rng(40 + lmax)
    rmax = lmax * (lmax + 2)
    k = 2 * np.pi / 0.8
    eps_b = 1.7
    center = np.array([0.1, -0.2, 0.05])
    for trial in range(3):
        p_inc = rng.normal(size=(2, rmax)) + 1j * rng.normal(size=(2, rmax))
        p_scat = rng.normal(size=(2, rmax)) + 1j * rng.normal(size=(2, rmax))
        F_o = oracle.force(p_scat, p_inc, k, eps_b)
        T_o = oracle.torque(p_scat, p_inc, k, eps_b)
        for R in (2.5 / k, 4.0 / k):
            F, T = _mst(p_inc, p_scat, lmax, center, k, eps_b, R)
            assert np.max(np.abs(F - F_o)) <= 1e-8 * np.max(np.abs(F_o)), (R, F, F_o)
            assert np.max(np.abs(T - T_o)) <= 1e-8 * np.max(np.abs(T_o)), (R, T, T_o)


def test_force_torque_maxwell_stress_solved_trimer(H, oracle):
    """Every particle of a solved, unsymmetric trimer at oblique incidence with
    elliptical light: all six components of force and torque are nonzero."""
    pos = np.array([[0, 0, 0], [260 * nm, 40 * nm, -30 * nm], [-60 * nm, 230 * nm, 90 * nm]])
    eps_b = 1.33 ** 2
    lmax = 3
    r = H.solve_cluster(pos, [75 * nm, 60 * nm, 90 * nm], [-12 + 1.1j, 4.0 + 0.3j, -8 + 0.6j], 650 * nm, lmax,
                        polarization=(0.8, 0.35 + 0.5j), theta=1.1, phi=2.3, eps_b=eps_b, method="exact",
                        phase=0.7, amplitude=1.3)
    k = r["k"]
    for i in range(3):
        F_o, T_o = r["force"][i], r["torque"][i]
        assert np.min(np.abs(F_o)) > 1e-3 * np.max(np.abs(F_o))
        assert np.min(np.abs(T_o)) > 1e-3 * np.max(np.abs(T_o))
        F, T = _mst(r["p_inc"][i], r["p_scat"][i], lmax, pos[i], k, eps_b, 110 * nm)
        assert np.max(np.abs(F - F_o)) <= 1e-8 * np.max(np.abs(F_o)), (i, F, F_o)
        assert np.max(np.abs(T - T_o)) <= 1e-8 * np.max(np.abs(T_o)), (i, T, T_o)


@pytest.mark.parametrize("lmax,tol", [(3, 3e-3), (6, 2e-7), (8, 2e-10)])
def test_plane_wave_expansion_oblique(H, lmax, tol):
    rng = np.random.default_rng(60 + lmax)
    k = 2 * np.pi / 0.7
    for trial in range(4):
        th = np.arccos(rng.uniform(-1, 1))
        ph = rng.uniform(0, 2 * np.pi)
        pol = rng.normal(size=2) + 1j * rng.normal(size=2)
        amp, phase = rng.uniform(0.5, 2), rng.uniform(0, 6)
        r0 = 0.3 * rng.normal(size=3)
        p = H.plane_wave_structure(r0, k, lmax, pol, th, ph, amp, phase)[0]
        dirs = rng.normal(size=(8, 3))
        dirs /= np.linalg.norm(dirs, axis=1)[:, None]
        errs = []
        for rho in (0.4 / k, 0.2 / k):
            pts = r0 + rho * dirs
            E_exp = V.expand(p, lmax, "incident", pts, r0, k)
            E_pw, _ = V.plane_wave_E(pts, k, pol, th, ph, amp, phase)
            errs.append(np.linalg.norm(E_exp - E_pw) / np.linalg.norm(E_pw))
        # the truncated regular expansion misses O((k rho)^lmax) of the field
        assert errs[0] < tol, errs
        assert errs[0] / errs[1] > 0.4 * 2 ** lmax, errs


DIRECTIONS = [(np.pi / 2, 0.0), (np.pi / 2, np.pi), (np.pi / 2, np.pi / 2), (np.pi / 2, 3 * np.pi / 2),
              (np.pi, 0.0), (0.0, 0.0), (0.7, 2.1), (2.3, 4.0)]
POLARIZATIONS = [(1, 0), (0, 1), (1, 1j), (1, -1j), (0.3, 0.8 + 0.2j)]


def test_single_sphere_any_direction(H):
    from test_oracle_cluster import _single_mie_closed_form

    radius, eps, wl, lmax = 90 * nm, 4.0 + 0.3j, 650 * nm, 4
    eps_b = 1.33 ** 2
    Cs, Ca, Ce, Fk = _single_mie_closed_form(radius, eps, wl, lmax, eps_b)
    omega = 2 * np.pi * C_LIGHT / wl
    intensity = EPS0 * np.sqrt(eps_b) * C_LIGHT / 2  # unit amplitude
    center = np.array([30 * nm, -20 * nm, 10 * nm])
    for th, ph in DIRECTIONS:
        for pol in POLARIZATIONS:
            r = H.solve_cluster(center, radius, eps, wl, lmax, polarization=pol, theta=th, phi=ph, eps_b=eps_b)
            np.testing.assert_allclose(r["cross_sections"], [Cs, Ca, Ce], rtol=1e-12)
            E, k_hat = V.plane_wave_E(np.zeros((1, 3)), 1.0, pol, th, ph)
            E = E[:, 0]
            spin = np.real(np.cross(np.conj(E), E) / 1j) @ k_hat  # Im(E* x E).k, |E| = 1
            F, T = r["force"][0], r["torque"][0]
            assert F @ k_hat == pytest.approx(Fk, rel=1e-9)
            assert np.linalg.norm(F - (F @ k_hat) * k_hat) < 1e-9 * abs(Fk)
            T_expect = spin * intensity * Ca / omega
            assert T @ k_hat == pytest.approx(T_expect, rel=1e-9, abs=1e-9 * intensity * Ca / omega)
            assert np.linalg.norm(T - (T @ k_hat) * k_hat) < 1e-9 * intensity * Ca / omega
