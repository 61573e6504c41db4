"""Pins the oracle's special functions (reference cpp/src/special.cpp).

Pins (none retypes the oracle's formula):
- h_n^(1) upward recurrence (special.cpp:90-112) vs scipy.special j_n + i y_n
  golden vectors, and the closed forms for n = 0, 1.
- Legendre, no Condon-Shortley phase (special.cpp:180-204) vs
  scipy.special.lpmv * (-1)^m golden vectors; negative-m identity.
- pi/tau (special.cpp:209-270) vs central finite differences of P_n^m(cos t)
  and the theta = 0 / pi limits of the reference's own special cases.
- Wigner 3j and the Schulten-Gordon batch (special.cpp:272-363) vs sympy's
  exact values, including the reference's known cases
  (tests/test_jax_special.py:289-312) and selection rules (:308-312).
- gaunt_batch (special.cpp:365-440) vs sympy-exact a_q/b_q for n, v <= 5 and
  vs the direct a_func/b_func.
"""

import os

import numpy as np
import pytest


def test_hankel_vs_scipy(oracle, golden_dir):
    d = np.load(os.path.join(golden_dir, "special_scipy.npz"))
    for iz, z in enumerate(d["z"]):
        nmax = 15
        h = oracle.spherical_hn_recursion(nmax, z)
        ref = d["hn"][iz]
        # relative error per order; upward j_n loses digits only where
        # |y_n| >> |j_n|, which keeps |h_n| accurate.
        rel = np.abs(h - ref) / np.abs(ref)
        assert np.max(rel) < 1e-13, (z, rel)


def test_hankel_closed_forms(oracle):
    for z in [0.4, 1.7, 9.3]:
        h = oracle.spherical_hn_recursion(1, z)
        h0 = -1j * np.exp(1j * z) / z
        h1 = -(z + 1j) * np.exp(1j * z) / z ** 2
        assert abs(h[0] - h0) < 1e-15 * abs(h0)
        assert abs(h[1] - h1) < 1e-14 * abs(h1)


def test_legendre_vs_scipy(oracle, golden_dir):
    d = np.load(os.path.join(golden_dir, "special_scipy.npz"))
    nmax = d["leg"].shape[1] - 1
    for ix, x in enumerate(d["x"]):
        P = oracle.associated_legendre_recursion(nmax, x)
        for n in range(nmax + 1):
            for m in range(-n, n + 1):
                ref = d["leg"][ix, n, abs(m)]
                if m < 0:
                    from math import factorial
                    ref *= (-1) ** (-m) * factorial(n + m) / factorial(n - m)
                got = P[n * (n + 2) - n + m]
                assert abs(got - ref) <= 1e-13 * max(1.0, abs(ref)), (x, n, m, got, ref)
                # single-value entry point agrees with the table
                assert abs(oracle.associated_legendre(n, m, x) - got) <= 1e-14 * max(1.0, abs(got))


def test_legendre_no_condon_shortley(oracle):
    # P_1^1(x) = +sqrt(1-x^2) without the phase (special.cpp:152)
    x = 0.37
    assert oracle.associated_legendre(1, 1, x) == pytest.approx(np.sqrt(1 - x * x), rel=1e-15)
    assert oracle.associated_legendre(1, -1, x) == pytest.approx(-np.sqrt(1 - x * x) / 2, rel=1e-15)
    assert oracle.associated_legendre(2, 2, x) == pytest.approx(3 * (1 - x * x), rel=1e-14)


@pytest.mark.parametrize("n,m", [(1, 1), (1, -1), (1, 0), (2, 1), (2, -1), (2, 2), (3, 2), (3, -3), (5, 3), (5, -3)])
def test_pi_tau_finite_difference(oracle, n, m):
    for th in [0.3, 0.7, 1.0, 1.5, 2.5]:
        h = 1e-5
        P = lambda t: oracle.associated_legendre(n, m, np.cos(t))
        fd = (P(th + h) - P(th - h)) / (2 * h)
        assert oracle.tau_func(n, m, th) == pytest.approx(fd, rel=1e-8, abs=1e-10)
        assert oracle.pi_func(n, m, th) == pytest.approx(m / np.sin(th) * P(th), rel=1e-13, abs=1e-15)


@pytest.mark.parametrize("n", [1, 2, 3, 5])
def test_pi_tau_poles_are_limits(oracle, n):
    # the theta = 0 / pi special cases (special.cpp:210-224, 234-247) are the
    # limits of the general expressions
    for m in (-1, 0, 1):
        for t0, sgn in ((0.0, 1), (np.pi, -1)):
            t = t0 + sgn * 1e-4  # O(t^2) from the limit, away from cos-cancellation
            assert oracle.pi_func(n, m, t0) == pytest.approx(oracle.pi_func(n, m, t), rel=1e-6, abs=1e-6)
            # tau_n^0 leaves the pole linearly: |tau| ~ n(n+1)/2 * t
            assert oracle.tau_func(n, m, t0) == pytest.approx(oracle.tau_func(n, m, t), rel=1e-6, abs=n * (n + 1) * 1e-4)


def test_emn(oracle):
    from math import factorial
    for n in range(1, 6):
        for m in range(-n, n + 1):
            mag = np.sqrt((2 * n + 1) * factorial(n - m) / (n * (n + 1) * factorial(n + m)))
            assert oracle.Emn(m, n) == pytest.approx(1j ** n * mag, rel=1e-14)


def test_wigner_3j_known_cases(oracle, golden_dir):
    g = np.load(os.path.join(golden_dir, "gaunt_sympy.npz"))
    for case, val in zip(g["w3j_cases"], g["w3j_values"]):
        assert oracle.wigner_3j(*[int(c) for c in case]) == pytest.approx(val, rel=1e-12, abs=1e-15)


def test_wigner_3j_selection_rules(oracle):
    assert oracle.wigner_3j(1, 1, 1, 0, 0, 1) == 0.0  # m1+m2+m3 != 0
    assert oracle.wigner_3j(1, 1, 5, 0, 0, 0) == 0.0  # triangle violated


@pytest.mark.parametrize("j2,j3,m1,m2,m3", [(1, 1, 0, 0, 0), (2, 1, 0, 0, 0), (3, 2, -1, 0, 1), (3, 2, 0, 0, 0),
                                              (5, 3, -2, 1, 1), (4, 4, 0, 0, 0), (5, 5, -3, 1, 2)])
def test_wigner_3j_batch(oracle, j2, j3, m1, m2, m3):
    from sympy.physics.wigner import wigner_3j as w3
    jmin, jmax, vals = oracle.wigner_3j_batch(j2, j3, m1, m2, m3)
    ref = np.array([float(w3(j, j2, j3, m1, m2, m3)) for j in range(jmin, jmax + 1)])
    np.testing.assert_allclose(vals, ref, rtol=1e-11, atol=1e-15)


def test_gaunt_batch_vs_sympy(oracle, golden_dir):
    rows = np.load(os.path.join(golden_dir, "gaunt_sympy.npz"))["rows"]
    cache = {}
    worst = 0.0
    for m, n, u, v, q, kind, val in rows:
        key = (int(m), int(n), int(u), int(v))
        if key not in cache:
            cache[key] = oracle.gaunt_batch(*key)
        got = cache[key][int(kind)][int(q)]
        err = abs(got - val) / max(abs(val), 1e-3)
        worst = max(worst, err)
    assert worst < 1e-11, worst


def test_gaunt_batch_vs_direct(oracle):
    for (m, n, u, v) in [(0, 1, 0, 1), (1, 1, -1, 1), (1, 2, 0, 1), (1, 3, -1, 2), (2, 3, -1, 2), (-1, 2, 1, 3),
                         (2, 4, -2, 4), (1, 5, -1, 5)]:
        a, b = oracle.gaunt_batch(m, n, u, v)
        for q in range(len(a)):
            assert a[q] == pytest.approx(oracle.a_func(m, n, u, v, n + v - 2 * q), rel=1e-10, abs=1e-14)
        for q in range(1, len(b)):
            assert b[q] == pytest.approx(oracle.b_func(m, n, u, v, n + v - 2 * q), rel=1e-10, abs=1e-14)
