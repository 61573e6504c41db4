"""Generate the committed golden fixtures under tests/golden/.

Run from the repo root:  python tests/golden/make_golden.py [--reference /root/reference]

Fixtures (all small .npz files, inputs + expected outputs):
- special_scipy.npz   spherical h_n^(1) = j_n + i y_n (scipy.special) and
                      associated Legendre P_n^m without Condon-Shortley phase
                      (scipy.special.lpmv * (-1)^m) on fixed grids.
- gaunt_sympy.npz     a_q, b_q Gaunt coefficients for every (m, n, u, v) with
                      n, v <= 5, evaluated from the definitions of
                      cpp/src/special.cpp:276-296 with sympy's exact Wigner 3j
                      (sympy.physics.wigner), plus the known 3j cases of
                      tests/test_jax_special.py:289-312.
- mie_mpmath.npz      Mie coefficients a_n, b_n, c_n, d_n of 36 spheres
                      (size parameters 0.01-40; dielectric, absorbing,
                      plasmonic, metallic, magnetic spheres; three media) by
                      the formulas of src/miepy/mie_single/mie_sphere.py:16-76
                      evaluated in 30-digit arithmetic (mpmath Bessel
                      functions; psi' from psi_{n-1} - (n/z) psi_n).
- fdtd_au_dimer.npz   the reference's own FDTD data for the Au dimer
                      (tests/test_fdtd_comparison.py): wavelengths, inline eps,
                      FDTD cross-sections and forces.  The arrays are read from
                      that file's text with ast (the file is not imported).

No reference code is imported or executed.
"""

import argparse
import ast
import os

import numpy as np
from scipy import special
import sympy
from sympy.physics.wigner import wigner_3j

HERE = os.path.dirname(os.path.abspath(__file__))


def make_special():
    z = np.array([0.3, 0.5, 1.0, 2.3, 3.7, 5.0, 7.5, 10.0, 15.0, 25.0])
    nmax = 15
    hn = np.array([[special.spherical_jn(n, zz) + 1j * special.spherical_yn(n, zz) for n in range(nmax + 1)]
                   for zz in z])
    x = np.array([-0.95, -0.8, -0.5, -0.1, 0.0, 0.3, 0.5, 0.7, 0.9, 0.99])
    lmax_leg = 11
    leg = np.zeros((len(x), lmax_leg + 1, lmax_leg + 1))  # [x, n, m>=0]
    for ix, xx in enumerate(x):
        for n in range(lmax_leg + 1):
            for m in range(n + 1):
                leg[ix, n, m] = special.lpmv(m, n, xx) * (-1) ** m
    np.savez(os.path.join(HERE, "special_scipy.npz"), z=z, hn=hn, x=x, leg=leg)


def _a_exact(m, n, u, v, p):
    f = sympy.factorial
    num = f(n + m) * f(v + u) * f(p - m - u)
    den = f(n - m) * f(v - u) * f(p + m + u)
    val = (-1) ** (m + u) * (2 * p + 1) * sympy.sqrt(sympy.Rational(num, den)) * \
        wigner_3j(n, v, p, 0, 0, 0) * wigner_3j(n, v, p, m, u, -m - u)
    return float(sympy.N(val, 30))


def _b_exact(m, n, u, v, p):
    f = sympy.factorial
    num = f(n + m) * f(v + u) * f(p - m - u + 1)
    den = f(n - m) * f(v - u) * f(p + m + u + 1)
    val = (-1) ** (m + u) * (2 * p + 3) * sympy.sqrt(sympy.Rational(num, den)) * \
        wigner_3j(n, v, p, 0, 0, 0) * wigner_3j(n, v, p + 1, m, u, -m - u)
    return float(sympy.N(val, 30))


def make_gaunt(lmax=5):
    rows = []  # m, n, u, v, q, kind(0=a,1=b), value
    for n in range(1, lmax + 1):
        for v in range(1, lmax + 1):
            for m in range(-n, n + 1):
                for u in range(-v, v + 1):
                    qa = min(n, v, (n + v - abs(m + u)) // 2)
                    qb = min(n, v, (n + v + 1 - abs(m + u)) // 2)
                    for q in range(qa + 1):
                        rows.append((m, n, u, v, q, 0, _a_exact(m, n, u, v, n + v - 2 * q)))
                    for q in range(1, qb + 1):
                        rows.append((m, n, u, v, q, 1, _b_exact(m, n, u, v, n + v - 2 * q)))
    rows = np.array(rows, dtype=float)
    cases = [(1, 1, 0, 0, 0, 0), (1, 1, 2, 0, 0, 0), (2, 1, 3, 0, 0, 0), (2, 1, 3, 2, -1, -1),
             (1, 1, 2, 1, -1, 0), (3, 2, 1, 0, 0, 0), (3, 2, 3, 1, -2, 1), (5, 3, 4, 2, -1, -1),
             (7, 6, 5, 3, -2, -1), (11, 6, 9, 4, -2, -2), (10, 10, 10, 0, 0, 0), (9, 7, 5, -3, 1, 2)]
    w3 = np.array([float(sympy.N(wigner_3j(*c), 30)) for c in cases])
    np.savez(os.path.join(HERE, "gaunt_sympy.npz"), rows=rows, w3j_cases=np.array(cases), w3j_values=w3)


def make_mie(lmax=8):
    import mpmath as mp
    mp.mp.dps = 30
    rng = np.random.default_rng(11)
    eps_set = [2.56, 12.0 + 0.3j, 4.0 + 0.05j, -20 + 1.2j, -1.2 + 0.2j, -50 + 5j, 0.1 + 3j, 16.0]
    media = [(1.0, 1.0), (1.33 ** 2, 1.0), (2.25, 1.2)]
    N = 36
    x = np.exp(np.linspace(np.log(0.01), np.log(40.0), N))
    rng.shuffle(x)
    eps = np.array([eps_set[i % len(eps_set)] for i in range(N)], dtype=complex)
    mu = np.where(np.arange(N) % 5 == 0, 1.5 + 0.1j, 1.0 + 0j)
    eps_b = np.array([media[i % 3][0] for i in range(N)], dtype=complex)
    mu_b = np.array([media[i % 3][1] for i in range(N)], dtype=complex)
    k = 2 * np.pi * np.sqrt(np.real(eps_b * mu_b)) / 600e-9
    radius = x / k

    def psi(n, t):  # t j_n(t)
        return t * mp.sqrt(mp.pi / (2 * t)) * mp.besselj(n + mp.mpf(1) / 2, t)

    def chi(n, t):  # t y_n(t)
        return t * mp.sqrt(mp.pi / (2 * t)) * mp.bessely(n + mp.mpf(1) / 2, t)

    a = np.zeros((N, 2, lmax), dtype=complex)
    c = np.zeros((N, 2, lmax), dtype=complex)
    for i in range(N):
        xx = mp.mpf(float(k[i] * radius[i]))
        m = mp.sqrt(mp.mpc(eps[i]) / mp.mpc(eps_b[i]))
        mt = m * mp.mpc(mu_b[i]) / mp.mpc(mu[i])
        z = m * xx
        for n in range(1, lmax + 1):
            pz, pz1 = psi(n, z), psi(n - 1, z)
            px, px1 = psi(n, xx), psi(n - 1, xx)
            xi, xi1 = px + 1j * chi(n, xx), px1 + 1j * chi(n - 1, xx)
            dpz, dpx, dxi = pz1 - n / z * pz, px1 - n / xx * px, xi1 - n / xx * xi
            a[i, 0, n - 1] = complex((mt * pz * dpx - px * dpz) / (mt * pz * dxi - xi * dpz))
            a[i, 1, n - 1] = complex((pz * dpx - mt * px * dpz) / (pz * dxi - mt * xi * dpz))
            w = m * (px * dxi - xi * dpx)
            c[i, 0, n - 1] = complex(w / (pz * dxi - mt * xi * dpz))
            c[i, 1, n - 1] = complex(w / (mt * pz * dxi - xi * dpz))
    np.savez(os.path.join(HERE, "mie_mpmath.npz"), radius=radius, eps=eps, mu=mu, eps_b=eps_b, mu_b=mu_b, k=k,
             lmax=lmax, mie_scat=a, mie_int=c)


def _np_array_literals(path):
    """Return {name: list} for module-level `name = np.array([...])` assignments."""
    tree = ast.parse(open(path).read())
    out = {}
    for node in tree.body:
        if isinstance(node, ast.Assign) and len(node.targets) == 1 and isinstance(node.targets[0], ast.Name):
            call = node.value
            if isinstance(call, ast.Call) and getattr(call.func, "attr", None) == "array" and call.args:
                try:
                    out[node.targets[0].id] = ast.literal_eval(call.args[0])
                except ValueError:
                    pass
    return out


def make_fdtd(reference):
    path = os.path.join(reference, "tests", "test_fdtd_comparison.py")
    arrs = _np_array_literals(path)
    nm = 1e-9
    wavelengths = np.linspace(400 * nm, 1000 * nm, 40)  # test_fdtd_comparison.py:296
    np.savez(os.path.join(HERE, "fdtd_au_dimer.npz"), wavelengths=wavelengths,
             eps=np.array(arrs["eps"], dtype=complex),
             scat=np.array(arrs["scat"]), absorb=np.array(arrs["absorb"]), extinct=np.array(arrs["scat"]) + np.array(arrs["absorb"]),  # :110
             Fx=np.array(arrs["Fx"]), Fy=np.array(arrs["Fy"]), Fz=np.array(arrs["Fz"]),
             radius=75 * nm, separation=400 * nm, lmax=2)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    ap.add_argument("--only", default="special,gaunt,mie,fdtd")
    args = ap.parse_args()
    todo = args.only.split(",")
    if "special" in todo:
        make_special()
    if "gaunt" in todo:
        make_gaunt()
    if "mie" in todo:
        make_mie()
    if "fdtd" in todo:
        make_fdtd(args.reference)
    print("ok")
