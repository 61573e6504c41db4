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
    ap.add_argument("--only", default="special,gaunt,fdtd")
    args = ap.parse_args()
    todo = args.only.split(",")
    if "special" in todo:
        make_special()
    if "gaunt" in todo:
        make_gaunt()
    if "fdtd" in todo:
        make_fdtd(args.reference)
    print("ok")
