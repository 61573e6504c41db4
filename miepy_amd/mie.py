"""Mie coefficients of spheres, vectorized over particles and orders.

Same formulas as src/miepy/mie_single/mie_sphere.py:16-76 with the
Riccati-Bessel functions of src/miepy/special_functions.py:127-153
(psi = x j_n(x), xi = x h_n^(1)(x), derivatives by the product rule), but
evaluated in one scipy call per order over all particles instead of
2 N lmax scalar calls (src/miepy/sphere_cluster.py:99-125).
"""

import numpy as np
from scipy import special


def _riccati_1(n, x):
    jn = special.spherical_jn(n, x)
    jnp = special.spherical_jn(n, x, derivative=True)
    return x * jn, jn + x * jnp


def _riccati_3(n, x):
    j0, j1 = _riccati_1(n, x)
    yn = special.spherical_yn(n, x)
    ynp = special.spherical_yn(n, x, derivative=True)
    return j0 + 1j * (x * yn), j1 + 1j * (yn + x * ynp)


def mie_coefficients(radius, eps, mu, eps_b, mu_b, k, lmax, conducting=None):
    """Return (mie_scat, mie_int), each [N, 2, lmax] complex128.

    radius, eps, mu: [N]; eps_b, mu_b: medium; k: medium wavenumber;
    conducting: [N] bool (perfect-conductor limit, mie_sphere.py:34-36, 67-69).
    k, eps_b and mu_b may also be [N] arrays (one value per particle), so that
    particles of many clusters and wavelengths go through one call.
    """
    radius = np.asarray(radius, dtype=float)
    eps = np.asarray(eps, dtype=complex)
    mu = np.asarray(mu, dtype=complex)
    N = radius.shape[0]
    if conducting is None:
        conducting = np.zeros(N, dtype=bool)
    conducting = np.asarray(conducting, dtype=bool)
    x = k * radius
    m = (eps / eps_b) ** 0.5
    mt = m * mu_b / mu
    mie_scat = np.zeros((N, 2, lmax), dtype=complex)
    mie_int = np.zeros((N, 2, lmax), dtype=complex)
    for n in range(1, lmax + 1):
        jn0, jn1 = _riccati_1(n, x)
        yn0, yn1 = _riccati_3(n, x)
        jm0, jm1 = _riccati_1(n, m * x)
        with np.errstate(all="ignore"):
            a = (mt * jm0 * jn1 - jn0 * jm1) / (mt * jm0 * yn1 - yn0 * jm1)
            b = (jm0 * jn1 - mt * jn0 * jm1) / (jm0 * yn1 - mt * yn0 * jm1)
            c = (m * jn0 * yn1 - m * yn0 * jn1) / (jm0 * yn1 - mt * yn0 * jm1)
            d = (m * jn0 * yn1 - m * yn0 * jn1) / (mt * jm0 * yn1 - yn0 * jm1)
        if conducting.any():
            a = np.where(conducting, jn1 / yn1, a)
            b = np.where(conducting, jn0 / yn0, b)
            c = np.where(conducting, 0.0, c)
            d = np.where(conducting, 0.0, d)
        mie_scat[:, 0, n - 1] = a
        mie_scat[:, 1, n - 1] = b
        mie_int[:, 0, n - 1] = c
        mie_int[:, 1, n - 1] = d
    return mie_scat, mie_int
