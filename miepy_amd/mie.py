"""Mie coefficients of spheres, all particles and orders in one device call.

Same formulas as src/miepy/mie_single/mie_sphere.py:16-76 with the
Riccati-Bessel functions of src/miepy/special_functions.py:127-153 (psi =
x j_n(x), xi = x h_n^(1)(x)), evaluated by the k_mie kernel instead of
2 N lmax scalar scipy calls (src/miepy/sphere_cluster.py:99-125).
"""

import numpy as np


def mie_coefficients(radius, eps, mu, eps_b, mu_b, k, lmax, conducting=None):
    """Return (mie_scat, mie_int), each [N, 2, lmax] complex128 (host numpy).

    radius, eps, mu: [N]; eps_b, mu_b: medium; k: medium wavenumber;
    conducting: [N] bool (perfect-conductor limit, mie_sphere.py:34-36, 67-69).
    k, eps_b and mu_b may also be [N] arrays (one value per particle), so that
    particles of many clusters and wavelengths go through one call.
    """
    from . import device as D
    ms, mi = D.mie_coefficients(radius, eps, mu, k, eps_b, mu_b, lmax, conducting)
    return ms.cpu().numpy(), mi.cpu().numpy()
