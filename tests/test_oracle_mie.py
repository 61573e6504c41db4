"""Pins the oracle's Mie coefficients (src/miepy/mie_single/mie_sphere.py:16-76
restated with scipy Bessel functions) against 30-digit values
(tests/golden/mie_mpmath.npz): sizes 0.01-40, dielectric, absorbing,
plasmonic, metallic and magnetic spheres, three media.  The scipy route
loses a few digits for the tiny sizes (about 1e-10 relative at x = 0.01),
hence the 1e-9 bar; elsewhere it is within 1e-13."""

import os

import numpy as np


def test_oracle_mie_vs_mpmath(oracle, golden_dir):
    import oracle.host as H
    g = np.load(os.path.join(golden_dir, "mie_mpmath.npz"))
    lmax = int(g["lmax"])
    for i in range(g["radius"].shape[0]):
        a, c = H.mie_coefficients(g["radius"][i:i + 1], g["eps"][i:i + 1], g["mu"][i:i + 1], g["eps_b"][i],
                                  g["mu_b"][i], g["k"][i], lmax)
        for got, ref in ((a[0], g["mie_scat"][i]), (c[0], g["mie_int"][i])):
            scale = np.abs(ref).max(axis=-1, keepdims=True)
            err = np.abs(got - ref)
            assert np.all((err <= 1e-9 * np.abs(ref)) | (err <= 1e-13 * scale)), (i, np.max(err / np.abs(ref)))
