"""Cluster coefficients and per-multipole cluster cross-sections on the
device (src/miepy/vsh/cluster_coefficients.py:14-53, cpp/src/flux.cpp:32-57).

Bars: p_cluster equals the oracle's restatement (a loop over the scalar
incident-mode vsh_translation, itself pinned by the addition theorem) to
1e-11 of its largest entry for lmax_in, lmax_out in 1..6, one to twenty
particles, one at the origin, several sources; the cross-sections equal the
oracle's flux restatement to 1e-13; the reference's own tests
(tests/test_cluster_coefficients.py): for a monomer and a dimer the
per-multipole scattering summed over multipoles equals the cluster's total
scattering (its bars 1e-4 and 1e-3); `cluster` of spheres equals
`sphere_cluster`.  Separations are a wavelength or more: the reference's
upward recurrence for j_p (cpp/src/special.cpp:20-44), kept here, amplifies
rounding when kd is far below p, and two correct evaluations then differ in
those terms.
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


@pytest.mark.parametrize("N,lmax_in,lmax_out,S", [(1, 2, 2, 1), (3, 2, 4, 1), (5, 3, 2, 2), (8, 1, 6, 1),
                                                  (6, 4, 4, 1), (20, 2, 3, 3), (4, 6, 6, 1)])
def test_cluster_coefficients_match_oracle(dev, oracle, N, lmax_in, lmax_out, S):
    import oracle.host as H
    rng = np.random.default_rng(N * 10 + lmax_out)
    k = 2 * np.pi / 1.0
    origin = np.array([0.1, -0.2, 0.05])
    pos = origin + rng.uniform(-3, 3, size=(N, 3))
    if N > 1 and lmax_out >= lmax_in:
        pos[0] = origin  # one particle exactly at the origin (the reference raises for lmax_out < lmax_in)
    rmax_in = lmax_in * (lmax_in + 2)
    p = rng.normal(size=(S, N, 2, rmax_in)) + 1j * rng.normal(size=(S, N, 2, rmax_in))
    got = dev.cluster_coefficients(pos, p, k, origin, lmax_out).cpu().numpy()
    for s in range(S):
        ref = H.cluster_coefficients(pos, p[s], k, origin, lmax=lmax_out)
        assert np.max(np.abs(got[s] - ref)) <= 1e-11 * np.abs(ref).max()


def test_particle_at_origin_truncates(dev):
    """A particle at the origin adds its coefficients up to lmax_out (the
    reference's slice assignment raises when lmax_out < lmax_in)."""
    rng = np.random.default_rng(6)
    p = rng.normal(size=(1, 2, 2, 15)) + 1j * rng.normal(size=(1, 2, 2, 15))
    pos = np.array([[0.0, 0.0, 0.0], [1.7, 0.4, -0.9]])
    both = dev.cluster_coefficients(pos, p, 6.0, np.zeros(3), 2).cpu().numpy()[0]
    other = dev.cluster_coefficients(pos[1:], p[:, 1:], 6.0, np.zeros(3), 2).cpu().numpy()[0]
    assert np.allclose(both - other, p[0, 0, :, :8], rtol=0, atol=1e-13 * np.abs(both).max())


def test_cluster_cross_sections_match_oracle(dev, oracle):
    rng = np.random.default_rng(5)
    for lmax in (1, 3, 6):
        rmax = lmax * (lmax + 2)
        pc = rng.normal(size=(2, 2, rmax)) + 1j * rng.normal(size=(2, 2, rmax))
        p0 = rng.normal(size=(2, 2, rmax)) + 1j * rng.normal(size=(2, 2, rmax))
        k = 7.3e6
        C = dev.cluster_cross_sections(dev.to_device(pc, dev.C128), p0, k).cpu().numpy()
        for s in range(2):
            ref = oracle.cluster_cross_sections(pc[s], p0[s], k)
            for q in range(3):
                assert np.max(np.abs(C[s, q] - ref[q])) <= 1e-13 * np.abs(ref[q]).max()


def test_reference_monomer_and_dimer(dev):
    """tests/test_cluster_coefficients.py:12-83 of the reference."""
    import miepy_amd as M
    wavelengths = np.linspace(300 * nm, 1000 * nm, 5)
    mat = M.constant_material(3.7 ** 2)
    src = M.plane_wave.from_string(polarization="x")
    C1 = np.zeros(5)
    C2 = np.zeros((5, 2, 2))
    for i, wl in enumerate(wavelengths):
        c = M.sphere_cluster(position=[0, 0, 0], radius=75 * nm, material=mat, source=src, lmax=2, wavelength=wl)
        C1[i] = c.cross_sections().scattering
        C2[i] = c.cross_sections_per_multipole().scattering
    C2_sum = np.sum(C2, axis=(1, 2))
    assert np.linalg.norm(C1 - C2_sum) / 5 < 1e-4 * np.average(C1 + C2_sum) / 2
    sep = 200 * nm
    C1 = np.zeros(5)
    C2 = np.zeros((5, 2, 4))
    for i, wl in enumerate(wavelengths):
        c = M.sphere_cluster(position=[[-sep / 2, 0, 0], [sep / 2, 0, 0]], radius=75 * nm, material=mat, source=src,
                             lmax=2, wavelength=wl)
        C1[i] = c.cross_sections().scattering
        C2[i] = c.cross_sections_per_multipole(lmax=4).scattering
    C2_sum = np.sum(C2, axis=(1, 2))
    assert np.linalg.norm(C1 - C2_sum) / 5 < 1e-3 * np.average(C1 + C2_sum) / 2


def test_lossless_cluster_multipoles_absorb_nothing(dev):
    """Cext = Cscat per multipole sum for lossless spheres: the extinction of
    the cluster expansion equals its scattering (energy balance), up to the
    truncation of the cluster expansion."""
    import miepy_amd as M
    c = M.sphere_cluster(position=[[-100 * nm, 0, 0], [100 * nm, 0, 0]], radius=75 * nm,
                         material=M.constant_material(2.56), source=M.plane_wave.from_string(polarization="y"),
                         lmax=3, wavelength=700 * nm)
    C = c.cross_sections_per_multipole(lmax=6)
    tot = c.cross_sections()
    assert abs(np.sum(C.absorption)) < 1e-4 * tot.scattering
    assert abs(np.sum(C.extinction) - tot.extinction) < 1e-4 * tot.extinction


def test_multi_source_and_cluster_class(dev):
    import miepy_amd as M
    pos = np.array([[0, 0, 0], [300 * nm, 0, 0], [0, 320 * nm, 40 * nm]])
    mat = M.constant_material(eps=3.6 + 0.2j)
    srcs = [M.plane_wave.from_string(polarization="x"), M.plane_wave.from_string(polarization="y")]
    kw = dict(position=pos, radius=70 * nm, material=mat, wavelength=620 * nm, lmax=2, origin="auto", tol=1e-12)
    multi = M.sphere_cluster(source=srcs, **kw)
    Cm = multi.cross_sections_per_multipole(lmax=4)
    for s, src in enumerate(srcs):
        one = M.sphere_cluster(source=src, **kw)
        C1 = one.cross_sections_per_multipole(lmax=4)
        assert np.allclose(multi.p_cluster[s], one.p_cluster, rtol=0, atol=1e-10 * np.abs(one.p_cluster).max())
        assert np.allclose(Cm.scattering[s], C1.scattering, rtol=1e-9, atol=0)
    cl = M.cluster(particles=[M.sphere(p, 70 * nm, mat) for p in pos], source=srcs[0], wavelength=620 * nm, lmax=2,
                   origin="auto", tol=1e-12)
    Cc = cl.cross_sections_per_multipole(lmax=4)
    one = M.sphere_cluster(source=srcs[0], **kw)
    C1 = one.cross_sections_per_multipole(lmax=4)
    for q in range(3):
        assert np.allclose(Cc[q], C1[q], rtol=1e-9, atol=1e-12 * np.abs(C1.extinction).max())


@pytest.mark.parametrize("mode", ["outgoing", "ingoing", "incident", "interior"])
def test_vsh_translation_matches_oracle(dev, oracle, mode):
    """The scalar vsh_translation and its numpy form (cpp/src/vsh_translation.cpp:366-470)
    on the device vs the oracle's restatement, every (m, n, u, v) up to lmax 4,
    at two translation vectors (kd about 4 and 9), 1e-12 of the largest entry."""
    import miepy_amd as M
    from miepy_amd import backend as B
    lmax = 4
    tuples = [(m, n, u, v) for n in range(1, lmax + 1) for m in range(-n, n + 1)
              for v in range(1, lmax + 1) for u in range(-v, v + 1)]
    k = 2 * np.pi / 1.0
    geos = [(0.65, 0.7, 0.3), (1.4, 2.4, -2.0)]
    md = getattr(M.vsh_mode, mode)
    for rad, th, ph in geos:
        got = np.array([M.vsh_translation(m, n, u, v, rad, th, ph, k, md) for (m, n, u, v) in tuples[::23]])
        ref = np.array([oracle.vsh_translation(m, n, u, v, rad, th, ph, k, mode) for (m, n, u, v) in tuples[::23]])
        assert np.max(np.abs(got - ref)) <= 1e-12 * np.abs(ref).max()
        allv = dev.vsh_translation(tuples, rad, th, ph, k, md.value)[0]
        refv = np.array([oracle.vsh_translation(*t, rad, th, ph, k, mode) for t in tuples])
        assert np.max(np.abs(allv - refv)) <= 1e-12 * np.abs(refv).max()
    rads, ths, phs = (np.array(a) for a in zip(*geos))
    got = B.vsh_translation_numpy(2, 3, -1, 2, rads, ths, phs, k, md)
    assert got.shape == (2, 2)
    for g, (rad, th, ph) in enumerate(geos):
        A, Bv = oracle.vsh_translation(2, 3, -1, 2, rad, th, ph, k, mode)
        assert abs(got[0, g] - A) <= 1e-12 * abs(A) + 1e-15 and abs(got[1, g] - Bv) <= 1e-12 * abs(Bv) + 1e-15
