"""T-matrix of a rigid sphere cluster on the device
(src/miepy/tmatrix/common.py:199-364, the particle of
src/miepy/particles/sphere_cluster_particle.py) against the oracle's
restatement, and the reference's own tests of it.

Bars: the source matrix S equals the oracle's to 1e-12 of its largest
entry (lmax_p, lmax_out in 1..6, a sphere at the origin); the T-matrix equals
the oracle's (numpy solve) to 1e-10 of its largest entry, with per-sphere
lmax and lmax_cluster above and below the spheres' lmax; the BiCGSTAB
method equals the LU one to 1e-9; reciprocity
T[a,(n,m)][b,(v,u)] = (-1)^(m+u+n+v) T[b,(v,-u)][a,(n,-m)] to 1e-11 on a
30-sphere cluster; the reference's tests (tests/test_tmatrix.py:111-139,
tests/test_tmatrix_validation.py:1411-1485): a sphere_cluster_particle in
`cluster` gives the cross-sections of `sphere_cluster` (rtol 7e-4 at lmax 6,
1e-3 at lmax 4) and its cluster coefficients (rtol 1e-2), and a three-sphere
particle satisfies ext = scat + abs to 1e-3.  The reference's silver comes
from its material database; a constant silver-like permittivity at 800 nm
stands in for it.
"""

import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

nm = 1e-9
AG800 = -29.0 + 0.4j


@pytest.fixture(scope="module")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    import miepy_amd.device as D
    D.require_gpu()
    return D


def _reciprocity_error(T):
    rmax = T.shape[1]
    lc = int(round((1 + rmax) ** 0.5)) - 1
    modes = [(n, m) for n in range(1, lc + 1) for m in range(-n, n + 1)]
    idx = {nm_: r for r, nm_ in enumerate(modes)}
    sgn = np.array([(-1) ** (n + m) for (n, m) in modes])
    flip = np.array([idx[(n, -m)] for (n, m) in modes])
    # (-1)^(m+u+n+v) T[b, flip(s), a, flip(r)] for every (a, r, b, s)
    Tt = np.transpose(T[:, flip][:, :, :, flip], (2, 3, 0, 1))
    return np.abs(T - sgn[None, :, None, None] * sgn[None, None, None, :] * Tt).max()


@pytest.mark.parametrize("N,lmax_p,lmax_out", [(1, 1, 1), (3, 3, 2), (4, 2, 5), (5, 4, 4), (2, 6, 6)])
def test_source_matrix_matches_oracle(dev, oracle, N, lmax_p, lmax_out):
    import oracle.host as H
    rng = np.random.default_rng(N + 10 * lmax_p)
    k = 2 * np.pi / 1.0
    dirs = rng.normal(size=(N, 3))
    pos = dirs / np.linalg.norm(dirs, axis=1)[:, None] * rng.uniform(2.0, 3.0, size=(N, 1))
    pos[0] = 0.0 if N > 2 else pos[0]  # a sphere at the origin
    got = dev.cluster_source(pos, k, np.zeros(3), lmax_p, lmax_out).cpu().numpy()
    ref = H.cluster_source_matrix(pos, np.zeros(3), lmax_p, lmax_out, k)
    assert got.shape == ref.T.shape
    assert np.max(np.abs(got - ref.T)) <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("pos,radii,lmax,lmax_cluster,eps", [
    ([[0.95, 0.1, -0.15], [-0.85, 0.25, 0.3]], 0.08, 3, 4, 4 + 0.1j),
    ([[1.5, 0.0, 0.0], [-1.25, 0.5, 0.0], [0.0, 1.5, -0.5], [0.25, -1.0, 1.25]], [0.06, 0.08, 0.05, 0.07], [2, 3, 1, 3],
     3, [2.56, -15.8 + 1.07j, 4.0, 2.0 + 0.1j]),
    ([[0.0, 0.0, 0.0]], 0.1, 3, 5, 6.0),
    ([[0.0, 0.0, 0.0], [0.32, 0.0, 0.05], [-0.1, -0.3, 0.0]], 0.09, 2, 2, -15.8 + 1.07j),
    ([[1.05, -0.5, 0.0], [-1.05, 0.5, 0.0]], 0.1, 4, 2, 3.0),
])
def test_tmatrix_matches_oracle(dev, oracle, pos, radii, lmax, lmax_cluster, eps):
    """Spheres a wavelength or more from the origin except at low order: the
    reference's upward recurrence for j_p (cpp/src/special.cpp:20-44), kept
    in S and R, amplifies rounding when k |r_i| is far below p, and two
    correct evaluations then differ in those terms."""
    import oracle.host as H
    from miepy_amd import tmatrix
    got = tmatrix.tmatrix_sphere_cluster(pos, radii, lmax, lmax_cluster, 1.0, eps, 1.7)
    ref = H.tmatrix_sphere_cluster(pos, radii, lmax, lmax_cluster, 1.0, eps, 1.7)
    assert got.shape == ref.shape
    assert np.max(np.abs(got - ref)) <= 1e-10 * np.abs(ref).max()


def test_bicgstab_method_and_reciprocity(dev):
    from miepy_amd import tmatrix
    rng = np.random.default_rng(4)
    pos = []
    while len(pos) < 30:  # spheres of radius 60 nm, centres at least 150 nm apart
        p = rng.uniform(-0.6e-6, 0.6e-6, size=3)
        if all(np.linalg.norm(p - q) > 150 * nm for q in pos):
            pos.append(p)
    pos = np.array(pos)
    args = (pos, 60 * nm, 3, 4, 700 * nm, 3.5 + 0.05j, 1.0)
    T_lu = tmatrix.tmatrix_sphere_cluster(*args)
    T_it = tmatrix.tmatrix_sphere_cluster(*args, method="bicgstab", tol=1e-13)
    assert np.max(np.abs(T_it - T_lu)) <= 1e-9 * np.abs(T_lu).max()
    assert _reciprocity_error(T_lu) <= 1e-11 * np.abs(T_lu).max()


def test_reference_sphere_cluster_particle(dev):
    """tests/test_tmatrix.py:111-139 of the reference."""
    import miepy_amd as M
    L = 155 * nm
    lmax = 6
    Ag = M.constant_material(eps=AG800)
    medium = M.constant_material(index=1.5 ** 2)
    source = M.plane_wave([1, 0])
    cl = M.sphere_cluster(position=[[-L / 2, 0, 0], [L / 2, 0, 0]], material=Ag, radius=75 * nm, source=source,
                          lmax=lmax, medium=medium, wavelength=800 * nm)
    C1 = cl.cross_sections()
    cl.solve_cluster_coefficients()
    p1 = cl.p_cluster
    particle = M.sphere_cluster_particle(cl.position, cl.radius, Ag, lmax=lmax)
    c2 = M.cluster(particles=particle, source=source, lmax=lmax, medium=medium, wavelength=800 * nm)
    C2 = c2.cross_sections()
    p2 = c2.p_scat
    assert np.allclose(C1, C2, rtol=7e-4, atol=0), (C1, C2)
    assert np.allclose(p1, p2, rtol=1e-2, atol=1e-12)


def test_reference_tmatrix_validation_cases(dev):
    """tests/test_tmatrix_validation.py:1414-1485 of the reference."""
    import miepy_amd as M
    Ag = M.constant_material(eps=AG800)
    medium = M.constant_material(eps=1.5)
    source = M.plane_wave([1, 0])
    L = 155 * nm
    direct = M.sphere_cluster(position=[[-L / 2, 0, 0], [L / 2, 0, 0]], material=Ag, radius=75 * nm, source=source,
                              lmax=4, medium=medium, wavelength=800 * nm)
    particle = M.sphere_cluster_particle(direct.position, direct.radius, Ag, lmax=4)
    via = M.cluster(particles=particle, source=source, lmax=4, medium=medium, wavelength=800 * nm)
    assert np.allclose(direct.cross_sections(), via.cross_sections(), rtol=1e-3, atol=0)
    sep = 200 * nm
    three = M.sphere_cluster_particle([[-sep, 0, 0], [0, 0, 0], [sep, 0, 0]], [60 * nm, 75 * nm, 60 * nm], Ag, lmax=3)
    c = M.cluster(particles=three, source=source, lmax=3, medium=medium, wavelength=600 * nm).cross_sections()
    assert c.scattering > 0 and c.extinction > 0
    assert abs(c.extinction - c.scattering - c.absorption) / abs(c.extinction) < 1e-3
    with pytest.raises(ValueError):
        M.sphere_cluster_particle([[0, 0, 0]], 50 * nm, Ag, lmax=2)


def test_large_cluster_tmatrix_timing(dev):
    """200 spheres, lmax 3, lmax_cluster 6: n = 6000 and 96 right-hand sides."""
    import torch
    from miepy_amd import tmatrix
    from miepy_amd.workloads import random_cluster
    rng = np.random.default_rng(8)
    pos = random_cluster(200, 3e-6, 100 * nm, rng)
    pos -= pos.mean(axis=0)
    args = (pos, 100 * nm, 3, 6, 600 * nm, 3.0 + 0.02j, 1.0)
    tmatrix.tmatrix_sphere_cluster(*args)
    for method in ("exact", "bicgstab"):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        T = tmatrix.tmatrix_sphere_cluster(*args, method=method, tol=1e-10)
        dt = time.perf_counter() - t0
        print(f"tmatrix_sphere_cluster N=200 lmax 3 -> 6 ({method}): {dt * 1e3:.1f} ms")
        assert np.all(np.isfinite(T))
        assert _reciprocity_error(T) <= 1e-8 * np.abs(T).max()
