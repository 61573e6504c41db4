"""Clusters of particles with general T-matrices (the reference's `cluster`
path): particle_aggregate_tmatrix on the device vs the oracle's
restatement (dense random T-matrices, lmax 1-5, 1e-12 relative to the
largest entry), the block-diagonal product p_scat = T p_inc vs numpy, the
matrix-free operator with T-matrices vs the stored matrix (1e-12), and the
`cluster` class vs `sphere_cluster` for spheres (cross-sections and forces
1e-10) and vs the oracle's LU solve for dense T-matrices (1e-9)."""

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


def _geometry(N, seed):
    from miepy_amd.workloads import random_cluster
    rng = np.random.default_rng(seed)
    return random_cluster(N, 350 * nm * N ** (1 / 3) * 1.5, 75 * nm, rng), rng


def _random_tmatrices(rng, N, lmax, scale=0.05):
    rmax = lmax * (lmax + 2)
    bs = 2 * rmax
    tm = scale * (rng.normal(size=(N, bs, bs)) + 1j * rng.normal(size=(N, bs, bs)))
    return tm.reshape(N, 2, rmax, 2, rmax)


@pytest.mark.parametrize("N,lmax", [(3, 1), (6, 2), (7, 3), (5, 4), (3, 5)])
def test_particle_aggregate_matches_oracle(dev, oracle, N, lmax):
    pos, rng = _geometry(N, 60 + lmax)
    k = 2 * np.pi / (600 * nm)
    tm = _random_tmatrices(rng, N, lmax)
    ref = oracle.particle_aggregate_tmatrix(pos, tm, k)
    scale = np.abs(ref).max()
    got = dev.assemble_tmatrix(pos, tm, k).cpu().numpy()
    assert np.max(np.abs(got - ref)) <= 1e-12 * scale
    got_i = dev.assemble_tmatrix(pos, tm, k, add_identity=True).cpu().numpy()
    assert np.max(np.abs(got_i - ref - np.eye(ref.shape[0]))) <= 1e-12 * scale
    got32 = dev.assemble_tmatrix(pos, tm, k, dtype=dev.C64).cpu().numpy()
    assert np.max(np.abs(got32 - ref)) <= 1e-6 * scale


def test_sphere_tmatrices_equal_sphere_assembly(dev):
    N, lmax = 9, 3
    pos, rng = _geometry(N, 5)
    k = 2 * np.pi / (700 * nm)
    mie = rng.normal(size=(N, 2, lmax)) + 1j * rng.normal(size=(N, 2, lmax))
    rmax = lmax * (lmax + 2)
    nidx = np.concatenate([[n - 1] * (2 * n + 1) for n in range(1, lmax + 1)])
    tm = np.zeros((N, 2, rmax, 2, rmax), dtype=complex)
    for a in range(2):
        tm[:, a, np.arange(rmax), a, np.arange(rmax)] = mie[:, a, nidx]
    A1 = dev.assemble(pos, mie, k, add_identity=True).cpu().numpy()
    A2 = dev.assemble_tmatrix(pos, tm, k, add_identity=True).cpu().numpy()
    assert np.max(np.abs(A1 - A2)) <= 1e-13 * np.abs(A1).max()


def test_tmatrix_apply(dev):
    rng = np.random.default_rng(2)
    N, lmax, S = 11, 2, 3
    tm = _random_tmatrices(rng, N, lmax, scale=1.0)
    rmax = lmax * (lmax + 2)
    x = rng.normal(size=(S, N, 2, rmax)) + 1j * rng.normal(size=(S, N, 2, rmax))
    ref = np.einsum("naibj,snbj->snai", tm, x)
    got = dev.tmatrix_apply(tm, x).cpu().numpy()
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("N,lmax,S", [(8, 2, 1), (12, 3, 2), (40, 2, 5)])
def test_matfree_tmatrix_matches_stored(dev, N, lmax, S):
    pos, rng = _geometry(N, 70 + N)
    k = 2 * np.pi / (650 * nm)
    tm = _random_tmatrices(rng, N, lmax)
    n = 2 * lmax * (lmax + 2) * N
    A = dev.assemble_tmatrix(pos, tm, k, add_identity=True)
    X = dev.to_device(rng.normal(size=(S, n)) + 1j * rng.normal(size=(S, n)), dev.C128)
    ref = (X.cpu().numpy() @ A.cpu().numpy().T)
    op = dev.ClusterOperator(pos, None, k, max_rhs=S, tmatrix=tm)
    got = dev.matfree_matvec(op, X, add_identity=True).cpu().numpy()
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("mode", ["stored", "matfree"])
def test_cluster_of_spheres_equals_sphere_cluster(dev, mode):
    """The reference's tests/test_tmatrix.py:26-51 (cluster of spheres ==
    sphere_cluster, atol 4e-16 there): 1e-10 relative here, the two paths
    summing the same terms in a different order."""
    import miepy_amd as M
    pos = np.array([[0, 0, 0], [300 * nm, 0, 0], [0, 320 * nm, 40 * nm], [-150 * nm, -200 * nm, 100 * nm]])
    mat = M.constant_material(eps=3.6 + 0.2j)
    src = M.plane_wave.from_string(polarization="x")
    kw = dict(source=src, wavelength=620 * nm, lmax=3, medium=M.constant_material(index=1.33), tol=1e-12)
    sc = M.sphere_cluster(position=pos, radius=70 * nm, material=mat, **kw)
    cl = M.cluster(particles=[M.sphere(p, 70 * nm, mat) for p in pos], solver_mode=mode, **kw)
    np.testing.assert_allclose(cl.p_scat, sc.p_scat, rtol=0, atol=1e-10 * np.abs(sc.p_scat).max())
    np.testing.assert_allclose(np.array(cl.cross_sections()), np.array(sc.cross_sections()), rtol=1e-10)
    F1, F2 = cl.force(), sc.force()
    assert np.max(np.abs(F1 - F2)) <= 1e-10 * np.abs(F2).max()
    T1, T2 = cl.torque(), sc.torque()
    assert np.max(np.abs(T1 - T2)) <= 1e-9 * np.abs(T2).max() + 1e-40


@pytest.mark.parametrize("mode", ["stored", "matfree"])
def test_cluster_dense_tmatrices_match_oracle_lu(dev, oracle, mode):
    import miepy_amd as M
    N, lmax = 6, 2
    pos, rng = _geometry(N, 90)
    rmax = lmax * (lmax + 2)
    tm = _random_tmatrices(rng, N, lmax, scale=0.1)
    src = M.plane_wave.from_string(polarization="y")
    cl = M.cluster(particles=[M.tmatrix_particle(p, t) for p, t in zip(pos, tm)], source=src,
                   wavelength=600 * nm, lmax=lmax, solver_mode=mode, tol=1e-12)
    k = 2 * np.pi / (600 * nm)
    T = oracle.particle_aggregate_tmatrix(pos, tm, k)
    p_inc = oracle.lu_solve(np.eye(T.shape[0]) + T, cl.p_src.reshape(-1)).reshape(N, 2, rmax)
    assert np.linalg.norm(cl.p_inc - p_inc) <= 1e-9 * np.linalg.norm(p_inc)
    p_scat = np.einsum("naibj,nbj->nai", tm, p_inc)
    assert np.linalg.norm(cl.p_scat - p_scat) <= 1e-9 * np.linalg.norm(p_scat)
