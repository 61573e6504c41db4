"""Pins the oracle's VSH translation and aggregate T-matrix assembly.

Reference: cpp/src/vsh_translation.cpp:12-237, 428-470;
cpp/src/interactions.cpp:370-462.

Pins:
- the symmetry identities of the reference's tests/test_vsh_translation.py:15-48
  (position reversal, (m,n)<->(u,v) swaps);
- the VSH addition theorem: the field radiated by particle j's outgoing
  expansion, evaluated near particle i with an independent numpy/scipy VSH
  evaluator (src/miepy/vsh/expansion.py:13-42, vsh_functions.py:40-78
  conventions), equals the incident-type expansion around i whose
  coefficients are -X_ij p_j, where X_ij is the (i, j) block of the assembled
  matrix with unit Mie coefficients.  This fixes every A/B coefficient, sign
  and index of the assembly independently of the oracle's formulas;
- assembled blocks vs the scalar vsh_translation entry point;
- the general T-matrix assembly reduces to the sphere assembly for diagonal
  (Mie) T-matrices;
- cluster coefficients (src/miepy/vsh/cluster_coefficients.py:14-53, the
  incident-mode translation): the outgoing field of the particles equals the
  outgoing expansion of p_cluster about the origin, converging with
  lmax_out; a particle at the origin contributes its coefficients directly.
"""

import numpy as np
import pytest

import vsh_fields as V

_modes = V.modes


# ---- reference symmetry identities -------------------------------------------------

def test_reference_symmetries(oracle):
    r_ji, theta_ji, phi_ji, lmax, k = 1, 0.7, 0.3, 2, 0.8
    for n, m in _modes(lmax):
        for v, u in _modes(lmax):
            A, B = oracle.vsh_translation(m, n, u, v, r_ji, theta_ji, phi_ji, k)
            A2, B2 = oracle.vsh_translation(m, n, u, v, r_ji, np.pi - theta_ji, np.pi + phi_ji, k)
            assert np.allclose(A2, (-1) ** (n + v) * A) and np.allclose(B2, (-1) ** (n + v + 1) * B)
            A2, B2 = oracle.vsh_translation(-u, v, -m, n, r_ji, theta_ji, phi_ji, k)
            assert np.allclose(A2, (-1) ** (m + u) * A) and np.allclose(B2, (-1) ** (m + u + 1) * B)
            A2, B2 = oracle.vsh_translation(-u, v, -m, n, r_ji, np.pi - theta_ji, np.pi + phi_ji, k)
            assert np.allclose(A2, (-1) ** (m + u + n + v) * A) and np.allclose(B2, (-1) ** (m + u + n + v) * B)


# ---- independent VSH evaluator (numpy/scipy): tests/vsh_fields.py ----------------

_expand = V.expand


@pytest.mark.parametrize("lmax,tol", [(2, 1e-5), (3, 3e-8), (4, 1e-10)])
def test_addition_theorem(oracle, lmax, tol):
    rng = np.random.default_rng(7 + lmax)
    rmax = lmax * (lmax + 2)
    k = 2 * np.pi / 0.9
    pos = np.array([[0.11, -0.07, 0.21], [0.52, 0.33, -0.18]])  # d ~ 0.66, kd ~ 4.6
    mie_one = np.ones((2, 2, lmax), dtype=complex)
    X = oracle.sphere_aggregate_tmatrix(pos, mie_one, k).reshape(2, 2, rmax, 2, 2, rmax)
    for (i, j) in [(0, 1), (1, 0)]:
        p_j = rng.normal(size=(2, rmax)) + 1j * rng.normal(size=(2, rmax))
        c_i = -np.einsum("arbs,bs->ar", X[i, :, :, j], p_j)
        dirs = rng.normal(size=(6, 3))
        dirs /= np.linalg.norm(dirs, axis=1)[:, None]
        rels = []
        for f in (0.004, 0.001):
            pts = pos[i] + f * np.linalg.norm(pos[0] - pos[1]) * dirs
            E_out = _expand(p_j, lmax, "outgoing", pts, pos[j], k)
            E_in = _expand(c_i, lmax, "incident", pts, pos[i], k)
            rels.append(np.linalg.norm(E_out - E_in) / np.linalg.norm(E_out))
        # truncating the regular expansion at lmax leaves an O(rho^lmax) error
        assert rels[1] < tol, rels
        assert rels[0] / rels[1] > 0.5 * 4 ** lmax, rels


def test_blocks_match_scalar_translation(oracle):
    lmax, k = 3, 2 * np.pi / 0.7
    rmax = lmax * (lmax + 2)
    pos = np.array([[0.0, 0.0, 0.0], [0.3, -0.4, 0.25], [-0.2, 0.5, -0.1]])
    rng = np.random.default_rng(3)
    mie = rng.normal(size=(3, 2, lmax)) + 1j * rng.normal(size=(3, 2, lmax))
    T = oracle.sphere_aggregate_tmatrix(pos, mie, k).reshape(3, 2, rmax, 3, 2, rmax)
    for (i, j) in [(0, 1), (2, 0), (1, 2)]:
        d = pos[i] - pos[j]
        rad = np.linalg.norm(d)
        th, ph = np.arccos(d[2] / rad), np.arctan2(d[1], d[0])
        for r, (n, m) in enumerate(_modes(lmax)):
            for s, (v, u) in enumerate(_modes(lmax)):
                A, B = oracle.vsh_translation(m, n, u, v, rad, th, ph, k)
                for a in range(2):
                    for b in range(2):
                        want = (A, B)[(a + b) % 2] * mie[j, b, v - 1]
                        got = T[i, a, r, j, b, s]
                        assert abs(got - want) <= 1e-12 * max(1e-3, abs(want)), (i, j, n, m, v, u, a, b, got, want)
    # diagonal blocks are zero (interactions.cpp:379)
    for i in range(3):
        assert np.all(T[i, :, :, i] == 0)


def test_single_particle_is_zero(oracle):
    T = oracle.sphere_aggregate_tmatrix(np.zeros((1, 3)), np.ones((1, 2, 2), dtype=complex), 1.0)
    assert T.shape == (16, 16) and np.all(T == 0)


def test_particle_tmatrix_reduces_to_sphere(oracle):
    lmax, k = 2, 2 * np.pi / 0.8
    rmax = lmax * (lmax + 2)
    pos = np.array([[0.0, 0.0, 0.0], [0.35, 0.1, -0.2], [-0.1, 0.4, 0.3]])
    rng = np.random.default_rng(11)
    mie = rng.normal(size=(3, 2, lmax)) + 1j * rng.normal(size=(3, 2, lmax))
    nidx = np.concatenate([[n - 1] * (2 * n + 1) for n in range(1, lmax + 1)])
    tm = np.zeros((3, 2, rmax, 2, rmax), dtype=complex)
    for i in range(3):
        for a in range(2):
            tm[i, a, np.arange(rmax), a, np.arange(rmax)] = mie[i, a, nidx]
    T1 = oracle.sphere_aggregate_tmatrix(pos, mie, k)
    T2 = oracle.particle_aggregate_tmatrix(pos, tm, k)
    assert np.max(np.abs(T1 - T2)) <= 1e-13 * np.max(np.abs(T1))


def test_particle_tmatrix_is_bare_blocks_times_tmatrix(oracle):
    """For dense T-matrices each block is the pair's bare translation block
    (the sphere assembly with unit Mie coefficients) times the column
    particle's T-matrix (cpp/src/vsh_translation.cpp:129-157)."""
    lmax, k = 2, 2 * np.pi / 0.7
    rmax = lmax * (lmax + 2)
    bs = 2 * rmax
    pos = np.array([[0.0, 0.0, 0.0], [0.35, 0.1, -0.2], [-0.1, 0.4, 0.3], [0.2, -0.3, 0.25]])
    N = pos.shape[0]
    rng = np.random.default_rng(12)
    tm = rng.normal(size=(N, bs, bs)) + 1j * rng.normal(size=(N, bs, bs))
    X = oracle.sphere_aggregate_tmatrix(pos, np.ones((N, 2, lmax), dtype=complex), k)
    T = oracle.particle_aggregate_tmatrix(pos, tm.reshape(N, 2, rmax, 2, rmax), k)
    for i in range(N):
        for j in range(N):
            blk = T[i * bs:(i + 1) * bs, j * bs:(j + 1) * bs]
            ref = X[i * bs:(i + 1) * bs, j * bs:(j + 1) * bs] @ tm[j]
            assert np.max(np.abs(blk - ref)) <= 1e-12 * max(np.max(np.abs(ref)), 1e-300)


@pytest.mark.parametrize("lmax_in,lmax_out,tol", [(1, 4, 1e-2), (1, 8, 3e-6), (2, 8, 3e-5)])
def test_cluster_coefficients_addition_theorem(oracle, lmax_in, lmax_out, tol):
    """The outgoing field of every particle's expansion equals the outgoing
    expansion of p_cluster about the origin outside a sphere holding the
    particles (the incident-mode translation re-expands outgoing waves); the
    truncation at lmax_out leaves an error falling like (kd)^n / (2n+1)!!.
    kd is about 1.2 here: the reference's upward recurrence for j_p
    (cpp/src/special.cpp:20-44) loses accuracy when kd is much below p, so
    smaller separations would test that recurrence rather than the sums."""
    import oracle.host as H
    rng = np.random.default_rng(20 + lmax_out)
    k = 2 * np.pi / 1.0
    origin = np.array([0.02, -0.01, 0.03])
    pos = origin + np.array([[0.15, 0.06, -0.09], [-0.12, 0.09, 0.15], [0.0, -0.18, 0.03]])
    rmax = lmax_in * (lmax_in + 2)
    p = rng.normal(size=(3, 2, rmax)) + 1j * rng.normal(size=(3, 2, rmax))
    pc = H.cluster_coefficients(pos, p, k, origin, lmax=lmax_out)
    dirs = rng.normal(size=(8, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    pts = origin + 1.5 * dirs
    E_parts = sum(_expand(p[i], lmax_in, "outgoing", pts, pos[i], k) for i in range(3))
    E_cluster = _expand(pc, lmax_out, "outgoing", pts, origin, k)
    assert np.linalg.norm(E_cluster - E_parts) < tol * np.linalg.norm(E_parts)


def test_cluster_coefficients_particle_at_origin(oracle):
    import oracle.host as H
    rng = np.random.default_rng(3)
    p = rng.normal(size=(1, 2, 8)) + 1j * rng.normal(size=(1, 2, 8))
    pc = H.cluster_coefficients(np.zeros((1, 3)), p, 3.0, np.zeros(3), lmax=3)
    assert np.array_equal(pc[:, :8], p[0]) and not np.any(pc[:, 8:])
