"""Pins the oracle's restatement of the sphere-cluster T-matrix
(src/miepy/tmatrix/common.py:199-364: source matrix S, re-expansion matrix
R, T = R diag(mie) solve(I + T_agg, S)).

Pins, each independent of the oracle's formulas:
- S by the VSH addition theorem: origin mode (b, n, m) as a regular field
  equals the regular expansion about each sphere with S's column as
  coefficients (independent numpy/scipy VSH evaluator), the truncation error
  falling as rho^lmax_p near the sphere;
- R equals cluster_coefficients (itself pinned by the addition theorem) on
  random coefficients;
- a single sphere at the origin: T = diag(a_n, b_n) exactly, zero-padded or
  truncated to lmax_cluster;
- the cluster T-matrix applied to a plane wave's expansion about the origin
  equals the cluster coefficients of the direct cluster solve, converging
  with lmax_cluster;
- reciprocity: T[a, (n, m)][b, (v, u)] = (-1)^(m+u+n+v) T[b, (v, -u)][a, (n, -m)]
  (the i^n of the mode normalization gives the (-1)^(n+v)), which a wrong
  sign, index or argument order in S relative to R breaks;
- per-sphere lmax: a lower lmax changes the T-matrix and keeps it reciprocal
  (the zeroed Mie coefficients themselves are not pinned further).
"""

import numpy as np
import pytest

from vsh_fields import expand as _expand

K = 2 * np.pi
POS = np.array([[0.19, 0.02, -0.03], [-0.17, 0.05, 0.06]])  # k |r_i| about 1.2


@pytest.fixture(scope="module")
def H(oracle):
    import oracle.host as host
    return host


@pytest.mark.parametrize("lmax_out", [2, 4])
def test_source_matrix_addition_theorem(H, lmax_out):
    lmax_p = 3
    rmax_p, rmax_out = lmax_p * (lmax_p + 2), lmax_out * (lmax_out + 2)
    S = H.cluster_source_matrix(POS, np.zeros(3), lmax_p, lmax_out, K)
    rng = np.random.default_rng(lmax_out)
    dirs = rng.normal(size=(6, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    for i in range(2):
        for col in (0, 3, rmax_out + 5, 2 * rmax_out - 1):
            e = np.zeros(2 * rmax_out, dtype=complex)
            e[col] = 1.0
            rels = []
            for f in (0.02, 0.005):
                pts = POS[i] + f * dirs
                E0 = _expand(e.reshape(2, rmax_out), lmax_out, "incident", pts, np.zeros(3), K)
                E1 = _expand(S[i * 2 * rmax_p:(i + 1) * 2 * rmax_p, col].reshape(2, rmax_p), lmax_p, "incident", pts,
                             POS[i], K)
                rels.append(np.linalg.norm(E0 - E1) / np.linalg.norm(E0))
            assert rels[1] < 1e-5, (i, col, rels)
            assert rels[0] / rels[1] > 0.5 * 4 ** lmax_p, (i, col, rels)


def test_reexpansion_matrix_is_cluster_coefficients(H):
    rng = np.random.default_rng(2)
    R = H.cluster_reexpansion_matrix(POS, np.zeros(3), 3, 5, K)
    p = rng.normal(size=(2, 2, 15)) + 1j * rng.normal(size=(2, 2, 15))
    pc = H.cluster_coefficients(POS, p, K, np.zeros(3), lmax=5)
    assert np.max(np.abs(R @ p.reshape(-1) - pc.reshape(-1))) <= 1e-14 * np.abs(pc).max()


@pytest.mark.parametrize("lmax_cluster", [2, 3, 5])
def test_single_sphere_at_origin_is_mie(H, lmax_cluster):
    lmax = 3
    T = H.tmatrix_sphere_cluster(np.zeros((1, 3)), 0.08, lmax, lmax_cluster, 1.0, 4.0 + 0.2j, 1.0)
    rmax = lmax_cluster * (lmax_cluster + 2)
    assert T.shape == (2, rmax, 2, rmax)
    ref = np.zeros_like(T)
    for r, n in enumerate(H.n_index(lmax_cluster) + 1):
        if n <= lmax:
            a, b = H.mie_sphere_scattering_coefficients(0.08, n, 4.0 + 0.2j, 1, 1.0, 1, K)
            ref[0, r, 0, r], ref[1, r, 1, r] = a, b
    assert np.array_equal(T, ref)


def test_plane_wave_matches_direct_cluster_solve(H):
    errs = []
    ref = H.solve_cluster(POS, 0.08, 4 + 0.1j, 1.0, 3, method="exact")
    for lc in (4, 6, 8):
        T = H.tmatrix_sphere_cluster(POS, 0.08, 3, lc, 1.0, 4 + 0.1j, 1.0)
        rmax = lc * (lc + 2)
        pcl = H.cluster_coefficients(POS, ref["p_scat"], K, np.zeros(3), lmax=lc)
        p0 = H.plane_wave_structure(np.zeros((1, 3)), K, lc, (1, 0))[0]
        got = (T.reshape(2 * rmax, 2 * rmax) @ p0.reshape(-1)).reshape(2, rmax)
        errs.append(np.abs(got - pcl).max() / np.abs(pcl).max())
    assert errs[0] > 10 * errs[1] > 100 * errs[2] and errs[2] < 1e-6, errs


@pytest.mark.parametrize("eps", [2.56, -15.8 + 1.07j])
def test_reciprocity(H, eps):
    lc = 4
    pos = np.vstack([POS, [[0.02, -0.2, 0.11]]])
    T = H.tmatrix_sphere_cluster(pos, 0.07, 3, lc, 1.0, eps, 1.0)
    modes = [(n, m) for n in range(1, lc + 1) for m in range(-n, n + 1)]
    idx = {nm: r for r, nm in enumerate(modes)}
    err = 0.0
    for a in range(2):
        for b in range(2):
            for r, (n, m) in enumerate(modes):
                for s, (v, u) in enumerate(modes):
                    err = max(err, abs(T[a, r, b, s] - (-1) ** (m + u + n + v) * T[b, idx[(v, -u)], a, idx[(n, -m)]]))
    assert err <= 1e-13 * np.abs(T).max()


def test_per_sphere_lmax(H):
    """A sphere with lmax 1 beside an lmax-3 sphere (its coefficients above
    n = 1 zero, common.py:334-341): the T-matrix differs from the all-lmax-3
    one and stays reciprocal."""
    T = H.tmatrix_sphere_cluster(POS, 0.08, [3, 1], 3, 1.0, 2.56, 1.0)
    T3 = H.tmatrix_sphere_cluster(POS, 0.08, 3, 3, 1.0, 2.56, 1.0)
    assert np.abs(T - T3).max() > 1e-6 * np.abs(T3).max()
    modes = [(n, m) for n in range(1, 4) for m in range(-n, n + 1)]
    idx = {nm: r for r, nm in enumerate(modes)}
    err = max(abs(T[a, r, b, s] - (-1) ** (m + u + n + v) * T[b, idx[(v, -u)], a, idx[(n, -m)]])
              for a in range(2) for b in range(2) for r, (n, m) in enumerate(modes) for s, (v, u) in enumerate(modes))
    assert err <= 1e-13 * np.abs(T).max()
