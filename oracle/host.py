"""Restatement of the reference's Python host steps of the cluster solve.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Scalar, loop-per-particle code that follows the reference line by line:
- Mie coefficients: src/miepy/mie_single/mie_sphere.py:16-76 with the
  Riccati-Bessel helpers of src/miepy/special_functions.py:127-153;
- plane-wave source decomposition: src/miepy/sources/plane_waves.py:126-156,
  polarization normalization src/miepy/sources/sources.py:182,
  k-hat/TM/TE basis src/miepy/coordinates.py:58-68;
- sphere_cluster construction + solve: src/miepy/sphere_cluster.py:27-137,
  664-717 (p_scat = p_inc * mie_scat[:, :, n_idx], p_int uses the swapped
  polarization mie_int[:, ::-1, n_idx]);
- cross-sections (sphere_cluster.py:442-471) and force/torque (:521-559);
- cluster coefficients (src/miepy/vsh/cluster_coefficients.py:14-53);
- the sphere-cluster T-matrix with its source and re-expansion matrices
  (src/miepy/tmatrix/common.py:199-364).
"""

import numpy as np
from scipy import special

import oracle as O


# ---- Mie coefficients ---------------------------------------------------------

def riccati_1_single(n, x):
    """special_functions.py:127-132"""
    jn = special.spherical_jn(n, x)
    jnp = special.spherical_jn(n, x, derivative=True)
    return np.array([x * jn, jn + x * jnp])


def riccati_2_single(n, x):
    """special_functions.py:135-143"""
    yn = special.spherical_yn(n, x)
    ynp = special.spherical_yn(n, x, derivative=True)
    return np.array([x * yn, yn + x * ynp])


def riccati_3_single(n, x):
    """special_functions.py:146-153"""
    return riccati_1_single(n, x) + 1j * riccati_2_single(n, x)


def mie_sphere_scattering_coefficients(radius, n, eps, mu, eps_b, mu_b, k, conducting=False):
    """mie_sphere.py:16-45"""
    xvals = k * radius
    m = (eps / eps_b) ** 0.5
    jn = riccati_1_single(n, xvals)
    yn = riccati_3_single(n, xvals)
    if conducting:
        a = jn[1] / yn[1]
        b = jn[0] / yn[0]
    else:
        jnm = riccati_1_single(n, m * xvals)
        mt = m * mu_b / mu
        a = (mt * jnm[0] * jn[1] - jn[0] * jnm[1]) / (mt * jnm[0] * yn[1] - yn[0] * jnm[1])
        b = (jnm[0] * jn[1] - mt * jn[0] * jnm[1]) / (jnm[0] * yn[1] - mt * yn[0] * jnm[1])
    return a, b


def mie_sphere_interior_coefficients(radius, n, eps, mu, eps_b, mu_b, k, conducting=False):
    """mie_sphere.py:48-76"""
    xvals = k * radius
    m = (eps / eps_b) ** 0.5
    mt = m * mu_b / mu
    jn = riccati_1_single(n, xvals)
    yn = riccati_3_single(n, xvals)
    jnm = riccati_1_single(n, m * xvals)
    if conducting:
        return 0.0, 0.0
    c = (m * jn[0] * yn[1] - m * yn[0] * jn[1]) / (jnm[0] * yn[1] - mt * yn[0] * jnm[1])
    d = (m * jn[0] * yn[1] - m * yn[0] * jn[1]) / (mt * jnm[0] * yn[1] - yn[0] * jnm[1])
    return c, d


def mie_coefficients(radius, eps, mu, eps_b, mu_b, k, lmax, conducting=None):
    """sphere_cluster.py:99-125: per particle, per order loop."""
    N = len(radius)
    if conducting is None:
        conducting = [False] * N
    mie_scat = np.zeros([N, 2, lmax], dtype=complex)
    mie_int = np.zeros([N, 2, lmax], dtype=complex)
    for i in range(N):
        for n in range(1, lmax + 1):
            mie_scat[i, :, n - 1] = mie_sphere_scattering_coefficients(
                radius[i], n, eps[i], mu[i], eps_b, mu_b, k, conducting=conducting[i])
            mie_int[i, :, n - 1] = mie_sphere_interior_coefficients(
                radius[i], n, eps[i], mu[i], eps_b, mu_b, k, conducting=conducting[i])
    return mie_scat, mie_int


# ---- plane-wave source -------------------------------------------------------------

def sph_basis_vectors(theta, phi):
    """coordinates.py:58-68"""
    r_hat = np.array([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)])
    theta_hat = np.array([np.cos(theta) * np.cos(phi), np.cos(theta) * np.sin(phi), -1 * np.sin(theta)])
    phi_hat = np.array([-1 * np.sin(phi), np.cos(phi), np.zeros_like(phi)])
    return r_hat, theta_hat, phi_hat


def plane_wave_structure(position, k, lmax, polarization, theta=0.0, phi=0.0, amplitude=1.0, phase=0.0):
    """plane_waves.py:126-156 (polarization normalized as sources.py:181-182)."""
    pol = np.asarray(polarization, dtype=complex)
    pol = pol / np.linalg.norm(pol)
    k_hat, _, _ = sph_basis_vectors(theta, phi)
    position = np.asarray(position, dtype=float).reshape(-1, 3)
    N = len(position)
    rmax = O.lmax_to_rmax(lmax)
    pi_vals = np.empty(rmax)
    tau_vals = np.empty(rmax)
    Emn_vals = np.empty(rmax)
    m_vals = np.empty(rmax, dtype=int)
    r = 0
    for n in range(1, lmax + 1):
        for m in range(-n, n + 1):
            pi_vals[r] = O.pi_func(n, m, theta)
            tau_vals[r] = O.tau_func(n, m, theta)
            Emn_vals[r] = np.abs(O.Emn(m, n))
            m_vals[r] = m
            r += 1
    phases = k * (position @ k_hat) + phase
    exp_factors = amplitude * np.exp(1j * (phases[:, None] - m_vals * phi)) * Emn_vals
    p_src = np.empty([N, 2, rmax], dtype=complex)
    p_src[:, 0, :] = exp_factors * (tau_vals * pol[0] - 1j * pi_vals * pol[1])
    p_src[:, 1, :] = exp_factors * (pi_vals * pol[0] - 1j * tau_vals * pol[1])
    return p_src


def n_index(lmax):
    """sphere_cluster.py:11-18"""
    out = []
    for n in range(1, lmax + 1):
        out += [n - 1] * (2 * n + 1)
    return np.array(out, dtype=int)


# ---- full cluster solve ------------------------------------------------------------

def solve_cluster(position, radius, eps, wavelength, lmax, polarization=(1, 0), theta=0.0, phi=0.0,
                  eps_b=1.0, mu=None, mu_b=1.0, method="bicgstab", tol=1e-5, maxiter=1000,
                  interactions=True, conducting=None, amplitude=1.0, phase=0.0):
    """sphere_cluster(...) construction: Mie, p_src, T, solve, p_scat/p_int,
    cross-sections (summed per particle, sphere_cluster.py:442-471), forces and
    torques (force_all / torque_all with source=True).  Returns a dict."""
    position = np.asarray(np.atleast_2d(position), dtype=float)
    N = position.shape[0]
    radius = np.broadcast_to(np.asarray(radius, dtype=float), (N,)).copy()
    eps = np.broadcast_to(np.asarray(eps, dtype=complex), (N,)).copy()
    mu = np.ones(N, dtype=complex) if mu is None else np.broadcast_to(np.asarray(mu, dtype=complex), (N,)).copy()
    n_b = (eps_b * mu_b) ** 0.5
    k = 2 * np.pi * n_b / wavelength
    k = float(np.real(k))
    rmax = O.lmax_to_rmax(lmax)
    mie_scat, mie_int = mie_coefficients(radius, eps, mu, eps_b, mu_b, k, lmax, conducting)
    p_src = plane_wave_structure(position, k, lmax, polarization, theta, phi, amplitude, phase)
    T = None
    iters, err = 0, 0.0
    if interactions:
        T = O.sphere_aggregate_tmatrix(position, mie_scat, k)
        x, iters, err = O.solve_linear_system(T, p_src, method=method, tol=tol, maxiter=maxiter)
        p_inc = x.reshape(N, 2, rmax)
    else:
        p_inc = p_src.copy()
    nidx = n_index(lmax)
    p_scat = p_inc * mie_scat[:, :, nidx]
    p_int = p_inc * mie_int[:, ::-1, nidx]
    Cs = np.zeros(3)
    per_particle = np.zeros((N, 3))
    for i in range(N):
        C = O.particle_cross_sections(p_scat[i], p_inc[i], p_src[i], k)
        per_particle[i] = [np.sum(c) for c in C]
        Cs += per_particle[i]
    F = O.force_all(p_scat, p_inc, k, np.real(eps_b), np.real(mu_b))
    Tq = O.torque_all(p_scat, p_inc, k, np.real(eps_b), np.real(mu_b))
    return dict(k=k, mie_scat=mie_scat, mie_int=mie_int, p_src=p_src, p_inc=p_inc, p_scat=p_scat, p_int=p_int,
                T=T, iters=iters, err=err, cross_sections=Cs, per_particle=per_particle, force=F, torque=Tq)


# ---- cluster coefficients ------------------------------------------------------------

def cluster_coefficients(positions, p_scat, k, origin, lmax=None):
    """src/miepy/vsh/cluster_coefficients.py:14-53: every particle's outgoing
    expansion re-expanded about `origin` with the incident-mode translation
    coefficients of rij = origin - r_i (the scalar vsh_translation,
    cpp/src/vsh_translation.cpp:428-470, through the oracle); a particle at the
    origin adds its coefficients directly.  Returns p_cluster [2, rmax]."""
    positions = np.asarray(positions, dtype=float).reshape(-1, 3)
    p_scat = np.asarray(p_scat, dtype=complex)
    origin = np.asarray(origin, dtype=float)
    rmax_in = p_scat.shape[-1]
    lmax_in = O.rmax_to_lmax(rmax_in)
    lmax = lmax_in if lmax is None else lmax
    rmax = O.lmax_to_rmax(lmax)
    modes = lambda L: [(n, m) for n in range(1, L + 1) for m in range(-n, n + 1)]
    p_cluster = np.zeros((2, rmax), dtype=complex)
    for i in range(positions.shape[0]):
        if np.all(positions[i] == origin):
            p_cluster[0, :rmax_in] += p_scat[i, 0, :rmax_in]
            p_cluster[1, :rmax_in] += p_scat[i, 1, :rmax_in]
            continue
        d = origin - positions[i]
        rad = np.linalg.norm(d)
        theta = np.arccos(d[2] / rad)
        phi = np.arctan2(d[1], d[0])
        for r, (n, m) in enumerate(modes(lmax)):
            for rp, (v, u) in enumerate(modes(lmax_in)):
                a, b = p_scat[i, 0, rp], p_scat[i, 1, rp]
                A, B = O.vsh_translation(m, n, u, v, rad, theta, phi, k, mode="incident")
                p_cluster[0, r] += a * A + b * B
                p_cluster[1, r] += a * B + b * A
    return p_cluster


# ---- T-matrix of a sphere cluster ------------------------------------------------------

def _modes(lmax):
    return [(n, m) for n in range(1, lmax + 1) for m in range(-n, n + 1)]


def _sph(d):
    rad = np.linalg.norm(d)
    return rad, np.arccos(d[2] / rad), np.arctan2(d[1], d[0])


def cluster_source_matrix(positions, origin, lmax_p, lmax_out, k):
    """src/miepy/tmatrix/common.py:199-247: S [N 2 rmax_p, 2 rmax_out], origin
    mode (b, n, m) expanded about particle i with the incident-mode
    vsh_translation(u, v, m, n) of d = r_i - origin; a particle within 1e-15
    of the origin takes the identity on the common modes."""
    positions = np.asarray(positions, dtype=float).reshape(-1, 3)
    rmax_p, rmax_out = O.lmax_to_rmax(lmax_p), O.lmax_to_rmax(lmax_out)
    S = np.zeros((positions.shape[0] * 2 * rmax_p, 2 * rmax_out), dtype=complex)
    for i in range(positions.shape[0]):
        d = positions[i] - origin
        row = i * 2 * rmax_p
        if np.linalg.norm(d) < 1e-15:
            rmin = O.lmax_to_rmax(min(lmax_p, lmax_out))
            for a in range(2):
                for r in range(rmin):
                    S[row + a * rmax_p + r, a * rmax_out + r] = 1.0
            continue
        rad, theta, phi = _sph(d)
        for s, (v, u) in enumerate(_modes(lmax_p)):
            for r, (n, m) in enumerate(_modes(lmax_out)):
                A, B = O.vsh_translation(u, v, m, n, rad, theta, phi, k, mode="incident")
                S[row + s, r] += A
                S[row + s, rmax_out + r] += B
                S[row + rmax_p + s, r] += B
                S[row + rmax_p + s, rmax_out + r] += A
    return S


def cluster_reexpansion_matrix(positions, origin, lmax_p, lmax_out, k):
    """src/miepy/tmatrix/common.py:250-295: R [2 rmax_out, N 2 rmax_p], particle
    i's scattered mode (b, v, u) re-expanded about the origin with the
    incident-mode vsh_translation(m, n, u, v) of d = origin - r_i."""
    positions = np.asarray(positions, dtype=float).reshape(-1, 3)
    rmax_p, rmax_out = O.lmax_to_rmax(lmax_p), O.lmax_to_rmax(lmax_out)
    R = np.zeros((2 * rmax_out, positions.shape[0] * 2 * rmax_p), dtype=complex)
    for i in range(positions.shape[0]):
        d = origin - positions[i]
        col = i * 2 * rmax_p
        if np.linalg.norm(d) < 1e-15:
            rmin = O.lmax_to_rmax(min(lmax_p, lmax_out))
            for a in range(2):
                for r in range(rmin):
                    R[a * rmax_out + r, col + a * rmax_p + r] = 1.0
            continue
        rad, theta, phi = _sph(d)
        for r, (n, m) in enumerate(_modes(lmax_out)):
            for s, (v, u) in enumerate(_modes(lmax_p)):
                A, B = O.vsh_translation(m, n, u, v, rad, theta, phi, k, mode="incident")
                R[r, col + s] += A
                R[r, col + rmax_p + s] += B
                R[rmax_out + r, col + s] += B
                R[rmax_out + r, col + rmax_p + s] += A
    return R


def tmatrix_sphere_cluster(pos, radii, lmax, lmax_cluster, wavelength, eps, eps_m):
    """src/miepy/tmatrix/common.py:298-364: T = R diag(mie) solve(I + T_agg, S),
    [2, rmax_out, 2, rmax_out]; per-sphere lmax (Mie coefficients above it
    zero), the origin at (0, 0, 0)."""
    pos = np.asarray(pos, dtype=float).reshape(-1, 3)
    N = pos.shape[0]
    radii = np.broadcast_to(np.atleast_1d(np.asarray(radii, dtype=float)), (N,))
    lmax_arr = np.broadcast_to(np.atleast_1d(np.asarray(lmax, dtype=int)), (N,))
    eps_arr = np.broadcast_to(np.atleast_1d(np.asarray(eps, dtype=complex)), (N,))
    lmax_p = int(np.max(lmax_arr))
    k = 2 * np.pi * eps_m ** 0.5 / wavelength
    origin = np.zeros(3)
    rmax_p, rmax_out = O.lmax_to_rmax(lmax_p), O.lmax_to_rmax(lmax_cluster)
    mie = np.zeros((N, 2, lmax_p), dtype=complex)
    for i in range(N):
        for n in range(1, int(lmax_arr[i]) + 1):
            mie[i, :, n - 1] = mie_sphere_scattering_coefficients(radii[i], n, eps_arr[i], 1, eps_m, 1, k)
    size = N * 2 * rmax_p
    A = np.eye(size, dtype=complex) + O.sphere_aggregate_tmatrix(pos, mie, k).reshape(size, size)
    S = cluster_source_matrix(pos, origin, lmax_p, lmax_cluster, k)
    X = np.linalg.solve(A, S)
    n_idx = n_index(lmax_p)
    mie_diag = np.concatenate([np.concatenate([mie[i, 0, n_idx], mie[i, 1, n_idx]]) for i in range(N)])
    Y = mie_diag[:, None] * X
    R = cluster_reexpansion_matrix(pos, origin, lmax_p, lmax_cluster, k)
    return (R @ Y).reshape(2, rmax_out, 2, rmax_out)
