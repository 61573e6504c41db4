"""Independent numpy/scipy evaluator of VSH fields and of the Maxwell stress
tensor, used by the tests to pin the oracle without retyping its formulas.

Field conventions follow the reference's expansion code
(src/miepy/vsh/expansion.py:13-42: E = factor * sum_mn E_mn (p_e N + p_m M),
factor = i for outgoing and -i for incident waves; src/miepy/vsh/vsh_functions.py:40-78
for N and M; the plane wave of src/miepy/sources/plane_waves.py:93-124).  The
special functions come from scipy (lpmv, spherical_jn/yn) and the angular
derivative tau = dP/dtheta from finite differences, so nothing here shares
code or recurrences with the oracle.

H' below is curl(E) / (i k): the physical H times the medium impedance, the
normalization of the reference's plane-wave H_field (H' = k_hat x E).  The
time-averaged Maxwell stress tensor in a medium of permittivity eps is then
<T> = eps/2 Re[E E* + H' H'* - 1/2 I (|E|^2 + |H'|^2)].
"""

from math import factorial

import numpy as np
from scipy import special


def modes(lmax):
    return [(n, m) for n in range(1, lmax + 1) for m in range(-n, n + 1)]


def P(n, m, x):
    """Associated Legendre without Condon-Shortley phase (GSL NONE)."""
    if abs(m) > n:
        return 0.0 * x
    val = special.lpmv(abs(m), n, x) * (-1) ** abs(m)
    if m < 0:
        val = val * (-1) ** (-m) * factorial(n + m) / factorial(n - m)
    return val


def dP_dtheta(n, m, theta, h=1e-3):
    f = lambda t: P(n, m, np.cos(t))
    d1 = (f(theta + h) - f(theta - h)) / (2 * h)
    d2 = (f(theta + 2 * h) - f(theta - 2 * h)) / (4 * h)
    return (4 * d1 - d2) / 3  # Richardson: O(h^4)


def zn(n, x, mode, derivative=False):
    j = special.spherical_jn(n, x, derivative)
    if mode == "incident":
        return j
    return j + 1j * special.spherical_yn(n, x, derivative)


def _sph_frame(d):
    r = np.linalg.norm(d, axis=1)
    theta = np.arccos(np.clip(d[:, 2] / r, -1, 1))
    phi = np.arctan2(d[:, 1], d[:, 0])
    rh = np.array([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)])
    th = np.array([np.cos(theta) * np.cos(phi), np.cos(theta) * np.sin(phi), -np.sin(theta)])
    ph = np.array([-np.sin(phi), np.cos(phi), 0 * phi])
    return r, theta, phi, rh, th, ph


def vsh_cart(n, m, mode, pts, center, k):
    """N_mn, M_mn (vsh_functions.py:55-76) in Cartesian components, [3, npts]."""
    r, theta, phi, rh, th, ph = _sph_frame(pts - center)
    kr = k * r
    H, Hp = zn(n, kr, mode), zn(n, kr, mode, True)
    e = np.exp(1j * m * phi)
    Pv = P(n, m, np.cos(theta))
    pi = m / np.sin(theta) * Pv
    tau = dP_dtheta(n, m, theta)
    fac = (H + kr * Hp) * e / kr
    N = np.array([n * (n + 1) * Pv * H / kr * e, tau * fac, 1j * pi * fac])
    M = np.array([0 * kr, 1j * H * e * pi, -H * e * tau])
    to_cart = lambda S: S[0] * rh + S[1] * th + S[2] * ph
    return to_cart(N), to_cart(M)


def Emn(n, m):
    return 1j ** n * np.sqrt((2 * n + 1) * factorial(n - m) / (n * (n + 1) * factorial(n + m)))


def expand(p, lmax, mode, pts, center, k):
    """E = factor * sum E_mn (p_e N + p_m M) (expansion.py:13-42), [3, npts]."""
    factor = 1j if mode == "outgoing" else -1j
    E = 0
    for r, (n, m) in enumerate(modes(lmax)):
        N, M = vsh_cart(n, m, mode, pts, center, k)
        E = E + factor * Emn(n, m) * (p[0, r] * N + p[1, r] * M)
    return E


def expand_EH(p, lmax, mode, pts, center, k):
    """(E, H') of one expansion.  curl N = k M and curl M = k N for the
    reference's N and M (checked against finite differences by the tests), so
    H' = curl E / (i k) = -i factor * sum E_mn (p_e M + p_m N)."""
    factor = 1j if mode == "outgoing" else -1j
    E = 0
    H = 0
    for r, (n, m) in enumerate(modes(lmax)):
        N, M = vsh_cart(n, m, mode, pts, center, k)
        c = factor * Emn(n, m)
        E = E + c * (p[0, r] * N + p[1, r] * M)
        H = H - 1j * c * (p[0, r] * M + p[1, r] * N)
    return E, H


def plane_wave_E(pts, k, polarization, theta, phi, amplitude=1.0, phase=0.0):
    """plane_waves.py:93-106 with the polarization normalized (sources.py:181-182)."""
    pol = np.asarray(polarization, dtype=complex)
    pol = pol / np.linalg.norm(pol)
    k_hat = np.array([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)])
    n_tm = np.array([np.cos(theta) * np.cos(phi), np.cos(theta) * np.sin(phi), -np.sin(theta)])
    n_te = np.array([-np.sin(phi), np.cos(phi), 0.0])
    amp = amplitude * np.exp(1j * k * (pts @ k_hat)) * np.exp(1j * phase)
    E = (n_tm * pol[0] + n_te * pol[1])[:, None] * amp[None, :]
    return E, k_hat


def curl_fd(field, pts, h):
    """Central-difference curl (4th order) of field(pts) -> [3, npts]."""
    def d(axis):
        e = np.zeros(3)
        e[axis] = h
        f1 = (field(pts + e) - field(pts - e)) / (2 * h)
        f2 = (field(pts + 2 * e) - field(pts - 2 * e)) / (4 * h)
        return (4 * f1 - f2) / 3
    dx, dy, dz = d(0), d(1), d(2)
    return np.array([dy[2] - dz[1], dz[0] - dx[2], dx[1] - dy[0]])


def sphere_quadrature(center, R, ntheta, nphi):
    """Gauss-Legendre in cos(theta) x trapezoid in phi: exact for spherical
    harmonics of degree < 2 ntheta and |m| < nphi."""
    x, w = np.polynomial.legendre.leggauss(ntheta)
    phi = 2 * np.pi * np.arange(nphi) / nphi
    ct = np.repeat(x, nphi)
    st = np.sqrt(1 - ct ** 2)
    ph = np.tile(phi, ntheta)
    nrm = np.stack([st * np.cos(ph), st * np.sin(ph), ct], axis=1)
    wt = np.repeat(w, nphi) * (2 * np.pi / nphi) * R ** 2
    return center + R * nrm, nrm, wt


def maxwell_stress_force_torque(E, H, pts, nrm, wt, center, eps):
    """Time-averaged force and torque about ``center`` on what a closed surface
    holds: F = closed-surface integral of <T>.n, torque = integral of r x (<T>.n),
    <T> = eps/2 Re[E E* + H H* - 1/2 I (|E|^2 + |H|^2)] (H = H', impedance-normalized)."""
    n = nrm.T
    En = np.sum(E * n, axis=0)
    Hn = np.sum(H * n, axis=0)
    u = np.sum(np.abs(E) ** 2, axis=0) + np.sum(np.abs(H) ** 2, axis=0)
    Tn = 0.5 * eps * np.real(E * np.conj(En) + H * np.conj(Hn) - 0.5 * n * u)
    F = Tn @ wt
    tq = wt @ np.cross(pts - center, Tn.T)
    return F, tq
