"""Plane-wave sources and their VSH decomposition p_src[N, 2, rmax].

Mirrors src/miepy/sources/plane_waves.py (from_string :35-90, structure
:126-156), the polarization normalization of src/miepy/sources/sources.py:182
and the k-hat/TM/TE basis of src/miepy/coordinates.py:58-68.  The angular
functions pi_n^m, tau_n^m (cpp/src/special.cpp:209-270, including the
theta = 0 / pi limits) and |E_mn| (cpp/src/vsh_functions.cpp:56-58) are
tabulated here per wave (rmax values); the per-particle coefficients are
formed on the device (k_plane_wave).
"""

import math

import numpy as np


def _legendre_table(lmax, x):
    """P_n^m(x), m >= 0, no Condon-Shortley phase: dict[(n, m)]."""
    s = math.sqrt(max(0.0, (1 - x) * (1 + x)))
    P = {}
    for m in range(lmax + 1):
        pmm = 1.0
        for i in range(1, m + 1):
            pmm *= (2 * i - 1) * s
        P[(m, m)] = pmm
        if m + 1 <= lmax:
            P[(m + 1, m)] = x * (2 * m + 1) * pmm
        for l in range(m + 2, lmax + 1):
            P[(l, m)] = ((2 * l - 1) * x * P[(l - 1, m)] - (l + m - 1) * P[(l - 2, m)]) / (l - m)
    return P


def pi_tau(lmax, theta):
    """pi_n^m(theta), tau_n^m(theta) for all modes, in mode order: [rmax] each."""
    pis, taus = [], []
    at0, atpi = theta == 0, theta == math.pi
    if not (at0 or atpi):
        x = math.cos(theta)
        st = math.sin(theta)
        P = _legendre_table(lmax, x)
    for n in range(1, lmax + 1):
        for m in range(-n, n + 1):
            if at0 or atpi:
                sgn_pi = 1 if at0 else (-1) ** (n + 1)
                sgn_tau = 1 if at0 else (-1) ** n
                if m == 1:
                    pi, tau = sgn_pi * n * (n + 1) / 2.0, sgn_tau * n * (n + 1) / 2.0
                elif m == -1:
                    pi, tau = sgn_pi * 0.5, -sgn_tau * 0.5
                else:
                    pi, tau = 0.0, 0.0
            else:
                am = abs(m)
                leg = P[(n, am)]
                prev = P[(n - 1, am)] if n - 1 >= am else 0.0
                # (x^2-1) dP/dx = n x P_n - (n+m) P_{n-1};  tau = -sin(theta) dP/dx
                dleg = -st * (n * x * leg - (n + am) * prev) / (x * x - 1)
                if m < 0:
                    f = (-1) ** am * math.factorial(n - am) / math.factorial(n + am)
                    leg, dleg = leg * f, dleg * f
                pi, tau = m / st * leg, dleg
            pis.append(pi)
            taus.append(tau)
    return np.array(pis), np.array(taus)


def emn_abs(lmax):
    out = []
    for n in range(1, lmax + 1):
        for m in range(-n, n + 1):
            out.append(math.sqrt((2 * n + 1) * math.factorial(n - m) / (n * (n + 1) * math.factorial(n + m))))
    return np.array(out)


def sph_basis_vectors(theta, phi):
    r_hat = np.array([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)])
    theta_hat = np.array([np.cos(theta) * np.cos(phi), np.cos(theta) * np.sin(phi), -1 * np.sin(theta)])
    phi_hat = np.array([-1 * np.sin(phi), np.cos(phi), np.zeros_like(phi)])
    return r_hat, theta_hat, phi_hat


class plane_wave:
    def __init__(self, polarization, amplitude=1, phase=0, theta=0, phi=0, standing=False):
        self.amplitude = amplitude
        self.phase = phase
        self.theta = theta
        self.phi = phi
        self.standing = standing
        self.k_hat, self.n_tm, self.n_te = sph_basis_vectors(theta, phi)
        self.polarization = np.asarray(polarization, dtype=complex)
        self.polarization = self.polarization / np.linalg.norm(self.polarization)

    def __repr__(self):
        return f"plane_wave(polarization={self.polarization}, amplitude={self.amplitude}, theta={self.theta}, phi={self.phi})"

    @classmethod
    def from_string(cls, polarization, direction="z", amplitude=1, phase=0, standing=False):
        if direction in ["z", "+z"]:
            theta, phi = 0, 0
        elif direction == "-z":
            theta, phi = np.pi, 0
        elif direction in ["x", "+x"]:
            theta, phi = np.pi / 2, 0
        elif direction == "-x":
            theta, phi = np.pi / 2, np.pi
        elif direction in ["y", "+y"]:
            theta, phi = np.pi / 2, np.pi / 2
        elif direction == "-y":
            theta, phi = np.pi / 2, 3 * np.pi / 2
        else:
            raise ValueError(f"'{direction}' is not a valid direction of propagation. "
                             "Use one of ['x', 'y', 'z', '-x', '-y', '-z']")
        if polarization == direction[-1]:
            raise ValueError("polarization cannot be the same as the direction of propagation")
        if polarization == "x":
            pol = [1, 0] if direction[-1] == "z" else [0, -1]
        elif polarization == "y":
            pol = [0, 1]
        elif polarization == "z":
            pol = [-1, 0]
        elif polarization == "rhc":
            pol = [1, 1j]
        elif polarization == "lhc":
            pol = [1, -1j]
        else:
            raise ValueError(f"'{polarization}' is not a valid polarization. Use one of ['x', 'y', 'z', 'rhc', 'lhc']")
        return cls(polarization=pol, theta=theta, phi=phi, amplitude=amplitude, phase=phase, standing=standing)

    def waves(self):
        return [self]

    def mode_table(self, lmax):
        """T[2, rmax]: amplitude |E_mn| e^{-i m phi} (tau pol0 - i pi pol1, pi pol0 - i tau pol1),
        so that p_src[i] = e^{i (k r_i . khat + phase)} T (plane_waves.py:126-156)."""
        pi_vals, tau_vals = pi_tau(lmax, float(self.theta))
        E = emn_abs(lmax)
        m_vals = np.concatenate([np.arange(-n, n + 1) for n in range(1, lmax + 1)])
        ef = self.amplitude * np.exp(-1j * m_vals * self.phi) * E
        T = np.empty((2, pi_vals.size), dtype=complex)
        T[0] = ef * (tau_vals * self.polarization[0] - 1j * pi_vals * self.polarization[1])
        T[1] = ef * (pi_vals * self.polarization[0] - 1j * tau_vals * self.polarization[1])
        return T

    def structure(self, position, k, lmax):
        """p_src[N, 2, rmax] of this wave at the particle positions (computed on the device)."""
        from . import device as D
        return D.plane_wave_structure([self], position, k, lmax)[0].cpu().numpy()

    def __add__(self, other):
        return combined_source(self, other)


class combined_source:
    """Sum of sources (src/miepy/sources/sources.py:187-239)."""

    def __init__(self, *sources):
        self.sources = []
        for s in sources:
            self.sources += s.sources if isinstance(s, combined_source) else [s]

    def waves(self):
        return list(self.sources)

    def structure(self, position, k, lmax):
        from . import device as D
        return D.plane_wave_structure([self], position, k, lmax)[0].cpu().numpy()

    def __add__(self, other):
        return combined_source(self, other)
