"""Materials: permittivity/permeability as functions of wavelength.

Mirrors src/miepy/material_functions/create.py (constant_material :58-82,
dielectric :44-56, function_material :85-108, data_material :111-135),
src/miepy/materials/predefined.py (vacuum/air/water/metal) and
src/miepy/material_functions/material_struct.py:4-49 (per-particle data and
the medium wavenumber k_b = 2 pi n_b / wavelength).
"""

import numpy as np


class material:
    def __init__(self, name=None):
        self.name = name

    def eps(self, wavelength):
        raise NotImplementedError

    def mu(self, wavelength):
        raise NotImplementedError

    def index(self, wavelength):
        return np.sqrt(self.eps(wavelength) * self.mu(wavelength))

    def __repr__(self):
        return str(self.name)


class constant_material(material):
    def __init__(self, eps=1.0, mu=1.0, index=None, name=None):
        super().__init__(name)
        self.eps_value = eps
        self.mu_value = mu
        if index is not None:
            self.eps_value = index ** 2

    def eps(self, wavelength):
        return np.full(np.shape(wavelength), self.eps_value, dtype=np.result_type(self.eps_value, float))[()] \
            if np.ndim(wavelength) else np.asarray(self.eps_value)[()]

    def mu(self, wavelength):
        return np.full(np.shape(wavelength), self.mu_value, dtype=np.result_type(self.mu_value, float))[()] \
            if np.ndim(wavelength) else np.asarray(self.mu_value)[()]

    def __repr__(self):
        if self.name is None:
            return f"constant_material(eps={self.eps_value}, mu={self.mu_value})"
        return str(self.name)


def dielectric(index=None, eps=1.0, mu=1.0, name=None):
    if index is not None:
        return constant_material(index=index, name=name)
    return constant_material(eps=eps, mu=mu, name=name)


class function_material(material):
    def __init__(self, eps_function, mu_function=None, name=None):
        super().__init__(name)
        self.eps_function = eps_function
        self.mu_function = mu_function if mu_function is not None else (lambda wavelength: 1.0)

    def eps(self, wavelength):
        return np.vectorize(self.eps_function)(wavelength)[()]

    def mu(self, wavelength):
        return np.vectorize(self.mu_function)(wavelength)[()]


class data_material(material):
    """Linear interpolation of tabulated eps (and mu) data."""

    def __init__(self, wavelength, eps, mu=None, name=None):
        super().__init__(name)
        if np.isscalar(wavelength):
            raise ValueError("Material has only 1 data point. Use constant_material instead")
        self.wavelength = np.asarray(wavelength, dtype=float)
        self.eps_data = np.asarray(eps, dtype=complex)
        self.mu_data = np.ones_like(self.eps_data) if mu is None else np.asarray(mu, dtype=complex)

    def _interp(self, data, wavelength):
        w = np.asarray(wavelength, dtype=float)
        if np.any(w < self.wavelength[0]) or np.any(w > self.wavelength[-1]):
            raise ValueError("wavelength outside the tabulated range")
        return (np.interp(w, self.wavelength, data.real) + 1j * np.interp(w, self.wavelength, data.imag))[()]

    def eps(self, wavelength):
        return self._interp(self.eps_data, wavelength)

    def mu(self, wavelength):
        return self._interp(self.mu_data, wavelength)


class materials:
    @staticmethod
    def vacuum():
        return constant_material(index=1, name="vacuum")

    @staticmethod
    def air():
        return constant_material(index=1, name="air")

    @staticmethod
    def water():
        return constant_material(index=1.33, name="water")

    @staticmethod
    def metal():
        # perfect conductor: the Mie coefficients take their conducting limit
        return constant_material(eps=complex(1, 1e40), name="metal")


class material_struct:
    """Per-particle eps/mu/n and the medium's eps_b, mu_b, n_b, k_b at one wavelength."""

    def __init__(self, materials_, medium, wavelength=None):
        self.materials = list(materials_)
        self.medium = medium
        N = len(self.materials)
        self.eps = np.zeros(N, dtype=complex)
        self.mu = np.zeros(N, dtype=complex)
        self.n = np.zeros(N, dtype=complex)
        self.eps_b = self.mu_b = self.n_b = self.k_b = None
        self._wavelength = None
        self.wavelength = wavelength

    @property
    def wavelength(self):
        return self._wavelength

    @wavelength.setter
    def wavelength(self, value):
        self._wavelength = value
        if value is None:
            return
        self.eps_b = self.medium.eps(value)
        self.mu_b = self.medium.mu(value)
        self.n_b = (self.eps_b * self.mu_b) ** 0.5
        self.k_b = 2 * np.pi * self.n_b / value
        for i, m in enumerate(self.materials):
            self.eps[i] = m.eps(value)
            self.mu[i] = m.mu(value)
        self.n[...] = (self.eps * self.mu) ** 0.5
