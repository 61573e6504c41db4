"""Generalized Mie theory for a cluster of particles with general T-matrices.

Same constructor and result attributes as the reference's
src/miepy/cluster.py:13-112 (particles, source, wavelength, lmax, medium,
origin, interactions, method -> tmatrix, p_src, p_inc, p_scat;
cross_sections() :393, per-particle variants :379-422, force() :472,
torque() :492, update() :588), with the solve of :629-657 on the device:

  host:   per-particle T-matrices (particle.compute_tmatrix), p_src
  device: particle_aggregate_tmatrix -> BiCGSTAB on (I + T), or the
          matrix-free operator (no stored matrix) ->
          p_scat = T p_inc (einsum "naibj,nbj->nai", :650,657) ->
          per-particle cross-sections, force, torque.

Particles: `sphere` (T-matrix = diag(a_n, b_n), src/miepy/tmatrix/common.py:6-28),
`sphere_cluster_particle` (the T-matrix of a rigid sphere cluster,
src/miepy/tmatrix/common.py:298-364) and `tmatrix_particle` (a T-matrix
computed elsewhere; the EBCM solvers for non-spherical shapes are not part
of this library).  Extra keyword
arguments (the reference hard-codes these): solver_mode ('stored' |
'matfree'), tol, maxiter.
"""

import numpy as np

from . import device as D
from .backend import force_all as _force_all, solver, torque_all as _torque_all
from .flux import cross_sections as _cs
from .materials import constant_material, material_struct
from .mie import mie_coefficients


class sphere:
    """A sphere (src/miepy/particles/sphere.py): T-matrix diag(a_n, b_n)."""

    def __init__(self, position, radius, material):
        self.position = np.asarray(position, dtype=float)
        self.radius = float(radius)
        self.material = material

    def compute_tmatrix(self, lmax, wavelength, eps_m):
        rmax = lmax * (lmax + 2)
        eps = np.atleast_1d(self.material.eps(wavelength)).astype(complex)
        k = 2 * np.pi * np.sqrt(np.real(eps_m)) / wavelength
        conducting = np.array([getattr(self.material, "name", None) == "metal"])
        # particle and medium mu = 1, as tmatrix_sphere (src/miepy/tmatrix/common.py:22-24)
        an_bn, _ = mie_coefficients(np.array([self.radius]), eps, np.ones(1), eps_m, 1.0, k, lmax, conducting)
        n_idx = np.concatenate([[n - 1] * (2 * n + 1) for n in range(1, lmax + 1)])
        T = np.zeros((2, rmax, 2, rmax), dtype=complex)
        r = np.arange(rmax)
        T[0, r, 0, r] = an_bn[0, 0, n_idx]
        T[1, r, 1, r] = an_bn[0, 1, n_idx]
        self.tmatrix = T
        return T


class tmatrix_particle:
    """A particle given by its T-matrix [2, rmax, 2, rmax] in the medium at
    the cluster's wavelength (computed by any T-matrix method)."""

    def __init__(self, position, tmatrix, material=None):
        self.position = np.asarray(position, dtype=float)
        self.tmatrix = np.asarray(tmatrix, dtype=complex)
        self.material = material if material is not None else constant_material(eps=1.0)

    def compute_tmatrix(self, lmax, wavelength, eps_m):
        rmax = lmax * (lmax + 2)
        if self.tmatrix.shape != (2, rmax, 2, rmax):
            raise ValueError(f"tmatrix must be [2, {rmax}, 2, {rmax}] for lmax={lmax}")
        return self.tmatrix


class sphere_cluster_particle:
    """A rigid collection of non-overlapping spheres acting as one particle
    (src/miepy/particles/sphere_cluster_particle.py:9-70): its T-matrix is
    tmatrix_sphere_cluster of the spheres at their given coordinates, about
    the coordinate origin as in the reference, reduced to the cluster's lmax;
    the particle sits at the spheres' centre of mass.  Fixed orientation only
    (the T-matrix rotation of the reference's particle base is not part of
    this library)."""

    def __init__(self, position, radius, material, lmax, orientation=None):
        self.p_position = np.asarray(position, dtype=float)
        self.Nparticles = len(self.p_position)
        if self.p_position.ndim != 2 or self.Nparticles == 1:
            raise ValueError("a sphere_cluster_particle must have at least 2 particles")
        if orientation is not None:
            raise NotImplementedError("sphere_cluster_particle: only the fixed orientation is supported")
        self.com = np.average(self.p_position, axis=0)
        self.position = self.com
        self.p_radii = np.empty(self.Nparticles, dtype=float)
        self.p_radii[...] = radius
        self.p_material = np.empty(self.Nparticles, dtype=object)
        self.p_material[...] = material
        self.p_lmax = np.empty(self.Nparticles, dtype=int)
        self.p_lmax[...] = lmax
        self.lmax_cluster = int(np.max(self.p_lmax))
        self.material = self.p_material[0]
        self.tmatrix = None

    def compute_tmatrix(self, lmax, wavelength, eps_m, **kwargs):
        from .tmatrix import tmatrix_reduce_lmax, tmatrix_sphere_cluster
        eps = np.array([complex(np.squeeze(m.eps(wavelength))) for m in self.p_material])
        T = tmatrix_sphere_cluster(self.p_position, self.p_radii, self.p_lmax, self.lmax_cluster, wavelength, eps,
                                   eps_m)
        self.tmatrix = tmatrix_reduce_lmax(T, lmax)
        return self.tmatrix

    def enclosed_radius(self):
        return np.max(np.linalg.norm(self.p_position - self.position[np.newaxis], axis=1)) + np.max(self.p_radii)


class cluster:
    def __init__(self, *, particles, source, wavelength, lmax, medium=None, origin=None, symmetry=None,
                 interface=None, interactions=True, method=None, solver_mode="stored", tol=None, maxiter=1000):
        if interface is not None:
            raise NotImplementedError("interface with miepy.cluster not yet supported")
        if symmetry is not None:
            raise NotImplementedError("symmetries with miepy.cluster not yet supported")
        if solver_mode not in ("stored", "matfree"):
            raise ValueError("solver_mode must be 'stored' or 'matfree'")
        self.source = source
        self.wavelength = wavelength
        self.lmax = lmax
        self.rmax = lmax * (lmax + 2)
        self.interactions = interactions
        self.method = method
        self.solver_mode = solver_mode
        self.tol = tol
        self.maxiter = maxiter
        self.report = None
        if medium is None:
            self.medium = constant_material(eps=1.0, mu=1.0)
        else:
            self.medium = medium
            if np.imag(self.medium.eps(wavelength)) != 0 or np.imag(self.medium.mu(wavelength)) != 0:
                raise ValueError("medium must be non-absorbing")
        self.particles = particles if isinstance(particles, list) else [particles]
        self.Nparticles = len(self.particles)
        self.position = np.array([p.position for p in self.particles], dtype=float).reshape(-1, 3)
        self.material = np.array([p.material for p in self.particles], dtype=object)
        eps_m = self.medium.eps(wavelength)
        self.tmatrix = np.array([p.compute_tmatrix(lmax, wavelength, eps_m) for p in self.particles],
                                dtype=complex).reshape(self.Nparticles, 2, self.rmax, 2, self.rmax)
        self.auto_origin = False
        if origin is None:
            self.origin = np.zeros(3)
        elif isinstance(origin, str) and origin == "auto":
            self.auto_origin = True
            self.origin = np.average(self.position, axis=0)
        else:
            self.origin = np.asarray(origin)
        self.material_data = material_struct(self.material, self.medium, wavelength=wavelength)
        self.p_cluster = None
        self.solve()

    def solve(self, wavelength=None, source=None):
        md = self.material_data
        k = float(np.real(md.k_b))
        N, rmax, lmax = self.Nparticles, self.rmax, self.lmax
        self.p_src = self.source.structure(self.position, k, lmax)  # [N, 2, rmax]
        D.require_gpu()
        p_src_d = D.to_device(self.p_src.reshape(1, -1), D.C128)
        tm_d = D.to_device(self.tmatrix, D.C128)
        n = 2 * rmax * N
        if self.interactions and N > 1:
            exact = self.method == solver.exact or getattr(self.method, "name", None) == "exact"
            # exact without a stored matrix (matrix-free mode): BiCGSTAB to 1e-14
            tol = 1e-14 if exact else (1e-5 if self.tol is None else self.tol)
            maxiter = max(self.maxiter, 10 * n) if exact else self.maxiter
            if exact and self.solver_mode == "stored":
                # the reference's exact branch: LU of I + T (interactions.cpp:356-357)
                A = D.assemble_tmatrix(self.position, tm_d, k, add_identity=True)
                x = D.lu_solve(A, p_src_d, overwrite=True)
                self.report = None
                del A
            elif self.solver_mode == "stored":
                A = D.assemble_tmatrix(self.position, tm_d, k, add_identity=True)
                x, self.report = D.bicgstab(A, p_src_d, tol=tol, maxiter=maxiter)
                del A
            else:
                op = D.ClusterOperator(self.position, None, k, tmatrix=tm_d)
                x, self.report = D.bicgstab(op, p_src_d, tol=tol, maxiter=maxiter, add_identity=True)
            p_inc_d = x.reshape(1, n)
        else:
            p_inc_d = p_src_d.clone()
        p_scat_d = D.tmatrix_apply(tm_d, p_inc_d)
        self.p_inc = p_inc_d.cpu().numpy().reshape(N, 2, rmax)
        self.p_scat = p_scat_d.cpu().numpy().reshape(N, 2, rmax)
        self.p_int = None

    def update(self, position=None, orientation=None):
        if orientation is not None:
            raise NotImplementedError("particle orientations are carried by the given T-matrices")
        if position is not None:
            self.position = np.asarray(np.atleast_2d(position), dtype=float)
            for p, r in zip(self.particles, self.position):
                p.position = r
            if self.auto_origin:
                self.origin = np.average(self.position, axis=0)
        self.p_cluster = None
        self.solve()

    # ---- observables (src/miepy/cluster.py:379-510) ---------------------------------

    def solve_cluster_coefficients(self, lmax=None):
        """p_cluster [2, rmax]: the cluster's scattered field expanded about
        self.origin up to lmax (src/miepy/cluster.py:616-627)."""
        lmax = self.lmax if lmax is None else lmax
        k = float(np.real(self.material_data.k_b))
        self.p_cluster = D.cluster_coefficients(self.position, self.p_scat[None], k, self.origin,
                                                lmax).cpu().numpy()[0]
        return self.p_cluster

    def cross_sections_per_multipole(self, lmax=None):
        """Cluster (scattering, absorption, extinction) per multipole, each
        [2, lmax] (src/miepy/cluster.py:379-391)."""
        lmax = self.lmax if lmax is None else lmax
        self.solve_cluster_coefficients(lmax)
        k = float(np.real(self.material_data.k_b))
        p0 = self.source.structure(np.array([self.origin]), k, lmax)[0]
        C = D.cluster_cross_sections(D.to_device(self.p_cluster[None], D.C128), p0[None], k).cpu().numpy()[0]
        return _cs(C[0], C[1], C[2])

    def cross_sections_per_multipole_of_particle(self, i):
        from .backend import particle_cross_sections
        return particle_cross_sections(self.p_scat[i], self.p_inc[i], self.p_src[i], self.material_data.k_b)

    def cross_sections_of_particle(self, i):
        C = self.cross_sections_per_multipole_of_particle(i)
        return _cs(*[np.sum(A) for A in C])

    def cross_sections(self):
        from .backend import _xs_rows
        C = _xs_rows(self.p_scat.reshape(self.Nparticles, -1), self.p_inc.reshape(self.Nparticles, -1),
                     self.p_src.reshape(self.Nparticles, -1), self.material_data.k_b, self.lmax)
        tot = C.sum(axis=(0, 2, 3))
        return _cs(tot[0], tot[1], tot[2])

    def _p_inc(self, source):
        return self.p_inc if source else self.p_inc - self.p_src

    def force(self, source=True):
        md = self.material_data
        return _force_all(self.p_scat, self._p_inc(source), md.k_b, md.eps_b, md.mu_b)

    def torque(self, source=True):
        md = self.material_data
        return _torque_all(self.p_scat, self._p_inc(source), md.k_b, md.eps_b, md.mu_b)

    def force_on_particle(self, i, source=True):
        return self.force(source)[i]

    def torque_on_particle(self, i, source=True):
        return self.torque(source)[i]

    def __repr__(self):
        return (f"{self.__class__.__name__}:\n    Nparticles = {self.Nparticles}\n    source = {self.source}\n"
                f"    wavelength = {self.wavelength:.2e}\n    medium = {self.medium}\n    lmax = {self.lmax}\n")
