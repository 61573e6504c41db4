"""Generalized Mie theory for a sphere cluster, solved on one MI355X.

Same constructor and result attributes as the reference's
src/miepy/sphere_cluster.py:27-137 (position, radius, material, source,
wavelength, lmax, medium, origin, interactions, method -> mie_scat, mie_int,
p_src, p_inc, p_scat, p_int; cross_sections() :442, force() :521,
torque() :541, per-particle variants :456-519, update_position() :637), with
the solve of :664-717 running as HIP kernels on the device:

  host:   material data, Mie coefficients (vectorized scipy), p_src
  device: assemble (I + T) -> BiCGSTAB (FP64) or complex64 BiCGSTAB + FP64
          refinement ('mixed'); or, with no stored matrix, BiCGSTAB on the
          FP64 matrix-free operator ('matfree') or complex64 corrections on
          the FP16-split matrix-free operator + FP64 matrix-free refinement
          ('mixed_matfree', lmax <= 5) -> p_scat / p_int, cross-sections,
          force, torque in one fused kernel.

Additional keyword arguments (the reference hard-codes these,
cpp/src/interactions.cpp:362-363): precision ('fp64' | 'mixed' | 'matfree'
| 'mixed_matfree'), tol, maxiter.  `source` may also be a list of sources, solved together as
several right-hand sides sharing one matrix; results then carry a leading
source axis.
"""

import numpy as np

from . import device as D
from .backend import solver
from .flux import cross_sections as _cs
from .materials import constant_material, material_struct


def _per_particle(value, count, dtype=None):
    """One entry per particle: a scalar (or one object) is broadcast to
    `count` entries, a sequence of `count` entries is kept; anything else is a
    shape error.  Same accepted inputs as the reference's `atleast(value, dim=1,
    length=count)` (src/miepy/utils.py:6-21) for 1-d per-particle data."""
    arr = np.asarray(value, dtype=dtype)
    if arr.ndim == 0:
        out = np.empty(count, dtype=arr.dtype)
        out[...] = arr.item() if dtype is object else arr
        return out
    if arr.ndim != 1:
        raise ValueError(f"per-particle data must be a scalar or 1-d, got shape {arr.shape}")
    if arr.shape[0] == count:
        return arr
    if arr.shape[0] == 1:
        return np.repeat(arr, count)
    raise ValueError("The shapes of position, radius, and material do not match")


def n_index(lmax):
    """r -> n-1 (sphere_cluster.py:11-18)."""
    return np.concatenate([[n - 1] * (2 * n + 1) for n in range(1, lmax + 1)]).astype(int)


class sphere_cluster:
    def __init__(self, *, position, radius, material, source, wavelength, lmax, medium=None, origin=None,
                 symmetry=None, interface=None, interactions=True, method=None, precision="fp64", tol=None,
                 maxiter=1000, refine_tol=None):
        if symmetry is not None or interface is not None:
            raise NotImplementedError("periodic symmetry and interfaces are not part of the mi355x cluster solve")
        self.position = np.asarray(np.atleast_2d(position), dtype=float)
        self.radius = _per_particle(radius, self.position.shape[0], dtype=float)
        self.material = _per_particle(material, self.position.shape[0], dtype=object)
        self.Nparticles = self.radius.shape[0]
        self.symmetry = symmetry
        self.interface = interface
        self.source = source
        self.wavelength = wavelength
        self.lmax = lmax
        self.rmax = lmax * (lmax + 2)
        self.interactions = interactions
        self.method = method
        self.precision = precision
        self.tol = tol
        self.maxiter = maxiter
        self.refine_tol = refine_tol
        if precision not in ("fp64", "mixed", "matfree", "mixed_matfree"):
            raise ValueError("precision must be 'fp64', 'mixed', 'matfree' or 'mixed_matfree'")
        if precision == "mixed_matfree" and lmax > 5:
            raise ValueError("precision 'mixed_matfree' covers lmax <= 5")

        self.auto_origin = False
        if origin is None:
            self.origin = np.zeros(3)
        elif isinstance(origin, str) and origin == "auto":
            self.auto_origin = True
            self.origin = np.average(self.position, axis=0)
        else:
            self.origin = np.asarray(origin)

        if medium is None:
            self.medium = constant_material(eps=1.0, mu=1.0)
        else:
            self.medium = medium
            if (np.imag(self.medium.eps(self.wavelength)) != 0) or (np.imag(self.medium.mu(self.wavelength)) != 0):
                raise ValueError("medium must be non-absorbing")

        self._set_wavelength(wavelength)
        self._multi = isinstance(source, (list, tuple))
        self.p_cluster = None
        self.report = None
        self.solve()

    # ---- solve ---------------------------------------------------------------------

    def _set_wavelength(self, wavelength):
        """Material data at `wavelength` and every sphere's Mie coefficients
        (computed on the device)."""
        self.wavelength = wavelength
        self.material_data = material_struct(self.material, self.medium, wavelength=self.wavelength)
        md = self.material_data
        conducting = np.array([getattr(m, "name", None) == "metal" for m in self.material])
        self._mie_d = D.mie_coefficients(self.radius, md.eps, md.mu, float(np.real(md.k_b)), md.eps_b, md.mu_b,
                                         self.lmax, conducting)
        self.mie_scat, self.mie_int = (a.cpu().numpy() for a in self._mie_d)

    def _sources(self):
        return list(self.source) if self._multi else [self.source]

    def solve(self, wavelength=None, source=None):
        """Solve for p_inc, p_scat, p_int and the observables
        (src/miepy/sphere_cluster.py:664-677).  A `wavelength` or `source`
        given here replaces the current one first (the material data and Mie
        coefficients follow the wavelength); the reference's signature accepts
        both and solves with its current state."""
        if source is not None:
            self.source = source
            self._multi = isinstance(source, (list, tuple))
            self.p_cluster = None
        if wavelength is not None and wavelength != self.wavelength:
            if np.imag(self.medium.eps(wavelength)) != 0 or np.imag(self.medium.mu(wavelength)) != 0:
                raise ValueError("medium must be non-absorbing")
            self._set_wavelength(wavelength)
            self.p_cluster = None
        md = self.material_data
        k = float(np.real(md.k_b))
        N, rmax, lmax = self.Nparticles, self.rmax, self.lmax
        srcs = self._sources()
        S = len(srcs)
        D.require_gpu()
        p_src_d = D.plane_wave_structure(srcs, self.position, k, lmax)  # [S, N, 2, rmax]
        p_src = p_src_d.cpu().numpy()
        mie_scat_d, mie_int_d = self._mie_d
        n = 2 * rmax * N
        if self.interactions and N > 1:
            exact = self.method == solver.exact or getattr(self.method, "name", None) == "exact"
            # exact without a stored matrix (matrix-free precisions): BiCGSTAB to 1e-14
            tol = 1e-14 if exact else (1e-5 if self.tol is None else self.tol)
            maxiter = max(self.maxiter, 10 * n) if exact else self.maxiter
            b = p_src_d.reshape(S, n)
            if exact and self.precision in ("fp64", "mixed"):
                # the reference's exact branch: LU of I + T (interactions.cpp:356-357), in FP64
                A = D.assemble(self.position, mie_scat_d, k, dtype=D.C128, add_identity=True)
                x = D.lu_solve(A, b, overwrite=True)
                self.report = None
                del A
            elif self.precision == "fp64":
                A = D.assemble(self.position, mie_scat_d, k, dtype=D.C128, add_identity=True)
                x, self.report = D.bicgstab(A, b, tol=tol, maxiter=maxiter)
                del A
            elif self.precision in ("matfree", "mixed_matfree"):
                op = D.ClusterOperator(self.position, mie_scat_d, k, max_rhs=S)
                if self.precision == "matfree":
                    x, self.report = D.bicgstab(op, b, tol=tol, maxiter=maxiter, add_identity=True)
                else:
                    rt = tol if self.refine_tol is None else self.refine_tol
                    x, self.report = D.solve_refined_matfree(op, b, tol=rt, maxiter=maxiter)
                del op
            else:
                A64, A32 = D.assemble_dual(self.position, mie_scat_d, k, add_identity=True)
                rt = tol if self.refine_tol is None else self.refine_tol
                x, self.report = D.solve_refined(A32, A64, b, tol=rt, maxiter=maxiter)
                del A32, A64
            p_inc_d = x.reshape(S, N, 2, rmax)
        else:
            p_inc_d = p_src_d.clone()
        obs = D.observables(p_inc_d, p_src_d, mie_scat_d, mie_int_d, k, md.eps_b)
        host = {key: (v.cpu().numpy() if v is not None else None) for key, v in obs.items()}
        p_inc = p_inc_d.cpu().numpy()
        sq = (lambda a: a) if self._multi else (lambda a: a[0])
        self.p_src = sq(p_src)
        self.p_inc = sq(p_inc)
        self.p_scat = sq(host["p_scat"])
        self.p_int = sq(host["p_int"])
        self._C_part = sq(host["C_part"])      # [(S,) N, 3, 2, lmax]
        self._C_total = sq(host["C_total"])    # [(S,) 3]
        self._F = sq(host["F"])                # [(S,) N, 3]
        self._T = sq(host["T"])

    def update_position(self, position):
        self.position = np.asarray(np.atleast_2d(position), dtype=float)
        self.p_cluster = None
        if self.auto_origin:
            self.origin = np.average(self.position, axis=0)
        self.solve()

    # ---- observables -----------------------------------------------------------------

    def cross_sections(self):
        """Cluster (scattering, absorption, extinction)."""
        C = self._C_total
        return _cs(C[..., 0], C[..., 1], C[..., 2])

    def solve_cluster_coefficients(self, lmax=None):
        """p_cluster [(S,) 2, rmax]: the cluster's scattered field expanded about
        self.origin up to lmax (src/miepy/sphere_cluster.py:651-662)."""
        lmax = self.lmax if lmax is None else lmax
        k = float(np.real(self.material_data.k_b))
        p_scat = self.p_scat if self._multi else self.p_scat[None]
        pc = D.cluster_coefficients(self.position, p_scat, k, self.origin, lmax).cpu().numpy()
        self.p_cluster = pc if self._multi else pc[0]
        return self.p_cluster

    def cross_sections_per_multipole(self, lmax=None):
        """Cluster (scattering, absorption, extinction) per multipole, each
        [(S,) 2, lmax], from the cluster coefficients about self.origin
        (src/miepy/sphere_cluster.py:428-440)."""
        lmax = self.lmax if lmax is None else lmax
        self.solve_cluster_coefficients(lmax)
        k = float(np.real(self.material_data.k_b))
        p0 = np.stack([s.structure(np.array([self.origin]), k, lmax)[0] for s in self._sources()])
        pc = self.p_cluster if self._multi else self.p_cluster[None]
        C = D.cluster_cross_sections(D.to_device(pc, D.C128), p0, k).cpu().numpy()
        C = C if self._multi else C[0]
        return _cs(C[..., 0, :, :], C[..., 1, :, :], C[..., 2, :, :])

    def cross_sections_per_multipole_of_particle(self, i):
        C = self._C_part[..., i, :, :, :]
        return _cs(C[..., 0, :, :], C[..., 1, :, :], C[..., 2, :, :])

    def cross_sections_of_particle(self, i):
        C = self.cross_sections_per_multipole_of_particle(i)
        return _cs(*[np.sum(A, axis=(-2, -1)) for A in C])

    def _forces(self, torque, source):
        if source:
            return self._T if torque else self._F
        from . import backend as B
        p_inc = self.p_inc - self.p_src
        md = self.material_data
        fn = B.torque_all if torque else B.force_all
        if self._multi:
            return np.stack([fn(self.p_scat[s], p_inc[s], md.k_b, md.eps_b, md.mu_b) for s in range(len(self.source))])
        return fn(self.p_scat, p_inc, md.k_b, md.eps_b, md.mu_b)

    def force(self, source=True):
        """F[N, 3] on every particle."""
        return self._forces(False, source)

    def torque(self, source=True):
        """T[N, 3] on every particle."""
        return self._forces(True, source)

    def force_on_particle(self, i, source=True):
        return self.force(source)[..., i, :]

    def torque_on_particle(self, i, source=True):
        return self.torque(source)[..., i, :]

    def __repr__(self):
        return (f"{self.__class__.__name__}:\n    Nparticles = {self.Nparticles}\n    source = {self.source}\n"
                f"    wavelength = {self.wavelength:.2e}\n    medium = {self.medium}\n    lmax = {self.lmax}\n"
                f"    origin = {self.origin}\n    precision = {self.precision}\n")
