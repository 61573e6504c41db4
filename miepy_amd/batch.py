"""Many independent cluster solves in one batched launch sequence.

Each cluster (positions, radii, permittivities, wavelength, lmax, sources)
becomes one system (I + T_c) p_inc = p_src with one right-hand side per
source.  Host work (Mie coefficients, p_src) is done once in `prepare`; the
device work (assembly of every matrix, one batched BiCGSTAB or one batched
complex64 + FP64-refinement solve over all clusters, one batched matrix-free
FP64 BiCGSTAB with no stored matrices, or one batched refinement whose
complex64 corrections and FP64 residuals both use the matrix-free operator
for clusters with lmax <= 5 and n >= matfree_min_n (each matrix-free
system launches its own operator kernels, so small clusters keep stored
matrices, whose products run as one batched launch); observables) is
`ClusterBatch.solve`.  The batched engine covers all clusters with one launch
per kernel and iteration, and a cluster that converged costs nothing further.
"""

import numpy as np
import torch

from . import device as D


class PreparedCluster:
    """Device-resident inputs of one cluster workload."""

    def __init__(self, w):
        self.w = w
        k = w.k
        conducting = None if w.conducting is None else np.asarray(w.conducting)
        self.pos = D.to_device(w.position, torch.float64)
        self.mie_scat, self.mie_int = D.mie_coefficients(w.radius, w.eps, np.ones(w.N), k, w.medium_eps, 1.0, w.lmax,
                                                         conducting)
        self.p_src = D.plane_wave_structure(list(w.sources), self.pos, k, w.lmax)
        self.S = self.p_src.shape[0]
        self.n = w.n
        self.rhs = self.p_src.reshape(self.S, self.n)  # one right-hand side per source
        self.k = k
        self.eps_b = float(np.real(w.medium_eps))
        self.A64 = None
        self.A32 = None
        self.op = None

    def matrix_bytes(self, precision, matfree_min_n=6000):
        if precision == "matfree" or (precision == "mixed_matfree" and self.w.lmax <= 5 and
                                      self.n >= matfree_min_n):
            return 0
        return 16 * self.n * self.n + (8 * self.n * self.n if precision.startswith("mixed") else 0)


class ClusterBatch:
    def __init__(self, workloads, precision="mixed", inner_tol=1e-5, maxiter=2000, max_refine=8, tol_scale=1.0,
                 matfree_min_n=6000):
        if precision not in ("fp64", "mixed", "matfree", "mixed_matfree"):
            raise ValueError("precision must be 'fp64', 'mixed', 'matfree' or 'mixed_matfree'")
        self.precision = precision
        self.inner_tol = inner_tol
        self.maxiter = maxiter
        self.max_refine = max_refine
        self.tol_scale = tol_scale  # solve to tol_scale x each workload's tolerance (reference solutions)
        self.clusters = [PreparedCluster(w) for w in workloads]
        for c in self.clusters:
            if precision == "matfree" or (precision == "mixed_matfree" and c.w.lmax <= 5 and
                                          c.n >= matfree_min_n):
                c.op = D.ClusterOperator(c.pos, c.mie_scat, c.k, max_rhs=c.S)
                continue
            c.A64 = torch.empty((c.n, c.n), dtype=D.C128, device=c.pos.device)
            if precision.startswith("mixed"):
                c.A32 = torch.empty((c.n, c.n), dtype=D.C64, device=c.pos.device)
        self.reports = None
        self.results = None
        self._obs_cache = {}
        self._asm_cache = {}

    def autotune(self, reps=3):
        """Pick the matvec tile of every (dtype, RHS-count class) present on
        one batched launch over this batch's stored matrices (their contents
        need not be assembled yet).  Returns {(dtype, S): (variant, ms)}."""
        out = {}
        for attr, dt in (("A64", "c128"), ("A32", "c64")):
            mats = [c for c in self.clusters if getattr(c, attr) is not None]
            for S in sorted({c.S for c in mats}):
                group = [getattr(c, attr) for c in mats if c.S == S]
                out[(dt, S)] = D.autotune_matvec_batched(group, S, reps=reps)
        return out

    def solve(self):
        """Assemble, solve and reduce every cluster; results stay on the device."""
        cs = self.clusters
        tols = [c.w.tol * self.tol_scale for c in cs]
        rhs = [c.rhs for c in cs]
        if self.precision == "matfree":
            xs, self.reports = D.bicgstab_batched([c.op for c in cs], rhs, tol=tols, maxiter=self.maxiter,
                                                  add_identity=True)
            return self._reduce(xs)
        stored = [c for c in cs if c.op is None]
        if stored:  # every stored matrix (and its complex64 rounding) in a few batched launches
            D.assemble_batched([(c.pos, c.mie_scat, c.k, c.A64, c.A32) for c in stored], add_identity=True,
                               cache=self._asm_cache)
        if self.precision == "fp64":
            xs, self.reports = D.bicgstab_batched([c.A64 for c in cs], rhs, tol=tols, maxiter=self.maxiter)
        else:
            xs, self.reports = D.solve_refined_batched([c.op if c.op is not None else c.A32 for c in cs],
                                                       [None if c.op is not None else c.A64 for c in cs], rhs,
                                                       tol=tols, inner_tol=self.inner_tol, maxiter=self.maxiter,
                                                       max_refine=self.max_refine)
        return self._reduce(xs)

    def _reduce(self, xs):
        """Observables of every cluster in a few batched launches (the
        clusters' static descriptor columns are kept between solves)."""
        cs = self.clusters
        self.results = D.observables_batched([(x, c.p_src, c.mie_scat, c.mie_int, c.k, c.eps_b)
                                              for c, x in zip(cs, xs)], cache=self._obs_cache)
        for obs, c, x in zip(self.results, cs, xs):
            obs.defer("p_inc", x, (c.S, c.w.N, 2, c.w.rmax))
        return self.results

    def converged(self):
        return all(r.converged for r in self.reports)

    def host_results(self):
        return [{k: (v.cpu().numpy() if v is not None else None) for k, v in r.items()} for r in self.results]
