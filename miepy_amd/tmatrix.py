"""T-matrix of a rigid sphere cluster about the origin (miepy.tmatrix,
src/miepy/tmatrix/common.py:298-364; tmatrix_reduce_lmax,
src/miepy/tmatrix/functions.py:6-24).

    T = R diag(mie) (I + T_agg)^-1 S

with the origin's modes as the right-hand sides: S [n, 2 rmax_out] expands
every origin mode (b, n, m) about every sphere (incident-mode translation of
r_i - origin), the cluster system (I + T_agg) X = S is solved for all
2 rmax_out columns at once, the Mie coefficients scale the solution and R
re-expands every sphere's scattered modes about the origin.  Every step runs
on the device: Mie coefficients, the stored aggregate matrix, S, the solve
(dense LU as the reference's numpy.linalg.solve, or batched BiCGSTAB with one
right-hand side per origin mode), and R diag(mie) X in one pass.
"""

import numpy as np

from . import device as D


def tmatrix_sphere_cluster(pos, radii, lmax, lmax_cluster, wavelength, eps, eps_m, method="exact", tol=1e-12,
                           maxiter=5000, **kwargs):
    """Cluster T-matrix [2, rmax(lmax_cluster), 2, rmax(lmax_cluster)] (host numpy).

    pos [N, 3]; radii, lmax (per-sphere lmax; coefficients above a sphere's
    lmax are zero) and eps scalar or [N]; eps_m the (real) medium
    permittivity.  method "exact" (dense LU, the reference's solve) or
    "bicgstab" (to `tol` relative, every origin mode a right-hand side of
    one batched solve).  Other keywords are ignored, as in the reference."""
    pos = np.asarray(pos, dtype=float).reshape(-1, 3)
    N = pos.shape[0]
    radii = np.broadcast_to(np.atleast_1d(np.asarray(radii, dtype=float)), (N,)).copy()
    lmax_arr = np.broadcast_to(np.atleast_1d(np.asarray(lmax, dtype=int)), (N,)).copy()
    eps_arr = np.broadcast_to(np.atleast_1d(np.asarray(eps, dtype=complex)), (N,)).copy()
    if np.imag(eps_m) != 0:
        raise ValueError("the medium must be non-absorbing (real eps_m)")
    eps_m = float(np.real(eps_m))
    if method not in ("exact", "bicgstab"):
        raise ValueError("method must be 'exact' or 'bicgstab'")
    lmax_p = int(np.max(lmax_arr))
    k = 2 * np.pi * np.sqrt(eps_m) / wavelength
    origin = np.zeros(3)
    D.require_gpu()
    mie, _ = D.mie_coefficients(radii, eps_arr, np.ones(N), k, eps_m, 1.0, lmax_p)
    if np.any(lmax_arr < lmax_p):
        keep = np.arange(1, lmax_p + 1)[None, :] <= lmax_arr[:, None]  # [N, lmax_p]
        mie = mie * D.to_device(keep.astype(float), mie.real.dtype).unsqueeze(1)
    A = D.assemble(pos, mie, k, add_identity=True)
    S = D.cluster_source(pos, k, origin, lmax_p, lmax_cluster)
    if method == "exact":
        X = D.lu_solve(A, S, overwrite=True)
    else:
        X, rep = D.bicgstab(A, S, tol=tol, maxiter=maxiter)
        if not rep.converged:
            raise RuntimeError(f"tmatrix_sphere_cluster: BiCGSTAB did not converge (status {rep.status})")
    del A
    return D.cluster_tmatrix(pos, mie, X, k, origin, lmax_cluster).cpu().numpy()


def tmatrix_reduce_lmax(tmatrix, lmax):
    """The leading [2, rmax(lmax), 2, rmax(lmax)] part of a T-matrix."""
    rmax_original = tmatrix.shape[1]
    lmax_original = int(round((1 + rmax_original) ** 0.5)) - 1
    if lmax > lmax_original:
        raise ValueError(f"cannot reduce tmatrix to lmax={lmax} since it is already smaller than this")
    rmax = lmax * (lmax + 2)
    return tmatrix[:, :rmax, :, :rmax]
