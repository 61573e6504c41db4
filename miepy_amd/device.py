"""Device-level operations: torch tensors in HBM -> libmiepy_amd C ABI.

torch provides device memory and the current HIP stream; every computation is
a hand-written gfx950 kernel behind include/miepy_amd.h.  There is no CPU
path: without a visible GPU or without the built library these raise.
"""

import ctypes
from collections.abc import MutableMapping

import numpy as np
import torch

from . import _lib as L

C128 = torch.complex128
C64 = torch.complex64
DTYPE_CODE = {C128: L.MPA_C128, C64: L.MPA_C64}


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("miepy_amd needs a HIP GPU (MI355X); none is visible")
    return torch.device("cuda", torch.cuda.current_device())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def to_device(a, dtype=None):
    dev = require_gpu()
    if isinstance(a, torch.Tensor):
        t = a.to(dev)
    else:
        t = torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    if dtype is not None:
        t = t.to(dtype)
    return t.contiguous()


def lmax_to_rmax(lmax):
    return lmax * (lmax + 2)


def assemble(positions, mie, k, dtype=C128, add_identity=False, out=None):
    """Aggregate T-matrix [n, n] (or I + T) on the device.

    positions [N, 3] f64, mie [N, 2, lmax] c128 (device tensors or arrays)."""
    lib = L.load()
    pos = to_device(positions, torch.float64).reshape(-1, 3)
    N = pos.shape[0]
    mie_d = to_device(mie, C128).reshape(N, 2, -1)
    lmax = mie_d.shape[-1]
    n = 2 * lmax_to_rmax(lmax) * N
    if out is None:
        out = torch.empty((n, n), dtype=dtype, device=pos.device)
    elif out.shape != (n, n) or out.dtype != dtype or not out.is_contiguous():
        raise ValueError("out must be a contiguous [n, n] tensor of the requested dtype")
    L.check(lib.mpa_sphere_aggregate_tmatrix(N, lmax, _ptr(pos), _ptr(mie_d), float(k), DTYPE_CODE[dtype],
                                             1 if add_identity else 0, _ptr(out), _stream()),
            "sphere_aggregate_tmatrix")
    return out


def assemble_tmatrix(positions, tmatrix, k, dtype=C128, add_identity=False, out=None):
    """Aggregate T-matrix [n, n] (or I + T) of particles with general
    T-matrices: positions [N, 3] f64, tmatrix [N, 2, rmax, 2, rmax] c128."""
    lib = L.load()
    pos = to_device(positions, torch.float64).reshape(-1, 3)
    N = pos.shape[0]
    tm = to_device(tmatrix, C128).reshape(N, -1)
    bs = int(round(tm.shape[1] ** 0.5))
    if bs * bs != tm.shape[1] or bs % 2:
        raise ValueError("tmatrix must be [N, 2, rmax, 2, rmax]")
    rmax = bs // 2
    lmax = int(round((rmax + 1) ** 0.5)) - 1
    if lmax * (lmax + 2) != rmax:
        raise ValueError(f"rmax={rmax} is not lmax(lmax+2)")
    tm = tm.reshape(N, bs, bs).contiguous()
    n = bs * N
    if out is None:
        out = torch.empty((n, n), dtype=dtype, device=pos.device)
    L.check(lib.mpa_particle_aggregate_tmatrix(N, lmax, _ptr(pos), _ptr(tm), float(k), DTYPE_CODE[dtype],
                                               1 if add_identity else 0, _ptr(out), _stream()),
            "particle_aggregate_tmatrix")
    return out


def tmatrix_apply(tmatrix, x, out=None):
    """y_j = Tm_j x_j per particle: tmatrix [N, 2, rmax, 2, rmax], x [S, N, 2, rmax] (or [N, 2, rmax])."""
    lib = L.load()
    tm = to_device(tmatrix, C128)
    N = tm.shape[0]
    bs = int(round(tm.reshape(N, -1).shape[1] ** 0.5))
    rmax = bs // 2
    lmax = int(round((rmax + 1) ** 0.5)) - 1
    xd = to_device(x, C128)
    x2 = xd.reshape(-1, N * bs)
    if out is None:
        out = torch.empty_like(x2)
    L.check(lib.mpa_tmatrix_apply(N, lmax, x2.shape[0], _ptr(tm), _ptr(x2), _ptr(out), _stream()), "tmatrix_apply")
    return out.reshape(xd.shape)


def assemble_dual(positions, mie, k, add_identity=False, out64=None, out32=None):
    """One assembly pass writing the c128 matrix and its complex64 rounding."""
    lib = L.load()
    pos = to_device(positions, torch.float64).reshape(-1, 3)
    N = pos.shape[0]
    mie_d = to_device(mie, C128).reshape(N, 2, -1)
    n = 2 * lmax_to_rmax(mie_d.shape[-1]) * N
    if out64 is None:
        out64 = torch.empty((n, n), dtype=C128, device=pos.device)
    if out32 is None:
        out32 = torch.empty((n, n), dtype=C64, device=pos.device)
    if out64.shape != (n, n) or out32.shape != (n, n) or out64.dtype != C128 or out32.dtype != C64:
        raise ValueError("out64 / out32 must be [n, n] complex128 / complex64")
    L.check(lib.mpa_sphere_aggregate_tmatrix_dual(N, mie_d.shape[-1], _ptr(pos), _ptr(mie_d), float(k),
                                                  1 if add_identity else 0, _ptr(out64), _ptr(out32), _stream()),
            "sphere_aggregate_tmatrix_dual")
    return out64, out32


def assemble_batched(clusters, add_identity=True, cache=None):
    """Aggregate T-matrices of many clusters in a few launches
    (mpa_sphere_aggregate_tmatrix_batched); each equals its own assemble() /
    assemble_dual() call bit for bit.  clusters: [(positions [N, 3] f64
    device, mie [N, 2, lmax] c128 device, k, out, out32 or None)] with out
    [n, n] complex128 or complex64 (out32: the complex64 rounding of a
    complex128 out, written in the same pass).  cache: a dict kept by a
    caller that assembles the same tensors again (the same clusters, inputs
    and outputs): their checked descriptors are reused."""
    lib = L.load()
    key = (len(clusters), bool(add_identity))
    if cache is not None and cache.get("key") == key:
        L.check(lib.mpa_sphere_aggregate_tmatrix_batched(len(clusters), cache["descs"], _stream()),
                "sphere_aggregate_tmatrix_batched")
        return
    descs = (L.MpaAsmCluster * len(clusters))()
    for q, (pos, mie, k, out, out32) in enumerate(clusters):
        N = pos.shape[0]
        lmax = mie.shape[-1]
        n = 2 * lmax_to_rmax(lmax) * N
        if (tuple(out.shape) != (n, n) or out.dtype not in (C128, C64) or not out.is_contiguous()
                or (out32 is not None and (tuple(out32.shape) != (n, n) or out32.dtype != C64 or out.dtype != C128))):
            raise ValueError("out must be a contiguous [n, n] complex128/complex64 tensor (out32: complex64 beside "
                             "a complex128 out)")
        if pos.dtype != torch.float64 or mie.dtype != C128 or not pos.is_contiguous() or not mie.is_contiguous():
            raise ValueError("positions must be contiguous float64 and mie contiguous complex128 device tensors")
        d = descs[q]
        d.N, d.lmax, d.dtype, d.add_identity = N, lmax, DTYPE_CODE[out.dtype], 1 if add_identity else 0
        d.positions, d.mie, d.k = pos.data_ptr(), mie.data_ptr(), float(k)
        d.out = out.data_ptr()
        d.out32 = out32.data_ptr() if out32 is not None else None
    L.check(lib.mpa_sphere_aggregate_tmatrix_batched(len(clusters), descs, _stream()), "sphere_aggregate_tmatrix_batched")
    if cache is not None:
        cache["key"], cache["descs"] = key, descs


def matvec(A, x, add_identity=False, out=None):
    """y = (I +) A x; x is [n] or [S, n] on the device, same dtype as A."""
    lib = L.load()
    n = A.shape[0]
    x2 = x.reshape(-1, n)
    if x2.dtype != A.dtype:
        raise ValueError("x must have the matrix dtype")
    y = torch.empty_like(x2) if out is None else out
    L.check(lib.mpa_matvec(n, x2.shape[0], DTYPE_CODE[A.dtype], _ptr(A), 1 if add_identity else 0, _ptr(x2.contiguous()),
                           _ptr(y), _stream()), "matvec")
    return y.reshape(x.shape)


def autotune_matvec(A, nrhs=1, reps=3):
    """Time every matvec tile variant for A's dtype and the RHS-count class of
    `nrhs` on A itself and keep the fastest for the process.  Returns
    (variant name, ms per launch)."""
    lib = L.load()
    n = A.shape[0]
    code = DTYPE_CODE[A.dtype]
    g = torch.Generator(device=A.device).manual_seed(0)
    x = torch.randn((nrhs, n), dtype=A.dtype, device=A.device, generator=g)
    y = torch.empty_like(x)
    v = ctypes.c_int(0)
    ms = ctypes.c_double(0)
    L.check(lib.mpa_autotune_matvec(n, nrhs, code, _ptr(A), _ptr(x), _ptr(y), int(reps), ctypes.byref(v),
                                    ctypes.byref(ms), _stream()), "autotune_matvec")
    return lib.mpa_matvec_variant_name(code, nrhs, v.value).decode(), ms.value


def autotune_matvec_batched(matrices, nrhs, reps=3):
    """Time every matvec tile variant of the RHS-count class of `nrhs` on one
    batched launch over `matrices` (the launch shape of the batched solvers)
    and keep the fastest for the process.  Returns (variant name, ms per
    batched launch)."""
    lib = L.load()
    mats = [A for A in matrices]
    if not mats or any(A.dtype != mats[0].dtype for A in mats):
        raise ValueError("autotune_matvec_batched: one dtype for every matrix")
    code = DTYPE_CODE[mats[0].dtype]
    g = torch.Generator(device=mats[0].device).manual_seed(0)
    xs = [torch.randn((nrhs, A.shape[0]), dtype=A.dtype, device=A.device, generator=g) for A in mats]
    ys = [torch.empty_like(x) for x in xs]
    cnt = len(mats)
    arr = lambda ts: (ctypes.c_void_p * cnt)(*[t.data_ptr() for t in ts])
    v = ctypes.c_int(0)
    ms = ctypes.c_double(0)
    L.check(lib.mpa_autotune_matvec_batched(cnt, (ctypes.c_int * cnt)(*[A.shape[0] for A in mats]),
                                            (ctypes.c_int * cnt)(*([int(nrhs)] * cnt)), code, arr(mats), arr(xs),
                                            arr(ys), int(reps), ctypes.byref(v), ctypes.byref(ms), _stream()),
            "autotune_matvec_batched")
    return lib.mpa_matvec_variant_name(code, nrhs, v.value).decode(), ms.value


def autotune_assembly(positions, mie, k, out, reps=2):
    """Time every assembly tile variant on this cluster (`out` [n, n] is
    overwritten) and keep the fastest for the process.  Returns
    (tile column blocks in use, ms per launch)."""
    lib = L.load()
    pos = to_device(positions, torch.float64).reshape(-1, 3)
    N = pos.shape[0]
    mie_d = to_device(mie, C128).reshape(N, 2, -1)
    lmax = mie_d.shape[-1]
    v = ctypes.c_int(0)
    ms = ctypes.c_double(0)
    L.check(lib.mpa_autotune_assembly(N, lmax, _ptr(pos), _ptr(mie_d), float(k), DTYPE_CODE[out.dtype], _ptr(out),
                                      int(reps), ctypes.byref(v), ctypes.byref(ms), _stream()), "autotune_assembly")
    return lib.mpa_assembly_tile(lmax), ms.value


class SolveReport:
    def __init__(self, iters, resid, status):
        self.iterations = iters
        self.residual = resid
        self.status = status

    @property
    def converged(self):
        return all(s == L.MPA_OK for s in self.status)

    def __repr__(self):
        return f"SolveReport(iterations={self.iterations}, residual={self.residual}, status={self.status})"


def bicgstab(A, b, tol=1e-5, maxiter=1000, add_identity=False, x=None, workspace=None, sync_every=1, check=True,
             M_inv_blocks=None, block_size=0, ref_loop=False):
    """BiCGSTAB (Eigen's algorithm) for (I +) A x = b on the device.

    b: [n] or [S, n] with A's dtype.  Returns (x, SolveReport).
    M_inv_blocks: [n, block_size] device tensor of A's dtype, the stacked
    inverses of the diagonal blocks, applied as a right preconditioner
    (cpp/src/interactions.cpp:291-345).  ref_loop: the reference's own loop
    (interactions.cpp:98-209, :230-288): x0 = b, tol absolute on ||r||, at
    most maxiter + 1 iterations, no restart; the report's residuals stay
    ||r||/||b||."""
    lib = L.load()
    n = A.shape[0]
    b2 = b.reshape(-1, n).contiguous()
    S = b2.shape[0]
    if isinstance(A, ClusterOperator):
        if M_inv_blocks is not None or ref_loop:
            raise NotImplementedError("preconditioned and reference-loop solves take a stored matrix")
        return _bicgstab_matfree(lib, A, b, b2, tol, maxiter, add_identity, x, workspace, sync_every, check)
    code = DTYPE_CODE[A.dtype]
    if b2.dtype != A.dtype:
        raise ValueError("b must have the matrix dtype")
    if M_inv_blocks is not None:
        if M_inv_blocks.dtype != A.dtype or not M_inv_blocks.is_cuda or M_inv_blocks.numel() != n * block_size \
                or block_size < 1 or n % block_size:
            raise ValueError("M_inv_blocks must be a device [n, block_size] tensor of the matrix dtype, "
                             "block_size dividing n")
        M_inv_blocks = M_inv_blocks.contiguous()
    pc = 1 if M_inv_blocks is not None else 0
    nbytes = lib.mpa_bicgstab_ex_workspace_bytes(n, S, code, pc)
    if workspace is None or workspace.numel() < nbytes:
        workspace = torch.empty(nbytes, dtype=torch.uint8, device=A.device)
    if x is None:
        x = torch.empty_like(b2)
    it = (ctypes.c_int * S)()
    res = (ctypes.c_double * S)()
    st = (ctypes.c_int * S)()
    rc = lib.mpa_bicgstab_ex(n, S, code, _ptr(A), 1 if add_identity else 0, _ptr(b2), _ptr(x), float(tol),
                             int(maxiter), L.MPA_SOLVE_REF_LOOP if ref_loop else 0,
                             _ptr(M_inv_blocks) if pc else None, int(block_size) if pc else 0,
                             _ptr(workspace), it, res, st, int(sync_every), _stream())
    report = SolveReport(list(it), list(res), list(st))
    if rc not in (L.MPA_OK, L.MPA_NOT_CONVERGED, L.MPA_NONFINITE) or (check and rc == L.MPA_NONFINITE):
        L.check(rc, "bicgstab")
    return x.reshape(b.shape), report


def lu_solve(A, b, add_identity=False, overwrite=False):
    """Dense LU with partial pivoting on the device (the `exact` method):
    x = (I +) A^-1 b.  A: [n, n] complex128; b: [n] or [S, n].  A is copied
    unless overwrite (then it holds the factors afterwards).  Raises
    numpy.linalg.LinAlgError for a singular matrix.  Returns x like b."""
    lib = L.load()
    if A.dtype != C128 or b.dtype != C128:
        raise ValueError("lu_solve is complex128")
    n = A.shape[0]
    F = A if overwrite else A.clone()
    b2 = b.reshape(-1, n).contiguous()
    ws = torch.empty(lib.mpa_lu_workspace_bytes(n), dtype=torch.uint8, device=A.device)
    info = ctypes.c_int(0)
    rc = lib.mpa_lu_factor(n, _ptr(F), 1 if add_identity else 0, _ptr(ws), ctypes.byref(info), _stream())
    if rc == L.MPA_BREAKDOWN:
        raise np.linalg.LinAlgError(f"singular matrix: no nonzero pivot in column {info.value - 1}")
    L.check(rc, "lu_factor")
    x = torch.empty_like(b2)
    L.check(lib.mpa_lu_solve(n, b2.shape[0], _ptr(F), _ptr(ws), _ptr(b2), _ptr(x), _stream()), "lu_solve")
    return x.reshape(b.shape)


def mie_coefficients(radius, eps, mu, k, eps_b, mu_b, lmax, conducting=None):
    """Mie coefficients of N spheres on the device (src/miepy/mie_single/mie_sphere.py:16-76):
    returns (mie_scat, mie_int), each [N, 2, lmax] complex128 (device).
    radius, eps, mu: [N]; k, eps_b, mu_b: scalars or [N] (one value per
    particle); conducting: [N] bool (perfect-conductor limit) or None."""
    lib = L.load()
    dev = require_gpu()
    radius = np.atleast_1d(np.asarray(radius, dtype=float))
    N = radius.shape[0]
    c = lambda a: np.broadcast_to(np.asarray(a, dtype=complex), (N,))
    r = lambda a: np.broadcast_to(np.asarray(np.real(a), dtype=float), (N,))
    reals = to_device(np.stack([radius, r(k)]), torch.float64)
    cplx = to_device(np.stack([c(eps), c(mu), c(eps_b), c(mu_b)]), C128)
    cond = None
    if conducting is not None and np.any(conducting):
        cond = torch.from_numpy(np.broadcast_to(np.asarray(conducting, dtype=np.int32), (N,)).copy()).to(dev)
    ms = torch.empty((N, 2, lmax), dtype=C128, device=dev)
    mi = torch.empty((N, 2, lmax), dtype=C128, device=dev)
    L.check(lib.mpa_mie_coefficients(N, int(lmax), _ptr(reals[0]), _ptr(cplx[0]), _ptr(cplx[1]), _ptr(reals[1]),
                                     _ptr(cplx[2]), _ptr(cplx[3]), _ptr(cond) if cond is not None else None,
                                     _ptr(ms), _ptr(mi), _stream()), "mie_coefficients")
    return ms, mi


def plane_wave_structure(sources, positions, k, lmax):
    """p_src [S, N, 2, rmax] complex128 (device) of the given sources (plane
    waves or sums of them) at the particle positions
    (src/miepy/sources/plane_waves.py:126-156).  The per-wave mode tables are
    built on the host (rmax values each); the per-particle phases and
    products run on the device."""
    lib = L.load()
    dev = require_gpu()
    pos = positions if isinstance(positions, torch.Tensor) else to_device(
        np.asarray(positions, dtype=float).reshape(-1, 3), torch.float64)
    pos = pos.to(torch.float64).reshape(-1, 3).contiguous()
    N = pos.shape[0]
    rmax = lmax_to_rmax(lmax)
    waves, begin = [], [0]
    for src in sources:
        waves += src.waves()
        begin.append(len(waves))
    W = len(waves)
    table = np.stack([w.mode_table(lmax) for w in waves])
    khat = np.stack([np.asarray(w.k_hat, dtype=float) for w in waves])
    phase = np.array([float(np.real(w.phase)) for w in waves])
    packed = to_device(np.concatenate([khat.reshape(-1), phase]), torch.float64)
    tab = to_device(table, C128)
    wb = torch.tensor(begin, dtype=torch.int32).to(dev)
    out = torch.empty((len(sources), N, 2, rmax), dtype=C128, device=dev)
    L.check(lib.mpa_plane_wave_structure(N, int(lmax), len(sources), _ptr(pos), float(np.real(k)), _ptr(wb),
                                         _ptr(packed), _ptr(packed[3 * W:]), _ptr(tab), _ptr(out), _stream()),
            "plane_wave_structure")
    return out


def cluster_coefficients(positions, p_scat, k, origin, lmax):
    """Scattered-field coefficients of the whole cluster about `origin`
    (src/miepy/vsh/cluster_coefficients.py:14-53) on the device.

    positions [N, 3]; p_scat [S, N, 2, rmax_in] complex128 (device or host);
    lmax: the output order (<= 6).  Returns p_cluster [S, 2, rmax] (device)."""
    lib = L.load()
    require_gpu()
    pos = to_device(np.asarray(positions, dtype=float).reshape(-1, 3), torch.float64)
    ps = p_scat if isinstance(p_scat, torch.Tensor) else to_device(np.asarray(p_scat, dtype=complex), C128)
    ps = ps.to(C128).contiguous()
    S, N, _, rmax_in = ps.shape
    if N != pos.shape[0]:
        raise ValueError("p_scat and positions disagree on the number of particles")
    lmax_in = int(round((1 + rmax_in) ** 0.5)) - 1
    rmax = lmax_to_rmax(lmax)
    out = torch.empty((S, 2, rmax), dtype=C128, device=ps.device)
    ws = torch.empty(lib.mpa_cluster_coefficients_workspace_bytes(N, lmax, S), dtype=torch.uint8, device=ps.device)
    o = np.asarray(origin, dtype=float).reshape(3)
    L.check(lib.mpa_cluster_coefficients(N, lmax_in, int(lmax), S, _ptr(pos), float(o[0]), float(o[1]), float(o[2]),
                                         _ptr(ps), float(k), _ptr(out), _ptr(ws), _stream()), "cluster_coefficients")
    return out


def cluster_source(positions, k, origin, lmax_p, lmax_out):
    """The source matrix of tmatrix_sphere_cluster (src/miepy/tmatrix/common.py:199-247)
    on the device, stored one origin mode per row: [2 rmax_out, N * 2 rmax_p]
    complex128 (device), the right-hand sides of (I + T) X = S."""
    lib = L.load()
    require_gpu()
    pos = to_device(np.asarray(positions, dtype=float).reshape(-1, 3), torch.float64)
    N = pos.shape[0]
    out = torch.empty((2 * lmax_to_rmax(lmax_out), N * 2 * lmax_to_rmax(lmax_p)), dtype=C128, device=pos.device)
    o = np.asarray(origin, dtype=float).reshape(3)
    L.check(lib.mpa_cluster_source(N, int(lmax_p), int(lmax_out), _ptr(pos), float(o[0]), float(o[1]), float(o[2]),
                                   float(k), _ptr(out), _stream()), "cluster_source")
    return out


def cluster_tmatrix(positions, mie, X, k, origin, lmax_out):
    """T = R diag(mie) X (src/miepy/tmatrix/common.py:250-295, 352-364): mie
    [N, 2, lmax_p] and X [2 rmax_out, N * 2 rmax_p] complex128 (device).
    Returns T [2, rmax_out, 2, rmax_out] (device)."""
    lib = L.load()
    require_gpu()
    pos = to_device(np.asarray(positions, dtype=float).reshape(-1, 3), torch.float64)
    N = pos.shape[0]
    mie = mie.to(C128).contiguous()
    lmax_p = mie.shape[2]
    rows = 2 * lmax_to_rmax(lmax_out)
    if mie.shape[:2] != (N, 2) or X.shape != (rows, N * 2 * lmax_to_rmax(lmax_p)) or X.dtype != C128:
        raise ValueError("cluster_tmatrix: mie must be [N, 2, lmax_p] and X [2 rmax_out, N * 2 rmax_p] complex128")
    X = X.contiguous()
    out = torch.empty((2, rows // 2, 2, rows // 2), dtype=C128, device=pos.device)
    ws = torch.empty(lib.mpa_cluster_coefficients_workspace_bytes(N, int(lmax_out), rows), dtype=torch.uint8,
                     device=pos.device)
    o = np.asarray(origin, dtype=float).reshape(3)
    L.check(lib.mpa_cluster_tmatrix(N, int(lmax_p), int(lmax_out), _ptr(pos), float(o[0]), float(o[1]), float(o[2]),
                                    _ptr(mie), _ptr(X), float(k), _ptr(out), _ptr(ws), _stream()), "cluster_tmatrix")
    return out


def vsh_translation(mnuv, rad, theta, phi, k, mode=0):
    """Translation coefficients on the device: mnuv [count, 4] ints (m, n, u,
    v); rad, theta, phi [nvec] (the translation vectors); mode 0 outgoing,
    1 ingoing, 2 incident, 3 interior.  Returns [nvec, count, 2] complex128
    (A, B) on the host (cpp/src/vsh_translation.cpp:366-470)."""
    lib = L.load()
    dev = require_gpu()
    mnuv = np.ascontiguousarray(np.asarray(mnuv, dtype=np.int32).reshape(-1, 4))
    n, v = mnuv[:, 1], mnuv[:, 3]
    if np.any(n < 1) or np.any(v < 1) or np.any(np.abs(mnuv[:, 0]) > n) or np.any(np.abs(mnuv[:, 2]) > v):
        raise ValueError("vsh_translation: need 1 <= n, v and |m| <= n, |u| <= v")
    lmax = int(max(n.max(), v.max()))
    geo = np.stack([np.atleast_1d(np.asarray(a, dtype=float)) for a in np.broadcast_arrays(rad, theta, phi)])
    nvec = geo.shape[1]
    g = to_device(geo, torch.float64)
    idx = torch.from_numpy(mnuv).to(dev)
    out = torch.empty((nvec, mnuv.shape[0], 2), dtype=C128, device=dev)
    L.check(lib.mpa_vsh_translation(mnuv.shape[0], _ptr(idx), lmax, nvec, _ptr(g[0]), _ptr(g[1]), _ptr(g[2]),
                                    float(k), int(mode), _ptr(out), _stream()), "vsh_translation")
    return out.cpu().numpy()


def cluster_cross_sections(p_cluster, p_src, k):
    """Per-multipole (scattering, absorption, extinction) of the cluster
    (cpp/src/flux.cpp:32-57) on the device: p_cluster, p_src [S, 2, rmax]
    complex128 (p_src: the source's coefficients at the origin).  Returns
    C [S, 3, 2, lmax] float64 (device)."""
    lib = L.load()
    pc = p_cluster.to(C128).contiguous()
    p0 = p_src if isinstance(p_src, torch.Tensor) else to_device(np.asarray(p_src, dtype=complex), C128)
    p0 = p0.to(C128).reshape(pc.shape).contiguous()
    S, _, rmax = pc.shape
    lmax = int(round((1 + rmax) ** 0.5)) - 1
    C = torch.empty((S, 3, 2, lmax), dtype=torch.float64, device=pc.device)
    L.check(lib.mpa_cluster_cross_sections(lmax, S, _ptr(pc), _ptr(p0), float(k), _ptr(C), _stream()),
            "cluster_cross_sections")
    return C


def observables(p_inc, p_src, mie_scat, mie_int, k, eps_b, want_int=True, want_torque=True):
    """Per-source observables of a solved cluster (all FP64, on the device).

    p_inc, p_src: [S, N, 2, rmax] c128.  Returns a dict of device tensors:
    p_scat, p_int, C_part [S, N, 3, 2, lmax], C_total [S, 3], F [S, N, 3], T [S, N, 3]."""
    lib = L.load()
    S, N, _, rmax = p_inc.shape
    lmax = mie_scat.shape[-1]
    dev = p_inc.device
    p_scat = torch.empty_like(p_inc)
    p_int = torch.empty_like(p_inc) if want_int else None
    C_part = torch.empty((S, N, 3, 2, lmax), dtype=torch.float64, device=dev)
    C_total = torch.empty((S, 3), dtype=torch.float64, device=dev)
    F = torch.empty((S, N, 3), dtype=torch.float64, device=dev)
    T = torch.empty((S, N, 3), dtype=torch.float64, device=dev) if want_torque else None
    nullp = ctypes.c_void_p(0)
    L.check(lib.mpa_cluster_observables(N, lmax, S, _ptr(p_inc), _ptr(p_src), _ptr(mie_scat),
                                        _ptr(mie_int) if want_int else nullp, float(k), float(np.real(eps_b)),
                                        _ptr(p_scat), _ptr(p_int) if want_int else nullp, _ptr(C_part), _ptr(C_total),
                                        _ptr(F), _ptr(T) if want_torque else nullp, _stream()),
            "cluster_observables")
    return dict(p_scat=p_scat, p_int=p_int, C_part=C_part, C_total=C_total, F=F, T=T)


# mpa_obs_cluster as a numpy record (pointers as 64-bit integers), so that the
# descriptors of many clusters are filled column by column
_OBS_DTYPE = np.dtype([("N", "<i4"), ("lmax", "<i4"), ("nsrc", "<i4"), ("pad", "<i4"), ("p_inc", "<u8"),
                       ("p_src", "<u8"), ("mie_scat", "<u8"), ("mie_int", "<u8"), ("k", "<f8"), ("eps_b", "<f8"),
                       ("p_scat", "<u8"), ("p_int", "<u8"), ("C_part", "<u8"), ("C_total", "<u8"), ("F", "<u8"),
                       ("T", "<u8")])
assert _OBS_DTYPE.itemsize == ctypes.sizeof(L.MpaObsCluster)


class _ObsOutputs:
    """The flat output arrays of one batched observables call and, per output
    kind, each cluster's offset and shape in them."""

    def __init__(self, flats, offsets, shapes):
        self.flats, self.offsets, self.shapes = flats, offsets, shapes

    def view(self, key, q):
        flat = self.flats[key]
        if flat is None:
            return None
        shape = self.shapes[key][q]
        off = int(self.offsets[key][q])
        return flat[off:off + int(np.prod(shape))].view(shape)


class ObsResult(MutableMapping):
    """One cluster's observables (p_scat, p_int, C_part, C_total, F, T): views
    into the batch's flat outputs, each made on its first access; other
    entries can be set as in a dict."""

    __slots__ = ("_out", "_q", "_d", "_lazy")
    KEYS = ("p_scat", "p_int", "C_part", "C_total", "F", "T")

    def __init__(self, out, q):
        self._out, self._q, self._d, self._lazy = out, q, {}, {}

    def defer(self, key, tensor, shape):
        """An entry that is tensor.view(shape), made on its first access."""
        self._lazy[key] = (tensor, shape)

    def __getitem__(self, key):
        if key in self._d:
            return self._d[key]
        if key in self._lazy:
            t, shape = self._lazy.pop(key)
            v = self._d[key] = t.view(shape)
            return v
        if key not in self.KEYS:
            raise KeyError(key)
        v = self._d[key] = self._out.view(key, self._q)
        return v

    def __setitem__(self, key, value):
        self._lazy.pop(key, None)
        self._d[key] = value

    def __delitem__(self, key):
        raise TypeError("observables results are not deleted")

    def _extra(self):
        return [k for k in list(self._d) + list(self._lazy) if k not in self.KEYS]

    def __iter__(self):
        yield from self.KEYS
        yield from dict.fromkeys(self._extra())

    def __len__(self):
        return len(self.KEYS) + len(dict.fromkeys(self._extra()))


def observables_batched(clusters, want_int=True, want_torque=True, cache=None):
    """observables() for many clusters in a few launches
    (mpa_cluster_observables_batched); each cluster's results equal its own
    observables() call bit for bit.  clusters: [(p_inc, p_src, mie_scat,
    mie_int, k, eps_b)] with p_src [S, N, 2, rmax] c128 and p_inc the same
    S x N x 2 x rmax values (any shape, contiguous).  Each output kind is one
    flat allocation; returns one mapping per cluster whose entries are views
    into them (the descriptors are filled column by column, so the host work
    per cluster is a few pointer reads).  cache: a dict kept by a caller that
    passes the same clusters' inputs (p_src, Mie coefficients, k, eps_b) again
    with new p_inc: their descriptor columns are reused."""
    lib = L.load()
    if not clusters:
        return []
    dev = clusters[0][0].device
    q = len(clusters)
    st = cache.get("static") if cache is not None else None
    if st is None or st["q"] != q or st["want"] != (want_int, want_torque):
        shp = [c[1].shape for c in clusters]
        S = np.array([t[0] for t in shp], dtype=np.int64)
        N = np.array([t[1] for t in shp], dtype=np.int64)
        R = np.array([t[3] for t in shp], dtype=np.int64)
        Lm = np.array([c[2].shape[-1] for c in clusters], dtype=np.int64)
        counts = {"p": S * N * 2 * R, "C_part": S * N * 6 * Lm, "C_total": 3 * S, "F": S * N * 3}
        offs = {k: np.concatenate(([0], np.cumsum(v)[:-1])) for k, v in counts.items()}
        d0 = np.zeros(q, dtype=_OBS_DTYPE)
        d0["N"], d0["lmax"], d0["nsrc"] = N, Lm, S
        d0["p_src"] = [c[1].data_ptr() for c in clusters]
        d0["mie_scat"] = [c[2].data_ptr() for c in clusters]
        if want_int:
            d0["mie_int"] = [c[3].data_ptr() for c in clusters]
        d0["k"] = [float(c[4]) for c in clusters]
        d0["eps_b"] = [float(np.real(c[5])) for c in clusters]
        S_, N_, R_, L_ = S.tolist(), N.tolist(), R.tolist(), Lm.tolist()
        shapes = {"p": [(a, b, 2, c) for a, b, c in zip(S_, N_, R_)],
                  "C_part": [(a, b, 3, 2, c) for a, b, c in zip(S_, N_, L_)],
                  "C_total": [(a, 3) for a in S_], "F": [(a, b, 3) for a, b in zip(S_, N_)]}
        st = {"q": q, "want": (want_int, want_torque), "counts": {k: int(v.sum()) for k, v in counts.items()},
              "offs": offs, "d0": d0, "shapes": shapes}
        if cache is not None:
            cache["static"] = st
    counts, offs, shapes = st["counts"], st["offs"], st["shapes"]
    p_scat_all = torch.empty(counts["p"], dtype=C128, device=dev)
    p_int_all = torch.empty(counts["p"], dtype=C128, device=dev) if want_int else None
    C_part_all = torch.empty(counts["C_part"], dtype=torch.float64, device=dev)
    C_total_all = torch.empty(counts["C_total"], dtype=torch.float64, device=dev)
    F_all = torch.empty(counts["F"], dtype=torch.float64, device=dev)
    T_all = torch.empty(counts["F"], dtype=torch.float64, device=dev) if want_torque else None
    d = st["d0"].copy()
    d["p_inc"] = [c[0].data_ptr() for c in clusters]
    d["p_scat"] = p_scat_all.data_ptr() + 16 * offs["p"]
    if want_int:
        d["p_int"] = p_int_all.data_ptr() + 16 * offs["p"]
    d["C_part"] = C_part_all.data_ptr() + 8 * offs["C_part"]
    d["C_total"] = C_total_all.data_ptr() + 8 * offs["C_total"]
    d["F"] = F_all.data_ptr() + 8 * offs["F"]
    if want_torque:
        d["T"] = T_all.data_ptr() + 8 * offs["F"]
    L.check(lib.mpa_cluster_observables_batched(q, d.ctypes.data_as(ctypes.POINTER(L.MpaObsCluster)), _stream()),
            "cluster_observables_batched")
    out = _ObsOutputs({"p_scat": p_scat_all, "p_int": p_int_all, "C_part": C_part_all, "C_total": C_total_all,
                       "F": F_all, "T": T_all},
                      {"p_scat": offs["p"], "p_int": offs["p"], "C_part": offs["C_part"],
                       "C_total": offs["C_total"], "F": offs["F"], "T": offs["F"]},
                      {"p_scat": shapes["p"], "p_int": shapes["p"], "C_part": shapes["C_part"],
                       "C_total": shapes["C_total"], "F": shapes["F"], "T": shapes["F"]})
    return [ObsResult(out, i) for i in range(q)]


def translation_table(lmax):
    """Host copy of the device translation table: (meta [rmax^2, 4], coef [rmax^2, 2 lmax + 1])."""
    lib = L.load()
    rmax = lmax_to_rmax(lmax)
    meta = np.zeros((rmax * rmax, 4), dtype=np.int32)
    coef = np.zeros((rmax * rmax, 2 * lmax + 1), dtype=np.complex128)
    L.check(lib.mpa_translation_table(lmax, meta.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                      coef.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "translation_table")
    return meta, coef


class RefineReport:
    def __init__(self, outer, inner, resid, status, fp64_fallback=False):
        self.outer = outer
        self.iterations = inner
        self.residual = resid
        self.status = status
        self.fp64_fallback = fp64_fallback  # True: refinement failed and the system was re-solved in FP64

    @property
    def converged(self):
        return all(s == L.MPA_OK for s in self.status)

    def __repr__(self):
        return (f"RefineReport(outer={self.outer}, inner_iterations={self.iterations}, residual={self.residual}, "
                f"status={self.status}" + (", fp64_fallback=True)" if self.fp64_fallback else ")"))


def solve_refined(A32, A64, b, tol=1e-10, inner_tol=1e-5, maxiter=1000, max_refine=8, add_identity=False,
                  workspace=None, x=None, fallback=True):
    """Mixed-precision solve: c64 BiCGSTAB corrections, FP64 refinement.

    A32 [n, n] c64, A64 [n, n] c128 (device), b [S, n] c128.  Returns (x c128, RefineReport).
    fallback: a system whose refinement does not converge is re-solved with
    FP64 BiCGSTAB on A64 (RefineReport.fp64_fallback)."""
    lib = L.load()
    n = A32.shape[0]
    b2 = b.reshape(-1, n).contiguous()
    S = b2.shape[0]
    if A32.dtype != C64 or A64.dtype != C128 or b2.dtype != C128:
        raise ValueError("solve_refined needs A32 complex64, A64 complex128, b complex128")
    nbytes = lib.mpa_refine_workspace_bytes(n, S)
    if workspace is None or workspace.numel() < nbytes:
        workspace = torch.empty(nbytes, dtype=torch.uint8, device=A32.device)
    if x is None:
        x = torch.empty_like(b2)
    outer = (ctypes.c_int * S)()
    inner = (ctypes.c_int * S)()
    res = (ctypes.c_double * S)()
    st = (ctypes.c_int * S)()
    ai = 1 if add_identity else 0
    rc = lib.mpa_solve_refined(n, S, _ptr(A32), ai, _ptr(A64), ai, _ptr(b2), _ptr(x), float(tol), float(inner_tol),
                               int(maxiter), int(max_refine), _ptr(workspace), outer, inner, res, st, _stream())
    if rc not in (L.MPA_OK, L.MPA_NOT_CONVERGED, L.MPA_NONFINITE):
        L.check(rc, "solve_refined")
    rep = RefineReport(list(outer), list(inner), list(res), list(st))
    if fallback and any(v != L.MPA_OK for v in rep.status):
        x, rep = _fp64_fallback(A64, b2, x, tol, maxiter, add_identity, rep)
    return x.reshape(b.shape), rep


def _fp64_fallback(A64, b2, x, tol, maxiter, add_identity, rep):
    """A refinement that did not reach tol (its complex64 corrections failed,
    e.g. on an ill-conditioned system) is redone with FP64 BiCGSTAB on the
    same FP64 operator, on the GPU."""
    x64, r64 = bicgstab(A64, b2, tol=tol, maxiter=max(maxiter, 10 * b2.shape[1]), add_identity=add_identity,
                        check=False)
    x.copy_(x64.reshape(x.shape))
    return x, RefineReport(rep.outer, [a + b for a, b in zip(rep.iterations, r64.iterations)], r64.residual,
                           r64.status, fp64_fallback=True)


class ClusterOperator:
    """Matrix-free (I +) T of a sphere cluster (mpa_cluster_operator).

    Holds the device positions [N, 3], Mie coefficients [N, 2, lmax] and the
    workspace for up to `max_rhs` right-hand sides; T itself is never
    stored, so memory is O(n)."""

    def __init__(self, positions, mie, k, max_rhs=1, tmatrix=None, workspace=None):
        """`workspace` (optional): a contiguous uint8 device tensor of at least
        mpa_matfree_workspace_bytes(N, lmax, max_rhs) bytes to use instead of
        allocating one (e.g. a block of the serving engine's arena)."""
        lib = L.load()
        self.pos = to_device(positions, torch.float64).reshape(-1, 3)
        self.N = self.pos.shape[0]
        if (mie is None) == (tmatrix is None):
            raise ValueError("give exactly one of mie (spheres) or tmatrix (general particles)")
        if mie is not None:
            self.mie = to_device(mie, C128).reshape(self.N, 2, -1)
            self.lmax = self.mie.shape[-1]
            self.tmatrix = None
        else:
            self.mie = None
            self.tmatrix = to_device(tmatrix, C128).reshape(self.N, -1)
            bs = int(round(self.tmatrix.shape[1] ** 0.5))
            self.lmax = int(round((bs // 2 + 1) ** 0.5)) - 1
            self.tmatrix = self.tmatrix.reshape(self.N, bs, bs).contiguous()
        self.k = float(k)
        self.n = 2 * lmax_to_rmax(self.lmax) * self.N
        self.shape = (self.n, self.n)
        self.dtype = C128
        self.device = self.pos.device
        self.max_rhs = int(max_rhs)
        nb = lib.mpa_matfree_workspace_bytes(self.N, self.lmax, self.max_rhs)
        if nb == 0:
            raise ValueError(f"matrix-free operator: unsupported N={self.N}, lmax={self.lmax}")
        if workspace is not None:
            if workspace.dtype != torch.uint8 or not workspace.is_contiguous() or workspace.numel() < nb:
                raise ValueError(f"matrix-free workspace needs {nb} contiguous bytes")
            self.workspace = workspace
        else:
            self.workspace = torch.empty(nb, dtype=torch.uint8, device=self.device)
        self.struct = L.MpaClusterOperator(self.N, self.lmax, self.pos.data_ptr(),
                                           None if self.mie is None else self.mie.data_ptr(), self.k,
                                           self.workspace.data_ptr(),
                                           None if self.tmatrix is None else self.tmatrix.data_ptr())

    def reserve(self, nrhs):
        if nrhs > self.max_rhs:
            self.__init__(self.pos, self.mie, self.k, nrhs, self.tmatrix)

    @staticmethod
    def workspace_bytes(N, lmax, nrhs):
        return int(L.load().mpa_matfree_workspace_bytes(int(N), int(lmax), int(nrhs)))


def matfree_matvec(op, x, add_identity=False, out=None):
    """y = (I +) T x with T recomputed inside the product; x [nrhs, n]
    complex128 (FP64 operator) or complex64 (FP32 matrix-core operator)."""
    lib = L.load()
    x2 = x.reshape(-1, op.n)
    if x2.dtype not in (C128, C64) or not x2.is_contiguous():
        raise ValueError("x must be contiguous complex128 or complex64 [nrhs, n]")
    op.reserve(x2.shape[0])
    if out is None:
        out = torch.empty_like(x2)
    fn = lib.mpa_matfree_matvec if x2.dtype == C128 else lib.mpa_matfree_matvec_c64
    L.check(fn(ctypes.byref(op.struct), x2.shape[0], 1 if add_identity else 0, _ptr(x2), _ptr(out), _stream()),
            "matfree_matvec")
    return out.reshape(x.shape)


def matfree_matvec_batched(ops, xs, add_identity=False, outs=None):
    """y_q = (I +) T_q x_q for every operator q in one batched call
    (mpa_matfree_matvec_batched); xs [nrhs_q, n_q], all complex128 or all
    complex64.  Returns the list of outputs."""
    lib = L.load()
    ops, xs = list(ops), [x.reshape(-1, op.n) for op, x in zip(ops, xs)]
    if len(ops) != len(xs) or not ops:
        raise ValueError("one right-hand-side block per operator")
    dt = xs[0].dtype
    if dt not in (C128, C64) or any(x.dtype != dt or not x.is_contiguous() for x in xs):
        raise ValueError("xs must be contiguous, all complex128 or all complex64")
    for op, x in zip(ops, xs):
        op.reserve(x.shape[0])
    outs = [torch.empty_like(x) for x in xs] if outs is None else [o.reshape(x.shape) for o, x in zip(outs, xs)]
    cnt = len(ops)
    structs = [op.struct for op in ops]
    optr = (ctypes.c_void_p * cnt)(*[ctypes.cast(ctypes.pointer(s), ctypes.c_void_p).value for s in structs])
    L.check(lib.mpa_matfree_matvec_batched(cnt, optr, (ctypes.c_int * cnt)(*[x.shape[0] for x in xs]), DTYPE_CODE[dt],
                                           1 if add_identity else 0,
                                           (ctypes.c_void_p * cnt)(*[x.data_ptr() for x in xs]),
                                           (ctypes.c_void_p * cnt)(*[o.data_ptr() for o in outs]), _stream()),
            "matfree_matvec_batched")
    return outs


def autotune_matfree(op, nrhs=1, dtype=C64, reps=3):
    """Time the matrix-core contraction variants of the matrix-free operator
    on `op` and keep the fastest for dtype.  Returns (variant, ms)."""
    lib = L.load()
    op.reserve(nrhs)
    g = torch.Generator(device=op.device).manual_seed(0)
    x = torch.randn((nrhs, op.n), dtype=dtype, device=op.device, generator=g)
    y = torch.empty_like(x)
    v = ctypes.c_int(0)
    ms = ctypes.c_double(0)
    L.check(lib.mpa_autotune_matfree(ctypes.byref(op.struct), nrhs, DTYPE_CODE[dtype], _ptr(x), _ptr(y), int(reps),
                                     ctypes.byref(v), ctypes.byref(ms), _stream()), "autotune_matfree")
    return v.value, ms.value


def solve_refined_matfree(op, b, tol=1e-10, inner_tol=1e-5, maxiter=1000, max_refine=8, fallback=True):
    """Mixed precision without a stored matrix: complex64 corrections with the
    FP16-split matrix-core operator, FP64 residuals with the FP64 operator
    (FP64 matrix-free BiCGSTAB if the refinement does not converge)."""
    xs, reps = solve_refined_batched([op], [None], [b.reshape(-1, op.n)], tol=tol, inner_tol=inner_tol,
                                     maxiter=maxiter, max_refine=max_refine, add_identity=True, fallback=fallback)
    return xs[0].reshape(b.shape), reps[0]


def bicgstab_matfree(op, b, tol=1e-5, maxiter=1000, add_identity=True, x=None, workspace=None, sync_every=1):
    """BiCGSTAB on (I +) T x = b with the matrix-free operator (FP64)."""
    return bicgstab(op, b, tol=tol, maxiter=maxiter, add_identity=add_identity, x=x, workspace=workspace,
                    sync_every=sync_every)


def _bicgstab_matfree(lib, op, b, b2, tol, maxiter, add_identity, x, workspace, sync_every, check):
    S, n = b2.shape
    if b2.dtype != C128:
        raise ValueError("the matrix-free operator solves complex128 systems")
    op.reserve(S)
    nbytes = lib.mpa_bicgstab_workspace_bytes(n, S, L.MPA_C128)
    if workspace is None or workspace.numel() < nbytes:
        workspace = torch.empty(nbytes, dtype=torch.uint8, device=op.device)
    if x is None:
        x = torch.empty_like(b2)
    it = (ctypes.c_int * S)()
    res = (ctypes.c_double * S)()
    st = (ctypes.c_int * S)()
    rc = lib.mpa_bicgstab_matfree(ctypes.byref(op.struct), S, 1 if add_identity else 0, _ptr(b2), _ptr(x),
                                  float(tol), int(maxiter), _ptr(workspace), it, res, st, int(sync_every), _stream())
    report = SolveReport(list(it), list(res), list(st))
    if rc not in (L.MPA_OK, L.MPA_NOT_CONVERGED, L.MPA_NONFINITE) or (check and rc == L.MPA_NONFINITE):
        L.check(rc, "bicgstab_matfree")
    return x.reshape(b.shape), report


def _sys_struct(A, b, x, state, tol, maxiter, add_identity):
    n = A.shape[0]
    s = L.MpaSystem()
    s.n = n
    s.nrhs = b.numel() // n
    s.add_identity = 1 if add_identity else 0
    if isinstance(A, ClusterOperator):
        A.reserve(s.nrhs)
        s.add_identity = 1  # a ClusterOperator stands for the cluster's T: its system is always I + T
        s.A = None
        s.op = ctypes.cast(ctypes.pointer(A.struct), ctypes.c_void_p)
        s.b = b.data_ptr()
        s.x = x.data_ptr()
        s.state = state.data_ptr()
        s.tol = float(tol)
        s.maxiter = int(maxiter)
        return s
    s.A = A.data_ptr()
    s.b = b.data_ptr()
    s.x = x.data_ptr()
    s.state = state.data_ptr()
    s.tol = float(tol)
    s.maxiter = int(maxiter)
    return s


def bicgstab_batched(matrices, rhs, tol=1e-5, maxiter=1000, add_identity=False, sync_every=1,
                     preconditioners=None, ref_loop=False):
    """Solve several independent systems in one batched BiCGSTAB.

    matrices: list of [n_i, n_i] tensors (one dtype) or ClusterOperator
    (matrix-free, complex128); rhs: list of [S_i, n_i].
    tol / maxiter: scalars or per-system lists.  preconditioners: None or a
    per-system list of None / (M_inv_blocks [n_i, bs] device tensor, bs), the
    right preconditioner of `bicgstab`; ref_loop: scalar or per-system list
    (the reference's own loop, as in `bicgstab`).
    Returns (list of x, list of SolveReport)."""
    lib = L.load()
    k = len(matrices)
    dtype = matrices[0].dtype
    code = DTYPE_CODE[dtype]
    tols = tol if isinstance(tol, (list, tuple)) else [tol] * k
    maxits = maxiter if isinstance(maxiter, (list, tuple)) else [maxiter] * k
    pcs = preconditioners if preconditioners is not None else [None] * k
    refs = ref_loop if isinstance(ref_loop, (list, tuple)) else [ref_loop] * k
    xs, states, structs, keep = [], [], (L.MpaSystem * k)(), []
    for i, (A, b) in enumerate(zip(matrices, rhs)):
        n = A.shape[0]
        b2 = b.reshape(-1, n)
        if A.dtype != dtype or b2.dtype != dtype or not b2.is_contiguous():
            raise ValueError("all matrices and right-hand sides need one dtype and contiguous storage")
        x = torch.empty_like(b2)
        st = torch.empty(lib.mpa_system_state_bytes_ex(n, b2.shape[0], code, 1 if pcs[i] is not None else 0),
                         dtype=torch.uint8, device=A.device)
        xs.append(x)
        states.append(st)
        structs[i] = _sys_struct(A, b2, x, st, tols[i], maxits[i], add_identity)
        structs[i].flags = L.MPA_SOLVE_REF_LOOP if refs[i] else 0
        if pcs[i] is not None:
            M, bs = pcs[i]
            if M.dtype != dtype or not M.is_contiguous() or M.numel() != n * bs:
                raise ValueError("each preconditioner is a contiguous [n, bs] tensor of the matrices' dtype")
            keep.append(M)
            structs[i].minv = M.data_ptr()
            structs[i].block_size = int(bs)
    scratch = torch.empty(lib.mpa_batch_scratch_bytes(k, structs, code), dtype=torch.uint8, device=matrices[0].device)
    total = sum(x.shape[0] for x in xs)
    it = (ctypes.c_int * total)()
    res = (ctypes.c_double * total)()
    stt = (ctypes.c_int * total)()
    rc = lib.mpa_bicgstab_batched(k, structs, code, _ptr(scratch), it, res, stt, int(sync_every), _stream())
    if rc not in (L.MPA_OK, L.MPA_NOT_CONVERGED, L.MPA_NONFINITE):
        L.check(rc, "bicgstab_batched")
    it_l, res_l, stt_l = list(it), list(res), list(stt)  # once, then a slice per system
    reports, o = [], 0
    for x in xs:
        S = x.shape[0]
        reports.append(SolveReport(it_l[o:o + S], res_l[o:o + S], stt_l[o:o + S]))
        o += S
    return xs, reports


def solve_refined_batched(A32s, A64s, rhs, tol=1e-10, inner_tol=1e-5, maxiter=1000, max_refine=8,
                          add_identity=False, fallback=True):
    """Mixed-precision solve of several systems at once (c64 corrections, FP64 refinement).

    A32s: list of [n_i, n_i] c64, A64s: list of [n_i, n_i] c128, rhs: list of [S_i, n_i] c128.
    An entry of A32s may be a ClusterOperator (its A64s entry None): that
    system (always I + T, whatever add_identity says for the stored ones)
    uses the complex64 matrix-core operator for its corrections and the FP64
    one for its residual, with no stored matrix."""
    lib = L.load()
    k = len(A32s)
    tols = tol if isinstance(tol, (list, tuple)) else [tol] * k
    structs = (L.MpaRefineSystem * k)()
    xs, states = [], []
    for i, (A32, A64, b) in enumerate(zip(A32s, A64s, rhs)):
        n = A32.shape[0]
        b2 = b.reshape(-1, n)
        op = A32 if isinstance(A32, ClusterOperator) else None
        if op is None and (A32.dtype != C64 or A64.dtype != C128):
            raise ValueError("solve_refined_batched needs c64 / c128 matrices")
        if b2.dtype != C128 or not b2.is_contiguous():
            raise ValueError("solve_refined_batched needs contiguous c128 right-hand sides")
        if op is not None:
            op.reserve(b2.shape[0])
        x = torch.empty_like(b2)
        st = torch.empty(lib.mpa_refine_state_bytes(n, b2.shape[0]), dtype=torch.uint8, device=b2.device)
        xs.append(x)
        states.append(st)
        s = L.MpaRefineSystem()
        # a ClusterOperator stands for the cluster's T: its system is always I + T
        s.n, s.nrhs, s.add_identity = n, b2.shape[0], 1 if (add_identity or op is not None) else 0
        if op is not None:
            s.A32 = s.A64 = None
            s.op = ctypes.cast(ctypes.pointer(op.struct), ctypes.c_void_p)
        else:
            s.A32, s.A64 = A32.data_ptr(), A64.data_ptr()
        s.b, s.x, s.state = b2.data_ptr(), x.data_ptr(), st.data_ptr()
        s.tol, s.inner_tol, s.maxiter, s.max_refine = float(tols[i]), float(inner_tol), int(maxiter), int(max_refine)
        structs[i] = s
    scratch = torch.empty(lib.mpa_refine_scratch_bytes(k, structs), dtype=torch.uint8, device=rhs[0].device)
    total = sum(x.shape[0] for x in xs)
    outer = (ctypes.c_int * total)()
    inner = (ctypes.c_int * total)()
    res = (ctypes.c_double * total)()
    stt = (ctypes.c_int * total)()
    rc = lib.mpa_solve_refined_batched(k, structs, _ptr(scratch), outer, inner, res, stt, _stream())
    if rc not in (L.MPA_OK, L.MPA_NOT_CONVERGED, L.MPA_NONFINITE):
        L.check(rc, "solve_refined_batched")
    # the per-RHS records as Python lists once, then a slice per system
    outer_l, inner_l, res_l, stt_l = list(outer), list(inner), list(res), list(stt)
    any_bad = fallback and any(v != L.MPA_OK for v in stt_l)
    reports, o = [], 0
    for i, x in enumerate(xs):
        S = x.shape[0]
        rep = RefineReport(outer_l[o:o + S], inner_l[o:o + S], res_l[o:o + S], stt_l[o:o + S])
        if any_bad and any(v != L.MPA_OK for v in rep.status):
            # the operator that stands for the FP64 system: the cluster operator (always I + T) or A64
            A = A32s[i] if isinstance(A32s[i], ClusterOperator) else A64s[i]
            ai = True if isinstance(A32s[i], ClusterOperator) else add_identity
            x, rep = _fp64_fallback(A, rhs[i].reshape(-1, x.shape[1]), x, tols[i], maxiter, ai, rep)
            xs[i] = x
        reports.append(rep)
        o += S
    return xs, reports
