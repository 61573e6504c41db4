"""The MI355X backend: the reference's dispatch functions, same signatures.

Each function here is what `miepy.interactions` / `flux` / `forces` would
call when the 'mi355x' backend is selected (the slots the JAX 'gpu' backend
takes in src/miepy/interactions.py:81-83,141-143 and src/miepy/flux.py:21-26):

  sphere_aggregate_tmatrix(positions, mie, k)       interactions.py:132-150
  solve_linear_system(tmatrix, p_src, method, M_inv_blocks, block_size)
                                                    interactions.py:71-92
  build_block_preconditioner_sphere / _particle     interactions.py:9-68
  bicgstab, bicgstab_profiled, bicgstab_preconditioned_profiled
                                                    cpp/src/interactions_pyb.cpp:39-69
  particle_cross_sections(p_scat, p_inc, p_src, k)  flux.py:12-32
  force_all / torque_all(p_scat, p_inc, k, eps_b, mu_b)  forces.py:44-81
  cluster_coefficients(positions, p_scat, k, origin, lmax)  vsh/cluster_coefficients.py:14-53
  cluster_cross_sections(p_cluster, p_src, k)       flux.py:35-56
  vsh_translation, vsh_translation_numpy, vsh_mode  cpp/src/vsh_translation_pyb.cpp:14-31

Like the JAX backend, the assembled matrix stays in HBM: it is returned as a
`DeviceTMatrix`, which has the reference's [N,2,rmax,N,2,rmax] shape,
supports `.reshape` and converts to numpy on demand (`np.asarray(T)`), and is
consumed in place by `solve_linear_system`.
"""

import enum

import numpy as np
import torch

from . import device as D


class solver(enum.Enum):
    """Mirror of the pybind enum miepy.solver (cpp/src/interactions_pyb.cpp:33-37)."""

    bicgstab = 0
    exact = 1


# Solver knobs the reference hard-codes (cpp/src/interactions.cpp:362-363).
options = dict(tol=1e-5, maxiter=1000)


class DeviceTMatrix:
    """Aggregate T-matrix resident in HBM (torch tensor [n, n])."""

    def __init__(self, data, shape):
        self.data = data
        self._shape = tuple(shape)

    @property
    def shape(self):
        return self._shape

    @property
    def dtype(self):
        return np.complex128 if self.data.dtype == torch.complex128 else np.complex64

    def reshape(self, *shape):
        if len(shape) == 1 and isinstance(shape[0], (tuple, list)):
            shape = tuple(shape[0])
        if int(np.prod(shape)) != self.data.numel():
            raise ValueError(f"cannot reshape {self._shape} to {shape}")
        return DeviceTMatrix(self.data, shape)

    def __array__(self, dtype=None, copy=None):
        a = self.data.cpu().numpy().reshape(self._shape)
        return a.astype(dtype) if dtype is not None else a


def sphere_aggregate_tmatrix(positions, mie, k):
    """positions[N,3], mie[N,2,lmax], k -> T[N,2,rmax,N,2,rmax] (device-resident)."""
    positions = np.asarray(positions, dtype=float).reshape(-1, 3)
    N = positions.shape[0]
    lmax = np.shape(mie)[-1]
    rmax = lmax * (lmax + 2)
    T = D.assemble(positions, np.asarray(mie, dtype=complex).reshape(N, 2, lmax), float(np.real(k)))
    return DeviceTMatrix(T, (N, 2, rmax, N, 2, rmax))


def particle_aggregate_tmatrix(positions, tmatrix, k):
    """positions[N,3], tmatrix[N,2,rmax,2,rmax], k -> T[N,2,rmax,N,2,rmax]
    (device-resident).  src/miepy/interactions.py:190-205."""
    positions = np.asarray(positions, dtype=float).reshape(-1, 3)
    N = positions.shape[0]
    rmax = np.shape(tmatrix)[-1]
    T = D.assemble_tmatrix(positions, np.asarray(tmatrix, dtype=complex).reshape(N, 2, rmax, 2, rmax),
                           float(np.real(k)))
    return DeviceTMatrix(T, (N, 2, rmax, N, 2, rmax))


def _device_matrix(tmatrix, n):
    if isinstance(tmatrix, DeviceTMatrix):
        return tmatrix.data.reshape(n, n)
    return D.to_device(np.asarray(tmatrix, dtype=complex).reshape(n, n), D.C128)


def _device_blocks(M_inv_blocks, n, block_size, dtype):
    M = M_inv_blocks.data if isinstance(M_inv_blocks, DeviceTMatrix) else M_inv_blocks
    if isinstance(M, torch.Tensor):
        M = M.to(dtype)
    else:
        M = D.to_device(np.asarray(M, dtype=complex), dtype)
    if block_size < 1 or n % block_size or M.numel() != n * block_size:
        raise ValueError(f"M_inv_blocks must be [{n}, block_size] with block_size dividing {n}")
    return M.reshape(n, block_size).contiguous()


def solve_linear_system(tmatrix, p_src, method=solver.bicgstab, M_inv_blocks=None, block_size=0, tol=None,
                        maxiter=None):
    """Solve (I + T) p_inc = p_src.  p_src[N,2,rmax] -> p_inc[N,2,rmax] (host numpy).

    method=bicgstab runs BiCGSTAB (default tolerance 1e-5 as the reference),
    right-preconditioned by the block-diagonal M_inv_blocks [n, block_size]
    when it is given (solve_linear_system_preconditioned,
    cpp/src/interactions.cpp:291-345); method=exact factors I + T by LU with
    partial pivoting on the device (interactions.cpp:356-357) and, like the
    reference, ignores the preconditioner."""
    p_src = np.asarray(p_src)
    shape = p_src.shape
    n = p_src.size
    A = _device_matrix(tmatrix, n)
    exact = (method == solver.exact) or (getattr(method, "name", None) == "exact")
    if exact:
        A64 = A if A.dtype == D.C128 else A.to(D.C128)
        x = D.lu_solve(A64, D.to_device(p_src.reshape(1, n), D.C128), add_identity=True)
        solve_linear_system.last_report = None
        return x.cpu().numpy().reshape(shape)
    t = options["tol"] if tol is None else tol
    mi = options["maxiter"] if maxiter is None else maxiter
    b = D.to_device(p_src.reshape(1, n), A.dtype)
    M = None if M_inv_blocks is None else _device_blocks(M_inv_blocks, n, int(block_size), A.dtype)
    x, report = D.bicgstab(A, b, tol=t, maxiter=mi, add_identity=True, M_inv_blocks=M,
                           block_size=int(block_size) if M is not None else 0)
    solve_linear_system.last_report = report
    return x.cpu().numpy().reshape(shape)


def build_block_preconditioner_sphere(mie_scat, lmax):
    """Block-diagonal M^-1 of a sphere cluster (src/miepy/interactions.py:9-40):
    each particle's block is diag(1 / (1 + mie_scat[i, pol, n-1])), 1 where
    |1 + mie_scat| < 1e-12.  Returns (M_inv_blocks [N*bs, bs], bs), bs = 2 rmax."""
    mie_scat = np.asarray(mie_scat, dtype=complex)
    N = mie_scat.shape[0]
    rmax = lmax * (lmax + 2)
    bs = 2 * rmax
    nidx = np.concatenate([np.full(2 * n + 1, n - 1) for n in range(1, lmax + 1)])
    d = 1.0 + mie_scat[:, :, nidx].reshape(N, bs)
    small = np.abs(d) < 1e-12
    inv = np.where(small, 1.0, 1.0 / np.where(small, 1.0, d))
    M = np.zeros((N, bs, bs), dtype=complex)
    M[:, np.arange(bs), np.arange(bs)] = inv
    return M.reshape(N * bs, bs), bs


def build_block_preconditioner_particle(tmatrix):
    """Block-diagonal M^-1 = inv(I + T_i) per particle (src/miepy/interactions.py:43-68);
    a singular block falls back to the identity.  Returns (M_inv_blocks [N*bs, bs], bs)."""
    tmatrix = np.asarray(tmatrix, dtype=complex)
    N = tmatrix.shape[0]
    bs = 2 * tmatrix.shape[-1]
    blocks = np.eye(bs) + tmatrix.reshape(N, bs, bs)
    M = np.empty_like(blocks)
    for i in range(N):
        try:
            M[i] = np.linalg.inv(blocks[i])
        except np.linalg.LinAlgError:
            M[i] = np.eye(bs)
    return M.reshape(N * bs, bs), bs


def _ref_loop(A, b, maxiter, tolerance, M_inv_blocks=None, block_size=0):
    """The reference's own BiCGSTAB loop on the device (x0 = b, absolute
    tolerance, at most maxiter + 1 iterations; cpp/src/interactions.cpp:98-209
    and, preconditioned, :230-288).  A is the full system matrix."""
    b = np.asarray(b, dtype=complex).reshape(-1)
    n = b.size
    Ad = A.data.reshape(n, n) if isinstance(A, DeviceTMatrix) else (
        A.reshape(n, n) if isinstance(A, torch.Tensor) else D.to_device(np.asarray(A, dtype=complex).reshape(n, n),
                                                                         D.C128))
    M = None if M_inv_blocks is None else _device_blocks(M_inv_blocks, n, int(block_size), Ad.dtype)
    x, rep = D.bicgstab(Ad, D.to_device(b.reshape(1, n), Ad.dtype), tol=float(tolerance), maxiter=int(maxiter),
                        add_identity=False, M_inv_blocks=M, block_size=int(block_size) if M is not None else 0,
                        ref_loop=True, check=False)
    residual = rep.residual[0] * float(np.linalg.norm(b))  # ||r|| (the engine reports ||r||/||b||)
    return x.cpu().numpy().reshape(n), int(rep.iterations[0]), residual


def bicgstab(A, b, maxiter, tolerance):
    """interactions.bicgstab(A, b, maxiter, tolerance) -> x
    (cpp/src/interactions_pyb.cpp:39-44, interactions.cpp:98-150)."""
    return _ref_loop(A, b, maxiter, tolerance)[0]


def bicgstab_profiled(A, b, maxiter=1000, tolerance=1e-5):
    """interactions.bicgstab_profiled -> (x, iterations, ||r||)
    (cpp/src/interactions_pyb.cpp:46-55, interactions.cpp:152-209)."""
    return _ref_loop(A, b, maxiter, tolerance)


def bicgstab_preconditioned_profiled(A, b, M_inv_blocks, block_size, maxiter=1000, tolerance=1e-5):
    """interactions.bicgstab_preconditioned_profiled -> (x, iterations, ||r||)
    (cpp/src/interactions_pyb.cpp:57-69, interactions.cpp:230-288)."""
    return _ref_loop(A, b, maxiter, tolerance, M_inv_blocks, block_size)


solve_linear_system.last_report = None


class vsh_mode(enum.Enum):
    """Mirror of the pybind enum miepy.vsh_mode (cpp/src/vsh_functions.hpp:10-15)."""

    outgoing = 0
    ingoing = 1
    incident = 2
    interior = 3


def _mode_value(mode):
    return mode.value if isinstance(mode, enum.Enum) else int(getattr(mode, "value", mode))


def vsh_translation(m, n, u, v, rad, theta, phi, k, mode):
    """miepy.cpp.vsh_translation.vsh_translation: the translation coefficients
    (A, B) of one (m, n, u, v) for one vector (cpp/src/vsh_translation.cpp:428-470),
    evaluated on the device."""
    out = D.vsh_translation([[m, n, u, v]], rad, theta, phi, float(np.real(k)), _mode_value(mode))
    return complex(out[0, 0, 0]), complex(out[0, 0, 1])


def vsh_translation_numpy(m, n, u, v, rad, theta, phi, k, mode):
    """miepy.cpp.vsh_translation.vsh_translation_numpy: one (m, n, u, v) over
    arrays of vectors -> [2, len] (row 0 A, row 1 B) (vsh_translation.cpp:366-426)."""
    out = D.vsh_translation([[m, n, u, v]], rad, theta, phi, float(np.real(k)), _mode_value(mode))
    return np.ascontiguousarray(out[:, 0, :].T)


def cluster_coefficients(positions, p_scat, k, origin, lmax=None):
    """miepy.cluster_coefficients (src/miepy/vsh/cluster_coefficients.py:14-53):
    p_scat[N,2,rmax] of the particles -> p_cluster[2,rmax(lmax)] about origin
    (host numpy), computed on the device."""
    p_scat = np.asarray(p_scat, dtype=complex)
    rmax_in = p_scat.shape[-1]
    lmax_in = int(round((1 + rmax_in) ** 0.5)) - 1
    lmax = lmax_in if lmax is None else int(lmax)
    pc = D.cluster_coefficients(positions, p_scat[None], float(np.real(k)), origin, lmax)
    return pc[0].cpu().numpy()


def cluster_cross_sections(p_cluster, p_src, k):
    """miepy.flux.cluster_cross_sections (src/miepy/flux.py:35-56,
    cpp/src/flux.cpp:32-57): (scattering, absorption, extinction), each [2, lmax]."""
    from .flux import cross_sections
    pc = np.asarray(p_cluster, dtype=complex).reshape(1, 2, -1)
    C = D.cluster_cross_sections(D.to_device(pc, D.C128), np.asarray(p_src, dtype=complex).reshape(pc.shape),
                                 float(np.real(k))).cpu().numpy()[0]
    return cross_sections(C[0], C[1], C[2])


def _rows(a):
    a = np.asarray(a, dtype=complex)
    return a.reshape(a.shape[0], -1) if a.ndim > 1 else a.reshape(1, -1)


def particle_cross_sections(p_scat, p_inc, p_src, k):
    """Per-multipole (scattering, absorption, extinction) of one particle, each [2, lmax]."""
    from .flux import cross_sections
    rows = [np.asarray(a, dtype=complex).reshape(1, -1) for a in (p_scat, p_inc, p_src)]
    lmax = int(-1 + (1 + rows[0].shape[1] // 2) ** 0.5)
    C = _xs_rows(*rows, k, lmax)[0]
    return cross_sections(C[0], C[1], C[2])


def _xs_rows(ps, pi, pr, k, lmax):
    import ctypes
    from . import _lib as L
    M = ps.shape[0]
    d = [D.to_device(a, D.C128) for a in (ps, pi, pr)]
    C = torch.empty((M, 3, 2, lmax), dtype=torch.float64, device=d[0].device)
    L.check(L.load().mpa_particle_cross_sections(M, lmax, *[ctypes.c_void_p(t.data_ptr()) for t in d], float(np.real(k)),
                                                 ctypes.c_void_p(C.data_ptr()), D._stream()),
            "particle_cross_sections")
    return C.cpu().numpy()


def _force_torque(p_scat, p_inc, k, eps_b, want_force):
    import ctypes
    from . import _lib as L
    ps, pi = _rows(p_scat), _rows(p_inc)
    M = ps.shape[0]
    lmax = int(-1 + (1 + ps.shape[1] // 2) ** 0.5)
    d = [D.to_device(a, D.C128) for a in (ps, pi)]
    out = torch.empty((M, 3), dtype=torch.float64, device=d[0].device)
    null = ctypes.c_void_p(0)
    ptr = ctypes.c_void_p(out.data_ptr())
    L.check(L.load().mpa_force_torque(M, lmax, ctypes.c_void_p(d[0].data_ptr()), ctypes.c_void_p(d[1].data_ptr()),
                                      float(np.real(k)), float(np.real(eps_b)), ptr if want_force else null,
                                      null if want_force else ptr, D._stream()), "force_torque")
    return out.cpu().numpy()


def force_all(p_scat, p_inc, k, eps_b, mu_b):
    """F[N,3] on every particle (mu_b unused, as cpp/src/forces.cpp:10)."""
    return _force_torque(p_scat, p_inc, k, eps_b, True)


def torque_all(p_scat, p_inc, k, eps_b, mu_b):
    """T[N,3] on every particle."""
    return _force_torque(p_scat, p_inc, k, eps_b, False)


def force(p_scat, p_inc, k, eps_b, mu_b):
    return force_all(np.asarray(p_scat).reshape(1, -1), np.asarray(p_inc).reshape(1, -1), k, eps_b, mu_b)[0]


def torque(p_scat, p_inc, k, eps_b, mu_b):
    return torque_all(np.asarray(p_scat).reshape(1, -1), np.asarray(p_inc).reshape(1, -1), k, eps_b, mu_b)[0]
