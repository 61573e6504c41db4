"""CPU FP64 oracle for the MiePy cluster solve (ctypes bindings).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py``, always as the checker.  The shipped
package ``miepy_amd`` never imports this module.

The native part (``oracle.cpp``, built by ``make -C oracle``) restates the
reference's C++ path (cpp/src/{special,vsh_translation,interactions,flux,
forces}.cpp) without GSL/Eigen; ``oracle.host`` restates the Python host steps
(Mie coefficients, plane-wave source decomposition, sphere_cluster.solve).

Parity status per function is recorded in the header of each test that pins
it; every exported function has at least one independent pin (scipy/sympy
golden vectors, closed forms, invariants, or the reference's own fixtures).
"""

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_d = ctypes.c_double
_i = ctypes.c_int
_pd = ctypes.POINTER(ctypes.c_double)
_pi = ctypes.POINTER(ctypes.c_int)

_SIGS = {
    "orc_spherical_hn_recursion": (None, [_i, _d, _pd]),
    "orc_spherical_jn_complex": (None, [_i, _d, _d, _pd]),
    "orc_associated_legendre_recursion": (None, [_i, _d, _pd]),
    "orc_associated_legendre": (_d, [_i, _i, _d]),
    "orc_pi_func": (_d, [_i, _i, _d]),
    "orc_tau_func": (_d, [_i, _i, _d]),
    "orc_Emn": (None, [_i, _i, _pd]),
    "orc_wigner_3j": (_d, [_i] * 6),
    "orc_wigner_3j_batch": (_i, [_i, _i, _i, _i, _i, _pi, _pd, _i]),
    "orc_a_func": (_d, [_i] * 5),
    "orc_b_func": (_d, [_i] * 5),
    "orc_gaunt_batch": (None, [_i, _i, _i, _i, _pd, _pd]),
    "orc_vsh_translation": (None, [_i, _i, _i, _i, _d, _d, _d, _d, _i, _pd]),
    "orc_sphere_aggregate_tmatrix": (None, [_pd, _pd, _i, _i, _d, _pd]),
    "orc_particle_aggregate_tmatrix": (None, [_pd, _pd, _i, _i, _d, _pd]),
    "orc_bicgstab_eigen": (_i, [_pd, _pd, _i, _d, _i, _pd, _pd]),
    "orc_bicgstab_custom": (_i, [_pd, _pd, _i, _i, _d, _pd, _pd]),
    "orc_bicgstab_eigen_pc": (_i, [_pd, _pd, _i, _pd, _i, _d, _i, _pd, _pd]),
    "orc_bicgstab_custom_pc": (_i, [_pd, _pd, _i, _pd, _i, _i, _d, _pd, _pd]),
    "orc_lu_solve": (_i, [_pd, _pd, _i, _pd]),
    "orc_solve_linear_system": (_i, [_pd, _pd, _i, _i, _d, _i, _pd, _pd]),
    "orc_gemv": (None, [_pd, _pd, _pd, _i]),
    "orc_particle_cross_sections": (None, [_pd, _pd, _pd, _i, _d, _pd, _pd, _pd]),
    "orc_cluster_cross_sections": (None, [_pd, _pd, _i, _d, _pd, _pd, _pd]),
    "orc_eps0": (_d, []),
    "orc_force": (None, [_pd, _pd, _i, _d, _d, _d, _pd]),
    "orc_torque": (None, [_pd, _pd, _i, _d, _d, _d, _pd]),
    "orc_force_all": (None, [_pd, _pd, _i, _i, _d, _d, _d, _pd]),
    "orc_torque_all": (None, [_pd, _pd, _i, _i, _d, _d, _d, _pd]),
    "orc_max_threads": (_i, []),
    "orc_set_threads": (None, [_i]),
}


def build(force=False):
    """Compile oracle.cpp into liboracle.so (g++ -fopenmp)."""
    src = os.path.join(_HERE, "oracle.cpp")
    if force or not os.path.exists(_LIB_PATH) or os.path.getmtime(_LIB_PATH) < os.path.getmtime(src):
        subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_pd)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def lmax_to_rmax(lmax):
    return lmax * (lmax + 2)


def rmax_to_lmax(rmax):
    return int(-1 + (1 + rmax) ** 0.5)


# ---- special functions -------------------------------------------------------

def spherical_hn_recursion(nmax, z):
    out = np.zeros(nmax + 1, dtype=np.complex128)
    lib().orc_spherical_hn_recursion(nmax, float(z), _p(out))
    return out


def spherical_jn_complex(n, z):
    out = np.zeros(2)
    lib().orc_spherical_jn_complex(n, float(np.real(z)), float(np.imag(z)), _p(out))
    return complex(out[0], out[1])


def associated_legendre_recursion(nmax, z):
    out = np.zeros(nmax * (nmax + 2) + 1)
    lib().orc_associated_legendre_recursion(nmax, float(z), _p(out))
    return out


def associated_legendre(n, m, z):
    return lib().orc_associated_legendre(n, m, float(z))


def pi_func(n, m, theta):
    return lib().orc_pi_func(n, m, float(theta))


def tau_func(n, m, theta):
    return lib().orc_tau_func(n, m, float(theta))


def Emn(m, n):
    out = np.zeros(2)
    lib().orc_Emn(m, n, _p(out))
    return complex(out[0], out[1])


def wigner_3j(j1, j2, j3, m1, m2, m3):
    return lib().orc_wigner_3j(j1, j2, j3, m1, m2, m3)


def wigner_3j_batch(j2, j3, m1, m2, m3):
    cap = 2 * (j2 + j3) + 4
    out = np.zeros(cap)
    jmin = ctypes.c_int(0)
    size = lib().orc_wigner_3j_batch(j2, j3, m1, m2, m3, ctypes.byref(jmin), _p(out), cap)
    return jmin.value, jmin.value + size - 1, out[:size].copy()


def a_func(m, n, u, v, p):
    return lib().orc_a_func(m, n, u, v, p)


def b_func(m, n, u, v, p):
    return lib().orc_b_func(m, n, u, v, p)


def gaunt_batch(m, n, u, v):
    qa = min(n, v, (n + v - abs(m + u)) // 2)
    qb = min(n, v, (n + v + 1 - abs(m + u)) // 2)
    a = np.zeros(max(qa + 1, 1))
    b = np.zeros(max(qb + 1, 1))
    lib().orc_gaunt_batch(m, n, u, v, _p(a), _p(b))
    return a[: qa + 1], b[: qb + 1]


VSH_MODE = {"outgoing": 0, "ingoing": 1, "incident": 2, "interior": 3}


def vsh_translation(m, n, u, v, rad, theta, phi, k, mode="outgoing"):
    out = np.zeros(4)
    lib().orc_vsh_translation(m, n, u, v, float(rad), float(theta), float(phi), float(k), VSH_MODE[mode], _p(out))
    return complex(out[0], out[1]), complex(out[2], out[3])


# ---- aggregate T-matrix and solve ------------------------------------------------

def sphere_aggregate_tmatrix(positions, mie, k, out=None):
    """positions[N,3], mie[N,2,lmax] -> T[n,n] complex128 (row-major).

    `out` may be a preallocated C-contiguous [n, n] complex128 array."""
    positions = _f64(positions).reshape(-1, 3)
    N = positions.shape[0]
    lmax = mie.shape[-1]
    rmax = lmax_to_rmax(lmax)
    n = 2 * rmax * N
    mie2 = _c128(mie).reshape(N, 2 * lmax)
    if out is None:
        out = np.empty((n, n), dtype=np.complex128)
    assert out.shape == (n, n) and out.dtype == np.complex128 and out.flags.c_contiguous
    lib().orc_sphere_aggregate_tmatrix(_p(positions), _p(mie2), N, lmax, float(k), _p(out))
    return out


def particle_aggregate_tmatrix(positions, tmatrix, k):
    """positions[N,3], tmatrix[N,2,rmax,2,rmax] -> T[n,n]."""
    positions = _f64(positions).reshape(-1, 3)
    N = positions.shape[0]
    rmax = tmatrix.shape[-1]
    n = 2 * rmax * N
    tm = _c128(tmatrix).reshape(N, 2 * rmax, 2 * rmax)
    out = np.empty((n, n), dtype=np.complex128)
    lib().orc_particle_aggregate_tmatrix(_p(positions), _p(tm), N, rmax, float(k), _p(out))
    return out


def bicgstab_eigen(A, b, tol=1e-5, maxiter=1000, M_inv_blocks=None, block_size=0):
    """Eigen-style BiCGSTAB on the full matrix A, optionally right-preconditioned
    by the block-diagonal M_inv_blocks [n, block_size].  Returns (x, iters, rel_err)."""
    A = _c128(A)
    b = _c128(b).reshape(-1)
    x = np.zeros_like(b)
    err = np.zeros(1)
    if M_inv_blocks is None:
        it = lib().orc_bicgstab_eigen(_p(A), _p(b), b.size, float(tol), int(maxiter), _p(x), _p(err))
    else:
        M = _c128(M_inv_blocks).reshape(b.size, block_size)
        it = lib().orc_bicgstab_eigen_pc(_p(A), _p(b), b.size, _p(M), int(block_size), float(tol), int(maxiter),
                                         _p(x), _p(err))
    return x, it, float(err[0])


def bicgstab(A, b, maxiter, tolerance, M_inv_blocks=None, block_size=0):
    """Custom BiCGSTAB (x0 = b, absolute tol), optionally right-preconditioned
    (cpp/src/interactions.cpp:230-289).  Returns (x, iters, residual)."""
    A = _c128(A)
    b = _c128(b).reshape(-1)
    x = np.zeros_like(b)
    res = np.zeros(1)
    if M_inv_blocks is None:
        it = lib().orc_bicgstab_custom(_p(A), _p(b), b.size, int(maxiter), float(tolerance), _p(x), _p(res))
    else:
        M = _c128(M_inv_blocks).reshape(b.size, block_size)
        it = lib().orc_bicgstab_custom_pc(_p(A), _p(b), b.size, _p(M), int(block_size), int(maxiter),
                                          float(tolerance), _p(x), _p(res))
    return x, it, float(res[0])


def build_block_preconditioner_sphere(mie_scat, lmax):
    """src/miepy/interactions.py:9-40: blocks diag(1 / (1 + mie_scat[i, pol, n-1])),
    identity where |1 + a| < 1e-12.  Returns (M_inv_blocks [N*bs, bs], bs)."""
    mie_scat = np.asarray(mie_scat, dtype=complex)
    N = mie_scat.shape[0]
    rmax = lmax_to_rmax(lmax)
    bs = 2 * rmax
    M = np.zeros((N * bs, bs), dtype=complex)
    for i in range(N):
        for pol in range(2):
            r = 0
            for n in range(1, lmax + 1):
                for _m in range(-n, n + 1):
                    idx = pol * rmax + r
                    d = 1.0 + mie_scat[i, pol, n - 1]
                    M[i * bs + idx, idx] = 1.0 if abs(d) < 1e-12 else 1.0 / d
                    r += 1
    return M, bs


def build_block_preconditioner_particle(tmatrix):
    """src/miepy/interactions.py:43-68: blocks inv(I + T_i), identity for a
    singular block.  Returns (M_inv_blocks [N*bs, bs], bs)."""
    tmatrix = np.asarray(tmatrix, dtype=complex)
    N = tmatrix.shape[0]
    bs = 2 * tmatrix.shape[-1]
    M = np.zeros((N * bs, bs), dtype=complex)
    for i in range(N):
        blk = np.eye(bs) + tmatrix[i].reshape(bs, bs)
        try:
            M[i * bs:(i + 1) * bs] = np.linalg.inv(blk)
        except np.linalg.LinAlgError:
            M[i * bs:(i + 1) * bs] = np.eye(bs)
    return M, bs


def lu_solve(A, b):
    A = _c128(A)
    b = _c128(b).reshape(-1)
    x = np.zeros_like(b)
    if lib().orc_lu_solve(_p(A), _p(b), b.size, _p(x)) != 0:
        raise np.linalg.LinAlgError("singular matrix")
    return x


def solve_linear_system(T, p_src, method="bicgstab", tol=1e-5, maxiter=1000):
    """(I + T) x = p_src.  method 'bicgstab' (Eigen restatement) or 'exact' (LU).

    Returns (x, iters, rel_err)."""
    T = _c128(T)
    b = _c128(p_src).reshape(-1)
    x = np.zeros_like(b)
    err = np.zeros(1)
    m = 1 if method == "exact" else 0
    it = lib().orc_solve_linear_system(_p(T), _p(b), b.size, m, float(tol), int(maxiter), _p(x), _p(err))
    if it < 0:
        raise np.linalg.LinAlgError("singular matrix")
    return x, it, float(err[0])


def gemv(A, x):
    A = _c128(A)
    x = _c128(x).reshape(-1)
    y = np.empty_like(x)
    lib().orc_gemv(_p(A), _p(x), _p(y), x.size)
    return y


# ---- reductions -------------------------------------------------------------------

def particle_cross_sections(p_scat, p_inc, p_src, k):
    ps, pi, pr = (_c128(a).reshape(-1) for a in (p_scat, p_inc, p_src))
    rmax = ps.size // 2
    lmax = rmax_to_lmax(rmax)
    C = [np.zeros(2 * lmax) for _ in range(3)]
    lib().orc_particle_cross_sections(_p(ps), _p(pi), _p(pr), rmax, float(k), *[_p(c) for c in C])
    return tuple(c.reshape(2, lmax) for c in C)


def cluster_cross_sections(p_cluster, p_src, k):
    pc, pr = (_c128(a).reshape(-1) for a in (p_cluster, p_src))
    rmax = pc.size // 2
    lmax = rmax_to_lmax(rmax)
    C = [np.zeros(2 * lmax) for _ in range(3)]
    lib().orc_cluster_cross_sections(_p(pc), _p(pr), rmax, float(k), *[_p(c) for c in C])
    return tuple(c.reshape(2, lmax) for c in C)


def eps0():
    return lib().orc_eps0()


def force(p_scat, p_inc, k, eps_b=1.0, mu_b=1.0):
    ps, pi = _c128(p_scat).reshape(-1), _c128(p_inc).reshape(-1)
    F = np.zeros(3)
    lib().orc_force(_p(ps), _p(pi), ps.size // 2, float(k), float(np.real(eps_b)), float(np.real(mu_b)), _p(F))
    return F


def torque(p_scat, p_inc, k, eps_b=1.0, mu_b=1.0):
    ps, pi = _c128(p_scat).reshape(-1), _c128(p_inc).reshape(-1)
    T = np.zeros(3)
    lib().orc_torque(_p(ps), _p(pi), ps.size // 2, float(k), float(np.real(eps_b)), float(np.real(mu_b)), _p(T))
    return T


def force_all(p_scat, p_inc, k, eps_b=1.0, mu_b=1.0):
    N = p_scat.shape[0]
    ps, pi = _c128(p_scat).reshape(N, -1), _c128(p_inc).reshape(N, -1)
    F = np.zeros((N, 3))
    lib().orc_force_all(_p(ps), _p(pi), N, ps.shape[1] // 2, float(k), float(np.real(eps_b)), float(np.real(mu_b)), _p(F))
    return F


def torque_all(p_scat, p_inc, k, eps_b=1.0, mu_b=1.0):
    N = p_scat.shape[0]
    ps, pi = _c128(p_scat).reshape(N, -1), _c128(p_inc).reshape(N, -1)
    T = np.zeros((N, 3))
    lib().orc_torque_all(_p(ps), _p(pi), N, ps.shape[1] // 2, float(k), float(np.real(eps_b)), float(np.real(mu_b)), _p(T))
    return T


def max_threads():
    return lib().orc_max_threads()


def set_threads(n):
    """OpenMP threads of the oracle's parallel loops (timing only)."""
    lib().orc_set_threads(int(n))
