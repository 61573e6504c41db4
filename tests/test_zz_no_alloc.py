"""mpa_bicgstab_noalloc: the BiCGSTAB solve with its workspace in the
library's static device scratch (no workspace argument, no allocation).

Two calls in a row, the second on a smaller system, each checked against
mpa_bicgstab (the same solve with a caller-provided workspace): iteration
counts, residuals and statuses equal, solutions equal bit for bit, for
complex128 and complex64.  The second call finds the first call's contents
in the scratch and must not depend on them."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

nm = 1e-9


@pytest.fixture(scope="module")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    import miepy_amd.device as D
    D.require_gpu()
    return D


def _system(D, N, lmax, S, seed, dtype):
    from miepy_amd.mie import mie_coefficients
    from miepy_amd.workloads import random_cluster, random_plane_waves
    rng = np.random.default_rng(seed)
    a = 80 * nm
    pos = random_cluster(N, 600 * nm * N ** (1 / 3), a, rng)
    k = 2 * np.pi / (700 * nm)
    ms, _ = mie_coefficients(np.full(N, a), np.full(N, -12 + 0.8j), np.ones(N), 1.0, 1.0, k, lmax)
    A = D.assemble(pos, ms, k, dtype=dtype, add_identity=True)
    b = np.stack([s.structure(pos, k, lmax).reshape(-1) for s in random_plane_waves(S, rng, random_pol=True)])
    return A, D.to_device(b, dtype)


def _call(D, L, name, A, b, tol, maxiter):
    import torch
    lib = L.load()
    S, n = b.shape
    code = L.MPA_C128 if A.dtype == D.C128 else L.MPA_C64
    x = torch.full_like(b, float("nan"))
    it = (ctypes.c_int * S)()
    res = (ctypes.c_double * S)()
    st = (ctypes.c_int * S)()
    if name == "noalloc":
        rc = lib.mpa_bicgstab_noalloc(n, S, code, D._ptr(A), 0, D._ptr(b), D._ptr(x), tol, maxiter, it, res, st,
                                      D._stream())
    else:
        ws = torch.empty(lib.mpa_bicgstab_workspace_bytes(n, S, code), dtype=torch.uint8, device=b.device)
        rc = lib.mpa_bicgstab(n, S, code, D._ptr(A), 0, D._ptr(b), D._ptr(x), tol, maxiter, D._ptr(ws), it, res, st,
                              1, D._stream())
    torch.cuda.synchronize()
    return rc, x.cpu().numpy(), list(it), list(res), list(st)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_noalloc_equals_workspace_entry_point(dev, dtype):
    from miepy_amd import _lib as L
    dt = dev.C128 if dtype == "c128" else dev.C64
    tol = 1e-10 if dtype == "c128" else 1e-5
    for N, lmax, S, seed in ((40, 3, 3, 1), (12, 2, 1, 2)):  # the second call is smaller
        print(f"[no_alloc] {dtype}: N={N} lmax={lmax} S={S}", flush=True)
        A, b = _system(dev, N, lmax, S, seed, dt)
        rc0, x0, it0, res0, st0 = _call(dev, L, "noalloc", A, b, tol, 2000)
        rc1, x1, it1, res1, st1 = _call(dev, L, "workspace", A, b, tol, 2000)
        assert rc0 == rc1 == 0, (rc0, rc1, L.last_error())
        assert it0 == it1 and res0 == res1 and st0 == st1
        assert np.array_equal(x0, x1)
        assert np.all(np.isfinite(x0))
