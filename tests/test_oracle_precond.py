"""Pins the oracle's block-preconditioned BiCGSTAB and the host builders of the
block preconditioner.

Reference: cpp/src/interactions.cpp:211-228 (apply_block_preconditioner),
:230-288 (bicgstab_preconditioned_profiled), :291-345
(BlockDiagonalPreconditioner + solve_linear_system_preconditioned, Eigen's
right-preconditioned BiCGSTAB); src/miepy/interactions.py:9-68 (builders).
The reference's tests do not cover this path; what pins it here:

- M^-1 = I reproduces the unpreconditioned iterates bit for bit (both loops),
  so the preconditioned code adds nothing but the two applications;
- right preconditioning with the exact inverse of a block-diagonal A makes
  A M^-1 = I: one iteration, x = A^-1 b (a left preconditioner, or an update
  with p in place of M^-1 p, fails this);
- coupled systems: the solution equals LU and the true residual meets the
  tolerance; the block-Jacobi preconditioner of a block-dominant system cuts
  the iteration count;
- builders: M_i (I + T_i) = I for the particle builder, M (1 + a_n) = 1 with
  the reference's clamp for spheres; the backend's vectorized builders equal
  the oracle's loops.
"""

import numpy as np
import pytest


def _block_dominant(rng, N, bs, coupling):
    n = N * bs
    A = coupling * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) / np.sqrt(n)
    for i in range(N):
        sl = slice(i * bs, (i + 1) * bs)
        A[sl, sl] = np.eye(bs) * (2.0 + rng.normal(size=bs) + 1j * rng.normal(size=bs)) \
            + 0.6 * (rng.normal(size=(bs, bs)) + 1j * rng.normal(size=(bs, bs)))
    return A


def _block_inverse(A, N, bs):
    M = np.zeros((N * bs, bs), dtype=complex)
    for i in range(N):
        sl = slice(i * bs, (i + 1) * bs)
        M[sl] = np.linalg.inv(A[sl, sl])
    return M


def test_identity_preconditioner_is_bit_identical(oracle):
    rng = np.random.default_rng(0)
    N, bs = 6, 8
    n = N * bs
    A = np.eye(n) + 0.3 * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) / np.sqrt(n)
    b = rng.normal(size=n) + 1j * rng.normal(size=n)
    M = np.tile(np.eye(bs), (N, 1)).astype(complex)
    x0, i0, e0 = oracle.bicgstab_eigen(A, b, 1e-12, 200)
    x1, i1, e1 = oracle.bicgstab_eigen(A, b, 1e-12, 200, M, bs)
    assert i0 == i1 and e0 == e1 and np.array_equal(x0, x1)
    y0, j0, r0 = oracle.bicgstab(A, b, 200, 1e-10)
    y1, j1, r1 = oracle.bicgstab(A, b, 200, 1e-10, M, bs)
    assert j0 == j1 and r0 == r1 and np.array_equal(y0, y1)


def test_exact_block_inverse_converges_in_one_iteration(oracle):
    rng = np.random.default_rng(1)
    N, bs = 5, 6
    A = _block_dominant(rng, N, bs, coupling=0.0)
    b = rng.normal(size=N * bs) + 1j * rng.normal(size=N * bs)
    M = _block_inverse(A, N, bs)
    ref = np.linalg.solve(A, b)
    x, it, err = oracle.bicgstab_eigen(A, b, 1e-12, 50, M, bs)
    assert it == 1 and np.linalg.norm(x - ref) <= 1e-12 * np.linalg.norm(ref)
    x, it, res = oracle.bicgstab(A, b, 50, 1e-10, M, bs)
    assert it == 1 and np.linalg.norm(x - ref) <= 1e-12 * np.linalg.norm(ref)


@pytest.mark.parametrize("coupling", [0.2, 0.6])
def test_preconditioned_solution_and_fewer_iterations(oracle, coupling):
    rng = np.random.default_rng(2)
    N, bs = 8, 6
    A = _block_dominant(rng, N, bs, coupling)
    b = rng.normal(size=N * bs) + 1j * rng.normal(size=N * bs)
    M = _block_inverse(A, N, bs)
    ref = oracle.lu_solve(A, b)
    x, it, err = oracle.bicgstab_eigen(A, b, 1e-11, 500, M, bs)
    assert np.linalg.norm(A @ x - b) <= 1.01e-11 * np.linalg.norm(b)
    assert np.linalg.norm(x - ref) <= 1e-8 * np.linalg.norm(ref)
    _, it_plain, _ = oracle.bicgstab_eigen(A, b, 1e-11, 500)
    assert it < it_plain
    y, jt, res = oracle.bicgstab(A, b, 500, 1e-9, M, bs)
    assert res < 1e-9 and abs(np.linalg.norm(A @ y - b) - res) <= 1e-6 * res + 1e-14
    assert np.linalg.norm(y - ref) <= 1e-6 * np.linalg.norm(ref)


def test_reference_loop_caps_at_maxiter_plus_one(oracle):
    rng = np.random.default_rng(3)
    n = 40
    A = np.eye(n) + 0.9 * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) / np.sqrt(n)
    b = rng.normal(size=n) + 1j * rng.normal(size=n)
    M = np.tile(np.eye(8), (5, 1)).astype(complex)
    x, it, res = oracle.bicgstab(A, b, 3, 0.0, M, 8)
    assert it == 4 and res > 0  # interactions.cpp:279 stops once current_iteration > maxiter


def test_particle_builder_inverts_blocks(oracle):
    from miepy_amd import backend as B
    rng = np.random.default_rng(4)
    N, lmax = 4, 2
    rmax = lmax * (lmax + 2)
    bs = 2 * rmax
    tm = 0.3 * (rng.normal(size=(N, 2, rmax, 2, rmax)) + 1j * rng.normal(size=(N, 2, rmax, 2, rmax)))
    tm[2] = 0.0
    tm[2, 0, 0, 0, 0] = -1.0  # I + T_2 singular: identity block
    M, b = oracle.build_block_preconditioner_particle(tm)
    assert b == bs and M.shape == (N * bs, bs)
    for i in range(N):
        blk = M[i * bs:(i + 1) * bs]
        if i == 2:
            assert np.array_equal(blk, np.eye(bs))
        else:
            assert np.allclose(blk @ (np.eye(bs) + tm[i].reshape(bs, bs)), np.eye(bs), atol=1e-12)
    M2, b2 = B.build_block_preconditioner_particle(tm)
    assert b2 == bs and np.allclose(M2, M, rtol=0, atol=1e-13)


def test_sphere_builder_inverts_diagonal(oracle):
    from miepy_amd import backend as B
    rng = np.random.default_rng(5)
    N, lmax = 3, 3
    rmax = lmax * (lmax + 2)
    mie = rng.normal(size=(N, 2, lmax)) + 1j * rng.normal(size=(N, 2, lmax))
    mie[1, 0, 1] = -1.0  # 1 + a_2 = 0: clamped to 1 (interactions.py:33-35)
    M, bs = oracle.build_block_preconditioner_sphere(mie, lmax)
    assert bs == 2 * rmax
    for i in range(N):
        blk = M[i * bs:(i + 1) * bs]
        assert np.count_nonzero(blk - np.diag(np.diag(blk))) == 0
        for pol in range(2):
            r = 0
            for n in range(1, lmax + 1):
                for _m in range(2 * n + 1):
                    v = blk[pol * rmax + r, pol * rmax + r]
                    d = 1 + mie[i, pol, n - 1]
                    assert (v == 1.0) if abs(d) < 1e-12 else abs(v * d - 1) < 1e-14
                    r += 1
    M2, bs2 = B.build_block_preconditioner_sphere(mie, lmax)
    assert bs2 == bs and np.array_equal(M2, M)


def test_cabi_rejects_bad_block_size():
    """Argument checks of mpa_bicgstab_ex run before any device call."""
    import ctypes
    from miepy_amd import _lib
    lib = _lib.load()
    dummy = ctypes.c_void_p(16)
    rc = lib.mpa_bicgstab_ex(12, 1, _lib.MPA_C128, dummy, 1, dummy, dummy, 1e-5, 10, 0, dummy, 5, dummy,
                             None, None, None, 1, None)
    assert rc == _lib.MPA_ERR_ARG and "block_size" in _lib.last_error()
    rc = lib.mpa_bicgstab_ex(12, 1, _lib.MPA_C128, dummy, 1, dummy, dummy, 1e-5, 10, 8, None, 0, dummy,
                             None, None, None, 1, None)
    assert rc == _lib.MPA_ERR_ARG
