"""A refinement correction that leaves the FP64 residual larger is taken back,
in the batched refinement (mpa_solve_refined_batched, several systems: the
flat gather/scatter of the per-RHS values) and in the iteration-level driver
that serves requests (mpa_refine_systems_start / _advance).

System 0 is built so that its first correction must fail: its complex64
operator is I while its FP64 system is 4 I, so the correction x = r raises
the residual from b to -3 b.  Taking it back leaves ||b - A x|| = ||b||
(without the revert it would be 3 ||b||), and the RHS stops as not
converged.  System 1, a physical cluster, converges beside it with exactly
the result it has alone."""

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


def _pair(dev):
    import torch
    from miepy_amd.mie import mie_coefficients
    from miepy_amd.workloads import random_cluster, random_plane_waves
    rng = np.random.default_rng(11)
    n0 = 96
    A64_0 = dev.to_device(4.0 * np.eye(n0, dtype=complex), dev.C128)
    A32_0 = dev.to_device(np.eye(n0, dtype=complex), dev.C64)
    b0 = dev.to_device(rng.normal(size=(2, n0)) + 1j * rng.normal(size=(2, n0)), dev.C128)
    N, lmax, a = 30, 3, 80 * nm
    pos = random_cluster(N, 600 * nm * N ** (1 / 3), a, rng)
    k = 2 * np.pi / (650 * nm)
    ms, _ = mie_coefficients(np.full(N, a), np.full(N, 4.0 + 0.1j), np.ones(N), 1.0, 1.0, k, lmax)
    A64_1, A32_1 = dev.assemble_dual(pos, ms, k, add_identity=True)
    srcs = random_plane_waves(3, rng, random_pol=True)
    b1 = dev.to_device(np.stack([s.structure(pos, k, lmax).reshape(-1) for s in srcs]), dev.C128)
    torch.cuda.synchronize()
    return (A32_0, A64_0, b0), (A32_1, A64_1, b1)


def _check_reverted(A64_0, b0, x0, status, resid):
    A = A64_0.cpu().numpy()
    b = b0.cpu().numpy()
    x = x0.cpu().numpy()
    for s in range(b.shape[0]):
        r = np.linalg.norm(b[s] - A @ x[s])
        assert r <= np.linalg.norm(b[s]) * (1 + 1e-12), (s, r / np.linalg.norm(b[s]))
        assert status[s] != 0
        assert resid[s] == pytest.approx(1.0, rel=1e-12)


def test_batched_refinement_reverts_a_failed_correction(dev):
    (A32_0, A64_0, b0), (A32_1, A64_1, b1) = _pair(dev)
    xs, reps = dev.solve_refined_batched([A32_0, A32_1], [A64_0, A64_1], [b0, b1], tol=1e-10, fallback=False)
    _check_reverted(A64_0, b0, xs[0], reps[0].status, reps[0].residual)
    assert reps[1].converged
    x1, rep1 = dev.solve_refined(A32_1, A64_1, b1, tol=1e-10, fallback=False)
    assert rep1.outer == reps[1].outer and rep1.iterations == reps[1].iterations
    assert np.array_equal(x1.cpu().numpy(), xs[1].cpu().numpy())


def _iterate(dev, systems, chunk=2):
    """Run systems [(A32, A64, b)] through mpa_refine_systems_start/_advance to the end."""
    import torch
    from miepy_amd import _lib as L
    lib = L.load()
    structs = (L.MpaRefineSystem * len(systems))()
    xs, keep = [], []
    for i, (A32, A64, b) in enumerate(systems):
        n, S = b.shape[1], b.shape[0]
        st = torch.empty(lib.mpa_refine_state_bytes(n, S), dtype=torch.uint8, device=b.device)
        x = torch.empty_like(b)
        rs = L.MpaRefineSystem()
        rs.n, rs.nrhs, rs.add_identity = n, S, 0
        rs.A32, rs.A64, rs.b, rs.x, rs.state = A32.data_ptr(), A64.data_ptr(), b.data_ptr(), x.data_ptr(), st.data_ptr()
        rs.tol, rs.inner_tol, rs.maxiter, rs.max_refine = 1e-10, 1e-5, 2000, 8
        structs[i] = rs
        xs.append(x)
        keep.append(st)
    scratch = torch.empty(lib.mpa_refine_scratch_bytes(len(systems), structs), dtype=torch.uint8, device="cuda")
    L.check(lib.mpa_refine_systems_start(len(systems), structs, dev._ptr(scratch), dev._stream()), "start")
    total = sum(b.shape[0] for _, _, b in systems)
    done = (ctypes.c_int * total)()
    outer = (ctypes.c_int * total)()
    its = (ctypes.c_int * total)()
    res = (ctypes.c_double * total)()
    stt = (ctypes.c_int * total)()
    for _ in range(10000):
        L.check(lib.mpa_refine_systems_advance(len(systems), structs, dev._ptr(scratch), chunk, done, outer, its,
                                               res, stt, dev._stream()), "advance")
        if all(done):
            break
    assert all(done)
    torch.cuda.synchronize()
    return xs, list(stt), list(res), list(outer), list(its)


def test_iteration_level_refinement_reverts_a_failed_correction(dev):
    sys0, sys1 = _pair(dev)
    xs, stt, res, outer, its = _iterate(dev, [sys0, sys1])
    S0 = sys0[2].shape[0]
    _check_reverted(sys0[1], sys0[2], xs[0], stt[:S0], res[:S0])
    assert all(v == 0 for v in stt[S0:])
    x1, stt1, res1, outer1, its1 = _iterate(dev, [sys1])
    assert outer1[0:] == outer[S0:] and its1 == its[S0:]
    assert np.array_equal(x1[0].cpu().numpy(), xs[1].cpu().numpy())
