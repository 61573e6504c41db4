"""The complex64 matrix-free operator at lmax 5 (k_matfree_mf16k2: the
FP16-split contraction with the 174 KB operand block of a column staged in K
slices through two LDS buffers by DMA, B operands formed slice by slice, two
workgroups per CU).

- against the FP64 matrix-free operator (which the oracle pins at lmax 5 in
  test_gpu_matfree.py) for 1, 3 and 8 sources: single-chunk (direct-write)
  and many-chunk jobs, idle 32-row waves (N not a multiple of 32), a pair of
  spheres; 3e-6 of the largest entry of each source's (I + T) x, 6e-6 of T x;
- pairs deep in the near field next to well separated pairs (the per-pair
  power-of-two scaling over h_10);
- lmax-5 jobs batched with lmax-3 and lmax-4 jobs in one batched product
  equal their products alone bit for bit;
- a complex64 + FP64-refinement solve through it (`mixed_matfree`) matches
  the FP64 stored solve: cross-sections and p_scat within 1e-8;
- the FP64 matrix-free product at lmax 5 (k_matfree_mf64k, the refinement's
  residuals) equals the assembled matrix times x and the FP64 LDS kernel to
  1e-12, with other data left in the workspace by a complex64 product.
"""

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


def _cluster(N, lmax, seed, spacing=300 * nm):
    from miepy_amd.workloads import random_cluster
    rng = np.random.default_rng(seed)
    pos = random_cluster(N, spacing * N ** (1 / 3) * 1.5, 60 * nm, rng)
    mie = (rng.normal(size=(N, 2, lmax)) + 1j * rng.normal(size=(N, 2, lmax))) * 0.3
    return pos, mie


def _errors(dev, pos, mie, k, S, seed, identity):
    n = 2 * mie.shape[2] * (mie.shape[2] + 2) * pos.shape[0]
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(S, n)) + 1j * rng.normal(size=(S, n))
    op = dev.ClusterOperator(pos, mie, k, max_rhs=S)
    ref = dev.matfree_matvec(op, dev.to_device(X, dev.C128), add_identity=identity).cpu().numpy()
    got = dev.matfree_matvec(op, dev.to_device(X, dev.C64), add_identity=identity).cpu().numpy()
    assert np.all(np.isfinite(got))
    return max(np.max(np.abs(got[s] - ref[s])) / np.max(np.abs(ref[s])) for s in range(S))


@pytest.mark.parametrize("N,S", [(2, 1), (45, 3), (33, 8), (260, 1), (300, 3)])
def test_lmax5_operator_matches_fp64(dev, N, S):
    pos, mie = _cluster(N, 5, seed=700 + N + S)
    k = 2 * np.pi / (600 * nm)
    for seed, identity, bar in ((N, True, 3e-6), (N + 1, False, 6e-6)):
        e = _errors(dev, pos, mie, k, S, seed, identity)
        print(f"[l5] N {N} S {S} identity {identity}: {e:.2e}", flush=True)
        assert e <= bar, (identity, e)


@pytest.mark.parametrize("kd_near,kd_far", [(0.05, 5.0), (2.0, 2000.0)])
def test_lmax5_operator_extreme_separations(dev, kd_near, kd_far):
    k = 2 * np.pi / (600 * nm)
    rng = np.random.default_rng(12)
    centers = rng.normal(size=(6, 3))
    centers *= kd_far / k / np.max(np.linalg.norm(centers[:, None] - centers[None], axis=-1))
    offs = np.array([[0, 0, 0], [kd_near / k, 0, 0], [0, kd_near / k, kd_near / k]])
    pos = (centers[:, None, :] + offs[None]).reshape(-1, 3)
    mie = (rng.normal(size=(len(pos), 2, 5)) + 1j * rng.normal(size=(len(pos), 2, 5))) * 1e-3
    e = _errors(dev, pos, mie, k, 1, 5, False)
    print(f"[l5] kd {kd_near} / {kd_far}: {e:.2e}", flush=True)
    assert e <= 6e-6


def test_lmax5_batched_with_other_lmax_equals_alone(dev):
    import torch
    k = 2 * np.pi / (600 * nm)
    ops, xs = [], []
    rng = np.random.default_rng(4)
    for lmax, N, S in ((5, 70, 2), (3, 90, 3), (5, 150, 1), (4, 40, 2)):
        pos, mie = _cluster(N, lmax, seed=N + lmax)
        op = dev.ClusterOperator(pos, mie, k, max_rhs=S)
        ops.append(op)
        xs.append(dev.to_device(rng.normal(size=(S, op.n)) + 1j * rng.normal(size=(S, op.n)), dev.C64))
    ys = dev.matfree_matvec_batched(ops, xs, add_identity=True)
    torch.cuda.synchronize()
    for op, x, y in zip(ops, xs, ys):
        one = dev.matfree_matvec(op, x, add_identity=True)
        assert np.array_equal(one.cpu().numpy(), y.cpu().numpy())


def test_lmax5_mixed_matfree_solve(dev):
    import miepy_amd as M
    from miepy_amd.workloads import scaled_random
    w = scaled_random(40, 5, seed=8, nsrc=2)
    kw = dict(position=w.position, radius=w.radius, material=w.materials(), source=w.sources,
              wavelength=w.wavelength, lmax=5, medium=w.medium())
    ref = M.sphere_cluster(**kw, precision="fp64", tol=1e-12)
    got = M.sphere_cluster(**kw, precision="mixed_matfree", tol=1e-10)
    C_ref, C = np.array(ref.cross_sections()), np.array(got.cross_sections())
    e_c = float(np.max(np.abs(C - C_ref)) / np.abs(C_ref).max())
    e_p = float(np.linalg.norm(got.p_scat - ref.p_scat) / np.linalg.norm(ref.p_scat))
    print(f"[l5] mixed_matfree vs FP64: cross-sections {e_c:.2e}, p_scat {e_p:.2e}", flush=True)
    assert e_c <= 1e-8 and e_p <= 1e-8


@pytest.mark.parametrize("N,S", [(3, 1), (70, 2), (130, 3), (200, 1)])
def test_lmax5_fp64_matrix_core_operator_matches_dense(dev, N, S):
    """The FP64 matrix-free product at lmax 5 (k_matfree_mf64k: one column's
    W block in a single LDS buffer filled by DMA) against the assembled matrix
    times x, and against the FP64 LDS kernel (MPA_OPT_MATFREE_FP64_KERNEL = 1,
    checked against the oracle's matrix in test_gpu_matfree.py), 1e-12 of the
    largest entry: one and several column chunks (some of them empty), rows
    past N in the last 64-row group, one to three sources, after a complex64
    product has left other data in the workspace."""
    import miepy_amd._lib as L
    lib = L.load()
    pos, mie = _cluster(N, 5, seed=900 + N)
    k = 2 * np.pi / (600 * nm)
    rng = np.random.default_rng(N)
    op = dev.ClusterOperator(pos, mie, k, max_rhs=S)
    X = dev.to_device(rng.normal(size=(S, op.n)) + 1j * rng.normal(size=(S, op.n)), dev.C128)
    dev.matfree_matvec(op, X.to(dev.C64))  # leaves other data in the shared workspace (partial rows included)
    got = dev.matfree_matvec(op, X, add_identity=True).cpu().numpy()
    A = dev.assemble(pos, mie, k, add_identity=True)
    dense = dev.matvec(A, X).cpu().numpy()
    del A
    assert np.max(np.abs(got - dense)) <= 1e-12 * np.max(np.abs(dense))
    saved = lib.mpa_get_option(L.MPA_OPT_MATFREE_FP64_KERNEL)
    try:
        lib.mpa_set_option(L.MPA_OPT_MATFREE_FP64_KERNEL, 1)
        lds = dev.matfree_matvec(op, X, add_identity=True).cpu().numpy()
    finally:
        lib.mpa_set_option(L.MPA_OPT_MATFREE_FP64_KERNEL, saved)
    assert np.max(np.abs(got - lds)) <= 1e-12 * np.max(np.abs(lds))
