"""Host-side pieces of the package vs the oracle's restatement of the
reference's Python host (Mie coefficients, plane-wave p_src, workloads,
backend registry).  CPU only."""

import numpy as np
import pytest

nm = 1e-9


@pytest.fixture(scope="module")
def H(oracle):
    import oracle.host as H
    return H


def test_mie_coefficients_match_reference_loop(H):
    from miepy_amd.mie import mie_coefficients
    rng = np.random.default_rng(0)
    N = 40
    radius = rng.uniform(50, 150, N) * nm
    eps = rng.uniform(2, 6, N) + 1j * rng.uniform(0, 0.1, N)
    eps[::3] = -20 + 1.2j
    k = 2 * np.pi * 1.33 / (640 * nm)
    a, c = mie_coefficients(radius, eps, np.ones(N), 1.33 ** 2, 1.0, k, 5)
    ra, rc = H.mie_coefficients(radius, eps, np.ones(N), 1.33 ** 2, 1.0, k, 5)
    np.testing.assert_allclose(a, ra, rtol=1e-12)
    np.testing.assert_allclose(c, rc, rtol=1e-12)


def test_mie_conducting_limit(H):
    from miepy_amd.mie import mie_coefficients
    k = 2 * np.pi / (600 * nm)
    a, c = mie_coefficients([80 * nm], [1 + 1e40j], [1.0], 1.0, 1.0, k, 3, conducting=[True])
    ra, rc = H.mie_coefficients([80 * nm], [1 + 1e40j], [1.0], 1.0, 1.0, k, 3, conducting=[True])
    np.testing.assert_allclose(a, ra, rtol=1e-13)
    assert np.all(c == 0)


def test_mie_rayleigh_limit():
    """a_1 -> -(2i/3) x^3 (m^2-1)/(m^2+2) for x -> 0 (Bohren & Huffman 5.1)."""
    from miepy_amd.mie import mie_coefficients
    eps = 3.0 + 0.5j
    x = 1e-3
    a, _ = mie_coefficients([x], [eps], [1.0], 1.0, 1.0, 1.0, 1)
    expect = -2j / 3 * x ** 3 * (eps - 1) / (eps + 2)
    assert abs(a[0, 0, 0] - expect) < 1e-5 * abs(expect)


def test_mie_lossless_energy_conservation():
    from miepy_amd.mie import mie_coefficients
    a, _ = mie_coefficients([120 * nm], [2.56], [1.0], 1.0, 1.0, 2 * np.pi / (500 * nm), 6)
    np.testing.assert_allclose(a.real, np.abs(a) ** 2, atol=1e-15)


@pytest.mark.parametrize("pol,theta,phi", [((1, 0), 0.0, 0.0), ((0, 1), np.pi, 0.0), ((1, 1j), 0.7, 2.1),
                                           ((0.3, -0.8j), 2.9, 5.5), ((1, 0), np.pi / 2, np.pi / 2)])
def test_plane_wave_structure(H, pol, theta, phi):
    import miepy_amd as M
    rng = np.random.default_rng(1)
    pos = rng.normal(size=(9, 3)) * 300 * nm
    k = 2 * np.pi / (700 * nm)
    got = M.plane_wave(pol, theta=theta, phi=phi).structure(pos, k, 4)
    ref = H.plane_wave_structure(pos, k, 4, pol, theta, phi)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-14)


def test_plane_wave_from_string():
    import miepy_amd as M
    s = M.plane_wave.from_string("x")
    assert s.theta == 0 and np.allclose(s.polarization, [1, 0])
    s = M.plane_wave.from_string("rhc")
    assert np.allclose(s.polarization, np.array([1, 1j]) / np.sqrt(2))
    with pytest.raises(ValueError):
        M.plane_wave.from_string("z")


def test_workloads():
    from miepy_amd import workloads as W
    h = W.hex37()
    assert h.N == 37 and h.n == 592
    w = W.rand1000()
    assert w.n == 30000
    d = np.linalg.norm(w.position[:, None] - w.position[None], axis=-1) + np.eye(w.N) * 1e9
    assert d.min() >= 2.2 * 100 * nm * (1 - 1e-12)
    assert np.all(w.position >= 100 * nm) and np.all(w.position <= 10e-6 - 100 * nm)
    assert np.array_equal(w.position, W.rand1000().position)  # seeded
    b = W.batch256(count=8)
    assert all(2 <= c.lmax <= 4 and 10 <= c.N <= 400 and len(c.sources) == 8 for c in b)


def test_backend_registry():
    import miepy_amd as M
    assert M.backends.get_backend() == "mi355x"
    with pytest.raises(ValueError):
        M.backends.set_backend("nope")
    sentinel = object()
    M.backends.register_backend("test-alt", sentinel)
    with M.backends.backend("test-alt"):
        assert M.backends.get_backend() == "test-alt"
        assert M.backends.get_backend_module() is sentinel
    assert M.backends.get_backend() == "mi355x"


def test_materials():
    import miepy_amd as M
    w = M.materials.water()
    assert w.eps(500e-9) == pytest.approx(1.33 ** 2)
    d = M.data_material([400e-9, 600e-9], [1 + 1j, 3 + 2j])
    assert d.eps(500e-9) == pytest.approx(2 + 1.5j)
    ms = M.material_struct([M.constant_material(4.0)], M.materials.water(), wavelength=500e-9)
    assert ms.k_b == pytest.approx(2 * np.pi * 1.33 / 500e-9)


def test_serving_arena_first_fit_and_coalescing():
    """The serving engine's HBM arena (run here on a host buffer): first-fit,
    256-byte alignment, typed views, and freed neighbours merge back."""
    import torch
    from miepy_amd.serve import Arena
    a = Arena(4096 + 100, device="cpu")
    assert a.size == 4096
    o1 = a.alloc(1000)
    o2 = a.alloc(1000)
    o3 = a.alloc(1000)
    assert (o1, o2, o3) == (0, 1024, 2048) and a.used == 3072
    assert a.alloc(2000) is None  # 1024 bytes left
    a.free(o2)
    assert a.alloc(1100) is None  # the hole is 1024: first fit does not split across blocks
    o4 = a.alloc(512)
    assert o4 == 1024  # first fit reuses the hole
    a.free(o1)
    a.free(o4)
    assert a.free_list[0] == (0, 2048)  # o1 + the hole merged
    a.free(o3)
    assert a.free_list == [(0, 4096)] and a.used == 0
    v = a.view(a.alloc(256), 256, torch.complex128, (4, 4))
    v[:] = 1 + 2j
    assert v.dtype == torch.complex128 and complex(v.sum()) == 16 + 32j
