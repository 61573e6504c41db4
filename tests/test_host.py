"""Host-side pieces of the package vs the oracle's restatement of the
reference's Python host (plane-wave mode tables, workloads, backend
registry, serving arena).  CPU only; the Mie coefficients and per-particle
p_src run on the device (tests/test_gpu_inputs.py)."""

import numpy as np
import pytest

nm = 1e-9


@pytest.fixture(scope="module")
def H(oracle):
    import oracle.host as H
    return H


def test_plane_wave_mode_table_at_origin(H):
    """The host-side per-wave table equals the oracle's p_src of a particle
    at the origin (phase 0): the device applies only e^{i k r . khat}."""
    import miepy_amd as M
    for pol, theta, phi in [((1, 0), 0.0, 0.0), ((1, 1j), 0.7, 2.1), ((0.3, -0.8j), np.pi, 5.5)]:
        w = M.plane_wave(pol, theta=theta, phi=phi, amplitude=1.3)
        ref = H.plane_wave_structure(np.zeros((1, 3)), 1e7, 4, pol, theta, phi, amplitude=1.3)[0]
        np.testing.assert_allclose(w.mode_table(4), ref, rtol=1e-13, atol=1e-15)


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
    # the reference's own name stays valid (src/miepy/backends/__init__.py:16-28)
    assert "cpu" in M.backends.available_backends()
    with M.backends.backend("cpu"):
        assert M.backends.get_backend() == "cpu"
        with pytest.raises(NotImplementedError):
            M.backends.get_backend_module()
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
