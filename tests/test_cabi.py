"""The C-ABI library loads and exports exactly what include/miepy_amd.h declares
(no GPU needed: nothing here calls into the device)."""

import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "miepy_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mpa_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = _declared()
    for must in ("mpa_sphere_aggregate_tmatrix", "mpa_matvec", "mpa_bicgstab", "mpa_solve_refined",
                 "mpa_cluster_observables", "mpa_force_torque", "mpa_particle_cross_sections"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from miepy_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libmiepy_amd.so is not built (python -m miepy_amd.build)")
    lib = _lib.load()  # loading needs the HIP runtime library, not a GPU
    for name in _declared():
        assert hasattr(lib, name), name
    # every declared function has a ctypes signature in the binding
    assert sorted(_lib.SIGNATURES) == _declared()


def test_library_reports_bad_arguments_without_gpu():
    """Argument validation happens before any device call."""
    import ctypes
    from miepy_amd import _lib
    lib = _lib.load()
    rc = lib.mpa_matvec(0, 1, 0, None, 0, None, None, None)
    assert rc == _lib.MPA_ERR_ARG
    assert "bad argument" in _lib.last_error()
    rc = lib.mpa_sphere_aggregate_tmatrix(0, 2, None, None, 1.0, 0, 0, None, None)
    assert rc == _lib.MPA_ERR_ARG
    # the allocation-free solve refuses a workspace larger than its static scratch before touching the device
    one = ctypes.c_void_p(256)
    it, res, st = (ctypes.c_int * 1)(), (ctypes.c_double * 1)(), (ctypes.c_int * 1)()
    assert lib.mpa_bicgstab_workspace_bytes(1000000, 1, _lib.MPA_C128) > 64 << 20
    rc = lib.mpa_bicgstab_noalloc(1000000, 1, _lib.MPA_C128, one, 0, one, one, 1e-8, 10, it, res, st, None)
    assert rc == _lib.MPA_ERR_ARG and "static scratch" in _lib.last_error()
    assert lib.mpa_bicgstab_workspace_bytes(2000, 300, _lib.MPA_C128) <= 64 << 20  # n * nrhs = 6e5 fits


def test_package_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    import numpy as np
    import miepy_amd
    with pytest.raises(RuntimeError, match="needs a HIP GPU"):
        miepy_amd.backend.sphere_aggregate_tmatrix(np.zeros((2, 3)), np.ones((2, 2, 2), complex), 1.0)
