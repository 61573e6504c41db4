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


def test_package_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    import numpy as np
    import miepy_amd
    with pytest.raises(RuntimeError, match="needs a HIP GPU"):
        miepy_amd.backend.sphere_aggregate_tmatrix(np.zeros((2, 3)), np.ones((2, 2, 2), complex), 1.0)
