"""Host logic of the serving engine and of the sphere-cluster T-matrix
helpers, on the CPU (no device calls).

- Routing (`ServingEngine._route`): FP64 engine: matrix-free for lmax <= 3 and
  n >= 6000 or matrices above a quarter of the budget ("auto"), as forced
  otherwise; mixed engine: refined systems on stored matrices or, for
  lmax <= 5, the complex64 matrix-free operator; lmax 6 systems that are to
  run matrix-free (forced, or matrices above a quarter of the budget) leave
  the refined set for the FP64 matrix-free set.
- tmatrix_reduce_lmax (src/miepy/tmatrix/functions.py:6-24): the leading
  block, and a ValueError above the T-matrix's own lmax.
"""

import types

import numpy as np
import pytest


def _w(lmax, N):
    return types.SimpleNamespace(lmax=lmax, N=N, n=2 * lmax * (lmax + 2) * N)


def _engine(precision, matrix_free, budget):
    return types.SimpleNamespace(precision=precision, matrix_free=matrix_free, budget=float(budget))


def test_serving_routes():
    from miepy_amd.serve import ServingEngine
    route = lambda eng, w: ServingEngine._route(eng, w)
    big = 200e9
    # FP64 engine
    e = _engine("fp64", "auto", big)
    assert route(e, _w(3, 100)) == (False, False)        # small: stored
    assert route(e, _w(3, 500)) == (True, False)         # lmax 3, n = 15000: matrix-free
    assert route(e, _w(4, 200)) == (False, False)        # lmax 4, n = 9600: stored fits
    assert route(_engine("fp64", "auto", 1e9), _w(5, 60)) == (True, False)  # 16 n^2 = 0.28 GB > budget / 4
    assert route(_engine("fp64", True, big), _w(2, 10)) == (True, False)
    assert route(_engine("fp64", False, big), _w(3, 5000)) == (False, False)
    # mixed engine
    e = _engine("mixed", "auto", big)
    assert route(e, _w(3, 100)) == (False, True)
    assert route(e, _w(4, 300)) == (True, True)          # lmax 4, n = 14400: complex64 matrix-free
    assert route(e, _w(5, 80)) == (False, True)          # lmax 5, n = 5600: stored
    assert route(e, _w(5, 100)) == (True, True)          # lmax 5, n = 7000: complex64 matrix-free
    small = _engine("mixed", "auto", 1.2e9)
    assert route(small, _w(5, 60)) == (True, True)       # lmax 5, 24 n^2 = 0.42 GB > budget / 4: matrix-free
    assert route(small, _w(6, 44)) == (True, False)      # lmax 6, 0.43 GB: FP64 matrix-free
    assert route(small, _w(5, 29)) == (False, True)      # 0.1 GB: stored, refined
    assert route(small, _w(6, 20)) == (False, True)
    assert route(_engine("mixed", True, big), _w(6, 10)) == (True, False)
    assert route(_engine("mixed", True, big), _w(5, 10)) == (True, True)
    assert route(_engine("mixed", True, big), _w(4, 10)) == (True, True)
    assert route(_engine("mixed", False, big), _w(5, 2000)) == (False, True)


def test_tmatrix_reduce_lmax():
    from miepy_amd.tmatrix import tmatrix_reduce_lmax
    rng = np.random.default_rng(0)
    T = rng.normal(size=(2, 24, 2, 24)) + 1j * rng.normal(size=(2, 24, 2, 24))
    R = tmatrix_reduce_lmax(T, 2)
    assert R.shape == (2, 8, 2, 8) and np.array_equal(R, T[:, :8, :, :8])
    assert np.array_equal(tmatrix_reduce_lmax(T, 4), T)
    with pytest.raises(ValueError):
        tmatrix_reduce_lmax(T, 5)


def test_request_final_status():
    """A request is "done" only when every system converged; wavelengths too
    large for the budget make it "partial", a failed system "failed", and a
    request none of whose systems fit "too_large"."""
    from miepy_amd.serve import Request
    r = Request(0, [None, None, None], 0.0)
    r._finish(1.0)
    assert r.status == "done" and r.latency == 1.0
    r = Request(0, [None, None], 0.0)
    r.rejected = 1
    r._finish(1.0)
    assert r.status == "partial"
    r.failed = 1
    r._finish(1.0)
    assert r.status == "failed"
    r = Request(0, [None, None], 0.0)
    r.rejected = 2
    r._finish(1.0)
    assert r.status == "too_large"
