"""CPU oracle for the Hawkes log-likelihood (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may import
this package, as the checker or as the timed CPU baseline. The product package
``hawkes_amd`` never imports it.

``liboracle.so`` (built from ``hawkes_oracle.c`` by ``oracle/Makefile``) holds:

* ``hawkes_oracle_rows`` / ``hawkes_oracle_total``: the long-double (or double) restatement of
  src/NewHawkes.hpp:520-547, 636-660, 662-719 with compensated sums and the exact gradient.
* ``hawkes_oracle_location_gradient``: d ll / d x (N x D), the exact derivative of the same
  model (the reference declares but never implements a location gradient,
  src/jni/dr_inference_hawkes_NativeHPHSingleton.h:92-96).
* ``hawkes_cpu_baseline``: the reference's serial, AVX and threaded CPU flavours
  (src/NewHawkes.hpp:899-980), timed as the CPU baseline.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

SERIAL, AVX, THREADED = 0, 1, 2
ROW_COLUMNS = ("sB", "sBr2", "sBdt2", "sS", "sSr2", "sSdt", "lambda", "ll", "p",
               "g_sigma", "g_tauX", "g_tauT", "g_omega", "g_theta", "g_mu0")


def lib():
    """Loads oracle/liboracle.so (built by ``make -C oracle``)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built; run `make -C oracle`")
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        L.hawkes_oracle_rows.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_long, dp, dp, dp,
                                         ctypes.c_long, ctypes.c_long, dp]
        L.hawkes_oracle_rows.restype = ctypes.c_int
        L.hawkes_oracle_total.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_long, dp, dp, dp,
                                          dp, dp, dp, dp]
        L.hawkes_oracle_total.restype = ctypes.c_int
        L.hawkes_cpu_baseline.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_long,
                                          dp, dp, dp, ctypes.c_long, ctypes.c_long, dp]
        L.hawkes_cpu_baseline.restype = ctypes.c_int
        L.hawkes_oracle_location_gradient.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_long, dp, dp, dp,
                                                      dp, dp]
        L.hawkes_oracle_location_gradient.restype = ctypes.c_int
        L.hawkes_oracle_num_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _prep(locations, times, params):
    loc = np.ascontiguousarray(locations, dtype=np.float64)
    if loc.ndim == 1:
        loc = loc[:, None]
    t = np.ascontiguousarray(times, dtype=np.float64)
    p = np.ascontiguousarray(params, dtype=np.float64)
    assert loc.shape[0] == t.shape[0] and p.shape == (6,)
    return loc, t, p


def total(locations, times, params, precision="long", probs=False):
    """Returns dict(ll, grad[6], scale[7], probs?) for all rows.

    ``scale[0]`` is sum_i |terms of ll_i| and ``scale[1+k]`` the same for gradient k: the
    magnitudes against which cancelling sums state their error.
    """
    loc, t, p = _prep(locations, times, params)
    N, D = loc.shape
    ll = np.zeros(1)
    g = np.zeros(6)
    sc = np.zeros(7)
    pr = np.zeros(N) if probs else None
    rc = lib().hawkes_oracle_total(0 if precision == "long" else 1, D, N, _p(loc), _p(t), _p(p),
                                   _p(ll), _p(g), _p(sc), _p(pr) if probs else None)
    if rc != 0:
        raise RuntimeError("oracle failed")
    out = {"ll": float(ll[0]), "grad": g, "scale": sc}
    if probs:
        out["probs"] = pr
    return out


def rows(locations, times, params, row_begin=0, row_end=None, precision="long"):
    """Per-row statistics (columns ``ROW_COLUMNS``) for rows [row_begin, row_end)."""
    loc, t, p = _prep(locations, times, params)
    N, D = loc.shape
    if row_end is None:
        row_end = N
    out = np.zeros((row_end - row_begin, 15))
    rc = lib().hawkes_oracle_rows(0 if precision == "long" else 1, D, N, _p(loc), _p(t), _p(p),
                                  row_begin, row_end, _p(out))
    if rc != 0:
        raise RuntimeError("oracle failed")
    return out


def location_gradient(locations, times, params, precision="long"):
    """Returns (grad, scale): d ll / d x as an N x D array in the caller's order, and per
    component the sum of |terms| (the magnitude against which a cancelling sum states its error)."""
    loc, t, p = _prep(locations, times, params)
    N, D = loc.shape
    g = np.zeros((N, D))
    sc = np.zeros((N, D))
    rc = lib().hawkes_oracle_location_gradient(0 if precision == "long" else 1, D, N, _p(loc), _p(t), _p(p),
                                               _p(g), _p(sc))
    if rc != 0:
        raise RuntimeError("oracle failed")
    return g, sc


def cpu_baseline(flavour, locations, times, params, row_begin=0, row_end=None, grad=True):
    """Runs one CPU baseline flavour (SERIAL, AVX, THREADED) over rows [row_begin, row_end).

    Returns an array of 7 (ll partial, 6 gradient partials) or 1 value when grad is False.
    """
    loc, t, p = _prep(locations, times, params)
    N, D = loc.shape
    if row_end is None:
        row_end = N
    out = np.zeros(7)
    rc = lib().hawkes_cpu_baseline(flavour, 1 if grad else 0, D, N, _p(loc), _p(t), _p(p),
                                   row_begin, row_end, _p(out))
    if rc != 0:
        raise RuntimeError("cpu baseline failed")
    return out if grad else out[:1]


def num_threads():
    return lib().hawkes_oracle_num_threads()
