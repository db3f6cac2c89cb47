"""Python mirror of the reference's engine API, over the C ABI of libhph.so.

Names and argument meanings follow the R API of suchard-group/hawkes:
``createEngine`` (src/hpHawkes.cpp:59-109), ``updateLocations`` (R/hph.R:102-110),
``setTimesData`` (R/hph.R:83-91), ``setParameters`` (R/hph.R:68-72), ``getLogLikelihood``
(R/hph.R:10-31), ``getProbsSelfExcite`` (R/hph.R:41-57); plus ``getLogLikeGradient`` and the
batched-proposal calls. As in R, the functions take and return the engine object, and the get
functions refuse to run before locations, times and parameters are set (R/hph.R:12-22).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import HphError, check, dptr

__all__ = ["Engine", "HphError", "createEngine", "updateLocations", "setTimesData", "setParameters",
           "getLogLikelihood", "getLogLikeGradient", "getProbsSelfExcite"]


class Engine:
    """Owns one hph_engine handle (one HIP stream, device buffers for one dataset)."""

    def __init__(self, embeddingDimension, locationCount, flags=_lib.OPENCL, device=0, threads=0):
        L = _lib.lib()
        h = ctypes.c_void_p()
        check(L.hph_create(int(embeddingDimension), int(locationCount), int(flags), int(device), int(threads),
                           ctypes.byref(h)))
        self._h = h
        self.embeddingDimension = int(embeddingDimension)
        self.locationCount = int(locationCount)
        self.flags = int(flags)
        self.timesInitialized = False
        self.locationsInitialized = False
        self.parametersInitialized = False

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().hph_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- data and parameters --
    def update_locations(self, locations, index=-1):
        a = np.ascontiguousarray(locations, dtype=np.float64).reshape(-1)
        check(_lib.lib().hph_update_locations(self._h, int(index), dptr(a), a.size))
        if index == -1:
            self.locationsInitialized = True

    def set_times(self, times):
        a = np.ascontiguousarray(times, dtype=np.float64).reshape(-1)
        check(_lib.lib().hph_set_times_data(self._h, dptr(a), a.size))
        self.timesInitialized = True

    def set_parameters(self, params):
        a = np.ascontiguousarray(params, dtype=np.float64).reshape(-1)
        check(_lib.lib().hph_set_parameters(self._h, dptr(a), a.size))
        self.parametersInitialized = True

    # -- evaluations --
    def log_likelihood(self):
        out = ctypes.c_double()
        check(_lib.lib().hph_get_sum_of_lik_contribs(self._h, ctypes.byref(out)))
        return out.value

    def gradient(self):
        g = np.zeros(6)
        check(_lib.lib().hph_get_log_like_gradient(self._h, dptr(g), 6))
        return g

    def log_likelihood_and_gradient(self):
        ll = np.zeros(1)
        g = np.zeros(6)
        check(_lib.lib().hph_get_log_likelihood_and_gradient(self._h, dptr(ll), dptr(g)))
        return float(ll[0]), g

    def probs_self_excite(self):
        p = np.zeros(self.locationCount)
        check(_lib.lib().hph_get_probs_self_excite(self._h, dptr(p), p.size))
        return p

    def location_gradient(self):
        """d loglik / d locations as an N x D array in the caller's event order."""
        g = np.zeros((self.locationCount, self.embeddingDimension))
        check(_lib.lib().hph_get_location_gradient(self._h, dptr(g), g.size))
        return g

    def set_parameters_batch(self, params):
        a = np.ascontiguousarray(params, dtype=np.float64).reshape(-1)
        check(_lib.lib().hph_set_parameters_batch(self._h, dptr(a), a.size))
        self._batch_k = a.size // 6

    def log_likelihood_batch(self, gradient=True):
        K = self._batch_k
        ll = np.zeros(K)
        g = np.zeros((K, 6)) if gradient else None
        check(_lib.lib().hph_get_log_likelihood_batch(self._h, dptr(ll), dptr(g) if gradient else None, K))
        return (ll, g) if gradient else ll

    def set_pair_kernel(self, kind):
        """_lib.PAIR_KERNEL_AUTO / PAIR_KERNEL_ROWWISE / PAIR_KERNEL_SYMMETRIC / PAIR_KERNEL_SPLIT."""
        check(_lib.lib().hph_set_pair_kernel(self._h, int(kind)))

    def set_pair_config(self, kind, rows_per_lane=2, target_blocks=4096):
        check(_lib.lib().hph_set_pair_config(self._h, int(kind), int(rows_per_lane), int(target_blocks)))

    def autotune(self, repeats=2):
        """Times the pair-pass configurations on the current data and parameters, keeps the
        fastest (process-wide, and in the file named by HPH_TUNE_CACHE); returns its ms."""
        best = np.zeros(1)
        check(_lib.lib().hph_autotune(self._h, int(repeats), dptr(best)))
        return float(best[0])

    def tuned_config(self):
        """(pair kernel kind, rows per lane, target grid size) used for a single evaluation."""
        out = (ctypes.c_int * 3)()
        check(_lib.lib().hph_get_tuned_config(self._h, out, 3))
        return tuple(out)

    # -- kernel timing (HIP events on the engine's stream) --
    def set_timing(self, enabled=True):
        check(_lib.lib().hph_set_timing(self._h, 1 if enabled else 0))

    def last_evaluation(self):
        """How the last single-vector call was formed: _lib.EVAL_FULL, EVAL_EXCITATION (cached
        background sums + excitation pass) or EVAL_CACHED (cached sums only)."""
        out = ctypes.c_int()
        check(_lib.lib().hph_get_last_evaluation(self._h, ctypes.byref(out)))
        return out.value

    def event_sums(self):
        """N x 6 per-event sums (sum B, B r^2, B dt^2, S, S r^2, S dt; natural units, caller's
        order) of the last log-likelihood-and-gradient evaluation at the current parameters."""
        out = np.zeros((self.locationCount, 6))
        check(_lib.lib().hph_get_event_sums(self._h, dptr(out), out.size))
        return out

    def last_pass(self):
        """(pair pass, base-2 form) of the last full evaluation: pass = _lib.PAIR_KERNEL_ROWWISE,
        PAIR_KERNEL_SYMMETRIC or PAIR_KERNEL_SPLIT; base-2 form True when the pass ran on scaled
        coordinates and times (2^-y of a sum of squares), False for the natural form."""
        out = (ctypes.c_int * 2)()
        check(_lib.lib().hph_get_last_pass(self._h, out, 2))
        return int(out[0]), bool(out[1])

    def vectors_per_launch(self):
        """Parameter vectors per launch of the last full evaluation (fewer than the batch's when
        its symmetric or split partial sums did not fit in device memory at once)."""
        out = (ctypes.c_int * 3)()
        check(_lib.lib().hph_get_last_pass(self._h, out, 3))
        return int(out[2])

    def exponent_in_table(self):
        """True when the last full evaluation's time-sorted symmetric pass ran the FP64 table
        exponential with the power of two folded into the table entry (every background exponent
        of the call within its range; results bitwise those of the table exponential with ldexp)."""
        out = (ctypes.c_int * 4)()
        check(_lib.lib().hph_get_last_pass(self._h, out, 4))
        return bool(out[3])

    def timing(self):
        """(pair-pass ms, finalize + reduction ms) of the last evaluation."""
        out = np.zeros(2)
        check(_lib.lib().hph_get_timing(self._h, dptr(out), 2))
        return float(out[0]), float(out[1])

    # -- state (src/AbstractHawkes.hpp:46-48) --
    def store_state(self):
        check(_lib.lib().hph_store_state(self._h))

    def restore_state(self):
        check(_lib.lib().hph_restore_state(self._h))

    def accept_state(self):
        check(_lib.lib().hph_accept_state(self._h))

    def internal_dimension(self):
        out = ctypes.c_int()
        check(_lib.lib().hph_get_internal_dimension(self._h, ctypes.byref(out)))
        return out.value


# ---- R-style functional API (R/hph.R, src/hpHawkes.cpp) ----

def createEngine(embeddingDimension, locationCount, tbb=0, simd=0, gpu=1, single=False):
    """src/hpHawkes.cpp:59-109. gpu > 0 selects device gpu-1 (the R API counts from 1);
    single selects the FP32-bulk mode. The CPU arguments tbb/simd are accepted and unused."""
    flags = _lib.OPENCL
    if single:
        flags |= _lib.FLOAT
    if tbb > 0:
        flags |= _lib.TBB
    if simd == 1:
        flags |= _lib.SSE
    elif simd == 2:
        flags |= _lib.AVX
    device = max(int(gpu) - 1, 0)
    return Engine(embeddingDimension, locationCount, flags=flags, device=device, threads=tbb)


def updateLocations(engine, locations):
    """R/hph.R:102-110: locations is N x D (row-major after R's transpose)."""
    locations = np.asarray(locations, dtype=np.float64)
    if locations.ndim == 1:
        locations = locations[:, None]
    if locations.shape != (engine.locationCount, engine.embeddingDimension):
        raise ValueError("locations must be locationCount x embeddingDimension")
    engine.update_locations(locations)
    return engine


def setTimesData(engine, times):
    """R/hph.R:83-91."""
    engine.set_times(times)
    return engine


def setParameters(engine, parameters):
    """R/hph.R:68-72: (sigmaXprec, tauXprec, tauTprec, omega, theta, mu0)."""
    engine.set_parameters(parameters)
    return engine


def _require_init(engine):
    if not engine.locationsInitialized:
        raise RuntimeError("locations not set")
    if not engine.timesInitialized:
        raise RuntimeError("times not set")
    if not engine.parametersInitialized:
        raise RuntimeError("parameters not set")


def getLogLikelihood(engine):
    """R/hph.R:10-31."""
    _require_init(engine)
    return engine.log_likelihood()


def getLogLikeGradient(engine):
    """Gradient with respect to (sigmaXprec, tauXprec, tauTprec, omega, theta, mu0)."""
    _require_init(engine)
    return engine.gradient()


def getProbsSelfExcite(engine):
    """R/hph.R:41-57."""
    _require_init(engine)
    return engine.probs_self_excite()


def getLocationGradient(engine):
    """d loglik / d locations (N x D); the reference declares it for its Java callers
    (src/jni/dr_inference_hawkes_NativeHPHSingleton.h:92-96) without an implementation."""
    _require_init(engine)
    return engine.location_gradient()
