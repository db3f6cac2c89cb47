"""Dataset batches over the C ABI: many independent datasets of uneven size, each with its own
parameter vectors, evaluated by one persistent launch (optionally a replayed HIP graph).

No counterpart in the reference; each (dataset, vector) result equals what an engine created
for that dataset returns for that vector (to rounding of the summation order).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, dptr


class Batch:
    def __init__(self, embeddingDimension, location_counts, flags=_lib.OPENCL, device=0):
        counts = np.ascontiguousarray(location_counts, dtype=np.int32)
        h = ctypes.c_void_p()
        check(_lib.lib().hph_batch_create(int(embeddingDimension), int(counts.size),
                                          counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), int(flags),
                                          int(device), ctypes.byref(h)))
        self._h = h
        self.D = int(embeddingDimension)
        self.counts = counts.astype(int)
        self._total = 0

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().hph_batch_destroy(self._h)
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

    def set_data(self, d, locations, times):
        loc = np.ascontiguousarray(locations, dtype=np.float64).reshape(-1)
        t = np.ascontiguousarray(times, dtype=np.float64).reshape(-1)
        check(_lib.lib().hph_batch_set_data(self._h, int(d), dptr(loc), loc.size, dptr(t), t.size))

    def set_parameters(self, params_per_dataset):
        """params_per_dataset: list (one per dataset) of K_d x 6 arrays."""
        ks = np.array([np.atleast_2d(p).shape[0] for p in params_per_dataset], dtype=np.int32)
        flat = np.ascontiguousarray(np.concatenate([np.atleast_2d(p) for p in params_per_dataset]), dtype=np.float64)
        check(_lib.lib().hph_batch_set_parameters(self._h, ks.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                                  dptr(flat), flat.size))
        self._ks = ks
        self._total = int(ks.sum())

    def log_likelihood(self, gradient=True):
        """Returns (loglik[total], grad[total, 6]) (or loglik only), in parameter order."""
        ll = np.zeros(self._total)
        g = np.zeros((self._total, 6)) if gradient else None
        check(_lib.lib().hph_batch_get_log_likelihood(self._h, dptr(ll), dptr(g) if gradient else None, self._total))
        return (ll, g) if gradient else ll

    def split(self, values):
        """Splits a per-vector result array into one array per dataset."""
        return np.split(values, np.cumsum(self._ks)[:-1])

    def set_graph(self, enabled=True):
        check(_lib.lib().hph_batch_set_graph(self._h, 1 if enabled else 0))

    def set_config(self, rows_per_lane=2, target_items=8192):
        check(_lib.lib().hph_batch_set_config(self._h, int(rows_per_lane), int(target_items)))

    def autotune(self, repeats=1):
        """Times the work decompositions on the current datasets and vectors and keeps the fastest
        for this table's shape (shared with every batch and engine of the process); returns the
        winner's device time in ms."""
        out = ctypes.c_double()
        check(_lib.lib().hph_batch_autotune(self._h, int(repeats), ctypes.byref(out)))
        return out.value

    def config(self):
        """(rows per lane, target items, source) of the next evaluation; source 0 built-in, 1 tuned,
        2 set by set_config."""
        out = (ctypes.c_int * 3)()
        check(_lib.lib().hph_batch_get_config(self._h, out, 3))
        return int(out[0]), int(out[1]), int(out[2])

    def timing(self):
        out = ctypes.c_double()
        check(_lib.lib().hph_batch_get_timing(self._h, ctypes.byref(out)))
        return out.value

    def last_split(self, dataset):
        """True when the dataset took the split passes in the last evaluation."""
        out = ctypes.c_int()
        check(_lib.lib().hph_batch_last_split(self._h, int(dataset), ctypes.byref(out)))
        return bool(out.value)
