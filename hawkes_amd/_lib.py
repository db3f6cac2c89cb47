"""ctypes binding of libhph.so (the C ABI declared in include/hph.h).

The library is built in-tree (``make -C hawkes_amd/csrc`` or ``__graft_entry__.build()``).
There is no CPU fallback: if the library is missing or cannot load, every call raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhph.so")

HPH_OK = 0
HPH_ERR_INVALID_ARGUMENT = 1
HPH_ERR_NOT_INITIALIZED = 2
HPH_ERR_DEVICE = 3
HPH_ERR_OUT_OF_MEMORY = 4
HPH_ERR_UNSUPPORTED = 5

FLOAT = 1 << 2
TBB = 1 << 3
OPENCL = 1 << 4
SSE = 1 << 5
AVX = 1 << 6
AVX512 = 1 << 7
ALL_PAIRS = 1 << 16  # extension: no exact-zero tile skipping

PAIR_KERNEL_AUTO, PAIR_KERNEL_ROWWISE, PAIR_KERNEL_SYMMETRIC, PAIR_KERNEL_SPLIT = 0, 1, 2, 3
EVAL_FULL, EVAL_EXCITATION, EVAL_CACHED = 0, 1, 2

# Every symbol include/hph.h declares, with (restype, argtypes).
_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_sz = ctypes.c_size_t
SIGNATURES = {
    "hph_last_error": (ctypes.c_char_p, []),
    "hph_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_long, ctypes.c_int, ctypes.c_int,
                                  ctypes.POINTER(_vp)]),
    "hph_destroy": (ctypes.c_int, [_vp]),
    "hph_set_times_data": (ctypes.c_int, [_vp, _dp, _sz]),
    "hph_update_locations": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _sz]),
    "hph_set_parameters": (ctypes.c_int, [_vp, _dp, _sz]),
    "hph_get_sum_of_lik_contribs": (ctypes.c_int, [_vp, _dp]),
    "hph_get_log_like_gradient": (ctypes.c_int, [_vp, _dp, _sz]),
    "hph_get_log_likelihood_and_gradient": (ctypes.c_int, [_vp, _dp, _dp]),
    "hph_get_probs_self_excite": (ctypes.c_int, [_vp, _dp, _sz]),
    "hph_get_location_gradient": (ctypes.c_int, [_vp, _dp, _sz]),
    "hph_store_state": (ctypes.c_int, [_vp]),
    "hph_restore_state": (ctypes.c_int, [_vp]),
    "hph_accept_state": (ctypes.c_int, [_vp]),
    "hph_get_internal_dimension": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int)]),
    "hph_set_parameters_batch": (ctypes.c_int, [_vp, _dp, _sz]),
    "hph_get_log_likelihood_batch": (ctypes.c_int, [_vp, _dp, _dp, _sz]),
    "hph_set_timing": (ctypes.c_int, [_vp, ctypes.c_int]),
    "hph_get_timing": (ctypes.c_int, [_vp, _dp, _sz]),
    "hph_get_last_evaluation": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int)]),
    "hph_get_last_pass": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int), _sz]),
    "hph_get_event_sums": (ctypes.c_int, [_vp, _dp, _sz]),
    "hph_set_pair_kernel": (ctypes.c_int, [_vp, ctypes.c_int]),
    "hph_set_pair_config": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "hph_autotune": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "hph_get_tuned_config": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int), _sz]),
    "hph_batch_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_long,
                                        ctypes.c_int, ctypes.POINTER(_vp)]),
    "hph_batch_destroy": (ctypes.c_int, [_vp]),
    "hph_batch_set_data": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _sz, _dp, _sz]),
    "hph_batch_set_parameters": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int), _dp, _sz]),
    "hph_batch_get_log_likelihood": (ctypes.c_int, [_vp, _dp, _dp, _sz]),
    "hph_batch_set_graph": (ctypes.c_int, [_vp, ctypes.c_int]),
    "hph_batch_set_config": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "hph_batch_get_timing": (ctypes.c_int, [_vp, _dp]),
    "hph_batch_autotune": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "hph_batch_get_config": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int), _sz]),
    "hph_batch_last_split": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "hph_log_likelihood_noalloc": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, ctypes.c_int, ctypes.c_long,
                                                  _vp, ctypes.c_int, _vp]),
}

_lib = None


class HphError(RuntimeError):
    """A non-zero status from libhph; ``code`` is the HPH_ERR_* value."""

    def __init__(self, code, message):
        super().__init__("hph error %d: %s" % (code, message))
        self.code = code


def lib():
    """Loads libhph.so once; raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("hawkes_amd/libhph.so is not built; run `make -C hawkes_amd/csrc` "
                               "or __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != HPH_OK:
        msg = lib().hph_last_error()
        raise HphError(rc, msg.decode() if msg else "")
    return rc


def dptr(a):
    return a.ctypes.data_as(_dp)
