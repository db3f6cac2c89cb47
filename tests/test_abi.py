"""The C ABI library loads and exports every symbol include/hph.h declares (no GPU needed),
and host-side argument checks return the documented status codes."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "hph.h")
LIB = os.path.join(ROOT, "hawkes_amd", "libhph.so")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hph_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_engine_api():
    syms = declared_symbols()
    for name in ("hph_create", "hph_destroy", "hph_set_times_data", "hph_update_locations",
                 "hph_set_parameters", "hph_get_sum_of_lik_contribs", "hph_get_log_like_gradient",
                 "hph_get_probs_self_excite", "hph_store_state", "hph_restore_state", "hph_accept_state",
                 "hph_get_internal_dimension", "hph_last_error"):
        assert name in syms


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "libhph.so not built"
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing


def test_python_binding_covers_header():
    from hawkes_amd import _lib
    assert sorted(_lib.SIGNATURES) == declared_symbols()
    L = _lib.lib()
    for name in declared_symbols():
        assert getattr(L, name) is not None


def test_flag_values_match_reference():
    # src/flags.h:6-13
    from hawkes_amd import _lib
    assert (_lib.FLOAT, _lib.TBB, _lib.OPENCL, _lib.SSE, _lib.AVX, _lib.AVX512) == (4, 8, 16, 32, 64, 128)
    text = open(HEADER).read()
    for name, v in (("FLOAT", 2), ("TBB", 3), ("OPENCL", 4), ("SSE", 5), ("AVX", 6), ("AVX512", 7)):
        assert re.search(r"#define HPH_FLAG_%s \(1L << %d\)" % (name, v), text)


def test_null_handles_are_rejected():
    from hawkes_amd import _lib
    L = _lib.lib()
    x = np.zeros(4)
    assert L.hph_set_times_data(None, _lib.dptr(x), 4) == _lib.HPH_ERR_INVALID_ARGUMENT
    assert L.hph_last_error()
    out = ctypes.c_double()
    assert L.hph_get_sum_of_lik_contribs(None, ctypes.byref(out)) == _lib.HPH_ERR_INVALID_ARGUMENT
    assert L.hph_destroy(None) == _lib.HPH_OK
    assert L.hph_create(2, 10, _lib.OPENCL, 0, 0, None) == _lib.HPH_ERR_INVALID_ARGUMENT


def test_create_validates_dimension_before_touching_a_device():
    from hawkes_amd import _lib
    L = _lib.lib()
    h = ctypes.c_void_p()
    assert L.hph_create(0, 10, _lib.OPENCL, 0, 0, ctypes.byref(h)) == _lib.HPH_ERR_INVALID_ARGUMENT
    assert L.hph_create(9, 10, _lib.OPENCL, 0, 0, ctypes.byref(h)) == _lib.HPH_ERR_INVALID_ARGUMENT
    assert L.hph_create(2, 0, _lib.OPENCL, 0, 0, ctypes.byref(h)) == _lib.HPH_ERR_INVALID_ARGUMENT
    assert not h.value


def test_create_without_device_reports_device_error():
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    from hawkes_amd import _lib
    L = _lib.lib()
    h = ctypes.c_void_p()
    assert L.hph_create(2, 10, _lib.OPENCL, 0, 0, ctypes.byref(h)) == _lib.HPH_ERR_DEVICE
    assert b"device" in L.hph_last_error().lower()
    assert not h.value


def test_cpp_caller_compiles_against_the_engine_header(tmp_path):
    """A C++ caller of the reference's interface (factory, setTimesData, updateLocations,
    setParameters, getSumOfLikContribs, getProbsSelfExcite) builds against include/hph_engine.hpp
    and links against libhph.so with the host compiler alone."""
    assert os.path.exists(LIB), "libhph.so not built"
    out = str(tmp_path / "dropin")
    subprocess.run(["bash", os.path.join(ROOT, "tests", "cpp", "build.sh"), out], check=True)
    assert os.path.exists(out)


def test_engine_subclass_of_the_reference_base_class_compiles(tmp_path):
    """An engine subclass written against src/AbstractHawkes.hpp:32-60 (the nine pure virtuals and
    the protected members) compiles against include/hph_engine.hpp unchanged and runs without a
    GPU; observationCount = N (N - 1) / 2 is 64-bit (4,999,950,000 at N = 100,000)."""
    out = str(tmp_path / "subclass")
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "subclass_main.cpp"), "-o", out], check=True)
    run = subprocess.run([out], capture_output=True, text=True, check=True)
    assert run.stdout.split() == ["2", "100000", "4999950000", "16", "7"]


def test_base_class_has_exactly_the_reference_virtuals():
    """hph::AbstractHawkes declares the nine pure virtuals of src/AbstractHawkes.hpp:42-51 and no
    others; the extensions live in hph::HawkesEngine."""
    text = open(os.path.join(ROOT, "include", "hph_engine.hpp")).read()
    body = text[text.index("class AbstractHawkes {"):text.index("typedef std::shared_ptr<hph::AbstractHawkes>")]
    virtuals = re.findall(r"virtual\s+[\w\s\*]+?\b(\w+)\s*\(", body)
    assert sorted(v for v in virtuals if v != "AbstractHawkes") == sorted([
        "getSumOfLikContribs", "updateLocations", "getProbsSelfExcite", "storeState", "restoreState",
        "acceptState", "setTimesData", "setParameters", "getInternalDimension"])
    for member in ("int embeddingDimension;", "int locationCount;", "long long observationCount;", "long flags;",
                   "int updatedLocation = -1;"):
        assert member in body
