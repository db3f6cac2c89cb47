"""hph_log_likelihood_noalloc: the allocation-free variant of hph_get_log_likelihood_batch, called
through the C ABI on device-resident inputs (torch tensors for the device memory).

Its kernels are the engine's row-wise pass with the decomposition an engine uses for a batch of K
vectors at target grid 4096, so the results equal the engine's batch results bit for bit; one
vector of each call is also checked against the oracle (1e-12, the FP64 parity bar). Two calls in
a row, the second smaller (fewer events, fewer vectors, another dimension), check that the shared
scratch array keeps nothing from the first call that the second would read.
"""
import ctypes

import numpy as np
import pytest

import oracle
from conftest import rel_err
from hawkes_amd import Engine, HphError, _lib, synth
from hawkes_amd._lib import check

pytestmark = pytest.mark.gpu


def noalloc(locs, times, P, flags=_lib.OPENCL, grad=True):
    import torch
    dev = torch.device("cuda:0")
    locs = np.ascontiguousarray(locs, dtype=np.float64)
    N, D = locs.shape
    P = np.ascontiguousarray(np.atleast_2d(P), dtype=np.float64)
    K = P.shape[0]
    l = torch.tensor(locs, device=dev)
    t = torch.tensor(np.ascontiguousarray(times, dtype=np.float64), device=dev)
    p = torch.tensor(P, device=dev)
    out = torch.full((K, 7), -1.0, dtype=torch.float64, device=dev)
    s = torch.cuda.current_stream()
    check(_lib.lib().hph_log_likelihood_noalloc(D, N, l.data_ptr(), t.data_ptr(), p.data_ptr(), K, flags,
                                                out.data_ptr(), 1 if grad else 0, ctypes.c_void_p(s.cuda_stream)))
    torch.cuda.synchronize()
    return out.cpu().numpy()


def engine_batch(locs, times, P, flags=_lib.OPENCL, grad=True):
    locs = np.asarray(locs, dtype=np.float64)
    e = Engine(locs.shape[1], locs.shape[0], flags=flags)
    e.set_pair_config(_lib.PAIR_KERNEL_ROWWISE, 2, 4096)
    e.update_locations(locs)
    e.set_times(times)
    e.set_parameters_batch(np.atleast_2d(P))
    r = e.log_likelihood_batch(gradient=grad)
    e.close()
    return r


def check_against_engine_and_oracle(locs, times, P, flags=_lib.OPENCL):
    out = noalloc(locs, times, P, flags=flags)
    ll, g = engine_batch(locs, times, P, flags=flags)
    print("N=%d D=%d K=%d: noalloc %r engine %r" % (len(times), locs.shape[1], len(P), out[0, 0], ll[0]), flush=True)
    assert np.array_equal(out[:, 0], ll) and np.array_equal(out[:, 1:], g)
    ref = oracle.total(locs, times, P[-1])
    assert rel_err(out[-1, 0], ref["ll"]) <= 1e-12
    gerr = np.abs(out[-1, 1:] - ref["grad"]) / np.maximum(np.abs(ref["grad"]), 1e-3 * ref["scale"][1:])
    assert gerr.max() <= 1e-12, gerr


def test_two_calls_in_a_row_match_the_batch_entry_point():
    locs, times = synth.cluster_process(4000, side=np.sqrt(5700 / 100), seed=301)
    check_against_engine_and_oracle(locs, times, synth.jitter_params(4, seed=302))
    rng = np.random.Generator(np.random.PCG64(303))
    n = 1200
    locs3 = rng.uniform(0.0, 3.0, size=(n, 3))
    times3 = np.sort(rng.uniform(0.0, 400.0, size=n))
    check_against_engine_and_oracle(locs3, times3, synth.jitter_params(2, base=[0.9, 0.7, 0.05, 0.4, 0.8, 0.6],
                                                                        seed=304))


def test_log_likelihood_only_and_fp32_bulk():
    locs, times = synth.cluster_process(2000, side=np.sqrt(2900 / 100), seed=305)
    P = synth.jitter_params(3, seed=306)
    ll_only = noalloc(locs, times, P, grad=False)[:, 0]
    assert np.array_equal(ll_only, engine_batch(locs, times, P, grad=False))
    out32 = noalloc(locs, times, P, flags=_lib.OPENCL | _lib.FLOAT)
    ll32, g32 = engine_batch(locs, times, P, flags=_lib.OPENCL | _lib.FLOAT)
    assert np.array_equal(out32[:, 0], ll32) and np.array_equal(out32[:, 1:], g32)


def test_rejects_what_the_array_cannot_hold_and_flags_bad_inputs():
    locs, times = synth.cluster_process(14000, seed=307)
    P = synth.jitter_params(200, seed=308)  # 200 x ~20k events x 6 partial sums: past 64 MiB
    with pytest.raises(HphError) as ex:
        noalloc(locs, times, P)
    assert ex.value.code == _lib.HPH_ERR_INVALID_ARGUMENT
    # times out of order: every result NaN (checked on the device), and the next call is clean
    small_locs, small_times = locs[:500], times[:500].copy()
    bad = small_times.copy()
    bad[100] = bad[101] + 1.0
    out = noalloc(small_locs, bad, P[:2])
    assert np.all(np.isnan(out))
    out = noalloc(small_locs, small_times, P[:2])
    ll, g = engine_batch(small_locs, small_times, P[:2])
    assert np.array_equal(out[:, 0], ll) and np.array_equal(out[:, 1:], g)
