"""Dataset batches: one launch over many datasets of uneven size and their parameter vectors."""
import math

import numpy as np
import pytest

import oracle
from conftest import rel_err
from hawkes_amd import Engine, _lib, synth
from hawkes_amd.batch import Batch

pytestmark = pytest.mark.gpu


def uneven_datasets(count, lo, hi, seed=666000):
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for d in range(count):
        n_target = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        b = max(1, int(round(0.7 * n_target)))
        locs, times = synth.cluster_process(b, side=math.sqrt(n_target / 100.0), seed=seed + d)
        out.append((locs, times))
    return out


def make_batch(data, params, flags=_lib.OPENCL):
    b = Batch(2, [len(t) for _, t in data], flags=flags)
    for d, (locs, times) in enumerate(data):
        b.set_data(d, locs, times)
    b.set_parameters(params)
    return b


def test_batch_matches_oracle_per_dataset():
    data = uneven_datasets(6, 200, 3000, seed=11)
    params = [synth.jitter_params(3, seed=700 + d) for d in range(len(data))]
    b = make_batch(data, params)
    ll, g = b.log_likelihood()
    lls, gs = b.split(ll), b.split(g)
    for d, (locs, times) in enumerate(data):
        for k in range(params[d].shape[0]):
            ref = oracle.total(locs, times, params[d][k])
            assert rel_err(lls[d][k], ref["ll"]) <= 1e-12, (d, k)
            err = np.abs(gs[d][k] - ref["grad"]) / np.maximum(np.abs(ref["grad"]), 1e-3 * ref["scale"][1:])
            assert err.max() <= 1e-12, (d, k, err)


def test_batch_matches_single_engines():
    data = uneven_datasets(5, 2000, 20000, seed=21)
    params = [synth.jitter_params(2, seed=800 + d) for d in range(len(data))]
    b = make_batch(data, params)
    ll, g = b.log_likelihood()
    lls, gs = b.split(ll), b.split(g)
    for d, (locs, times) in enumerate(data):
        e = Engine(2, len(times))
        e.set_pair_kernel(_lib.PAIR_KERNEL_SYMMETRIC)
        e.update_locations(locs)
        e.set_times(times)
        for k in range(2):
            e.set_parameters(params[d][k])
            l1, g1 = e.log_likelihood_and_gradient()
            assert rel_err(lls[d][k], l1) <= 1e-13
            assert np.all(np.abs(gs[d][k] - g1) <= 1e-12 * np.maximum(np.abs(g1), 1.0))
        e.close()


def test_split_datasets_match_single_engines():
    """Datasets of at least 65,536 events spread over 60 km take the split passes inside the batch
    (as a single engine's built-in choice does); the others share the time-sorted pass. Each
    vector's result equals a single engine's (same passes, other work decomposition: rounding)
    in eager and graph mode, and graph replays equal the eager evaluation bitwise."""
    data = [synth.cluster_process(47000, side=60.0, seed=51), synth.cluster_process(1400, side=4.0, seed=52),
            synth.cluster_process(48000, side=60.0, seed=53), synth.cluster_process(3000, side=6.0, seed=54)]
    params = [synth.jitter_params(2, sd=0.05, seed=960 + d) for d in range(len(data))]
    b = make_batch(data, params)
    ll, g = b.log_likelihood()
    assert [b.last_split(d) for d in range(len(data))] == [True, False, True, False]
    lls, gs = b.split(ll), b.split(g)
    for d, (locs, times) in enumerate(data):
        e = Engine(2, len(times))
        e.update_locations(locs)
        e.set_times(times)
        for k in range(2):
            e.set_parameters(params[d][k])
            l1, g1 = e.log_likelihood_and_gradient()
            print(d, k, lls[d][k], l1, flush=True)
            assert rel_err(lls[d][k], l1) <= 1e-13, (d, k)
            assert np.all(np.abs(gs[d][k] - g1) <= 1e-12 * np.maximum(np.abs(g1), 1e-3 * np.abs(g1).max())), (d, k)
        e.close()
    b.set_graph(True)
    for rep in range(3):
        got = b.log_likelihood()
        assert np.array_equal(got[0], ll) and np.array_equal(got[1], g), rep
    # vectors that put a large dataset back on the time-sorted pass (a wide background kernel)
    wide = [p.copy() for p in params]
    wide[0][:, 1] = 0.05
    b.set_parameters(wide)
    got = b.log_likelihood()
    assert [b.last_split(d) for d in range(len(data))] == [False, False, True, False]
    b.set_graph(False)
    eager = b.log_likelihood()
    assert np.array_equal(got[0], eager[0]) and np.array_equal(got[1], eager[1])
    ref = oracle.total(data[1][0], data[1][1], wide[1][0])
    assert rel_err(got[0][2], ref["ll"]) <= 1e-12


def test_split_dataset_fp32_bulk_error():
    """FP32-bulk batch with a dataset on the split passes: every vector within 1e-6 of the FP64
    single-engine result (and the split dataset's FP32 error of the same order as the others)."""
    data = [synth.cluster_process(47000, side=60.0, seed=55), synth.cluster_process(2000, side=4.5, seed=56)]
    params = [synth.jitter_params(2, sd=0.05, seed=970 + d) for d in range(len(data))]
    b = make_batch(data, params, flags=_lib.OPENCL | _lib.FLOAT)
    ll, g = b.log_likelihood()
    assert b.last_split(0) and not b.last_split(1)
    lls = b.split(ll)
    for d, (locs, times) in enumerate(data):
        e = Engine(2, len(times))
        e.update_locations(locs)
        e.set_times(times)
        for k in range(2):
            e.set_parameters(params[d][k])
            l64 = e.log_likelihood()
            print(d, k, lls[d][k], l64, rel_err(lls[d][k], l64), flush=True)
            assert rel_err(lls[d][k], l64) <= 1e-6, (d, k)
        e.close()


def test_graph_replays_match_eager():
    data = uneven_datasets(4, 500, 8000, seed=31)
    params = [synth.jitter_params(2, seed=900 + d) for d in range(len(data))]
    b = make_batch(data, params)
    eager = b.log_likelihood()
    b.set_graph(True)
    locs, times = synth.cluster_process(560, seed=3)  # eager work between replays on another engine
    e_other = Engine(2, len(times))
    e_other.update_locations(locs)
    e_other.set_times(times)
    e_other.set_parameters(synth.P0)
    for rep in range(4):
        got = b.log_likelihood()
        assert np.array_equal(got[0], eager[0]) and np.array_equal(got[1], eager[1]), rep
        e_other.log_likelihood_and_gradient()
        e_other.set_parameters(synth.P0)
    # new parameter values replay the same graph
    params2 = [p * 1.01 for p in params]
    b.set_parameters(params2)
    g2 = b.log_likelihood()
    b.set_graph(False)
    eager2 = b.log_likelihood()
    assert np.array_equal(g2[0], eager2[0]) and np.array_equal(g2[1], eager2[1])
    assert not np.array_equal(g2[0], eager[0])


def test_graph_recaptures_when_exponent_range_changes():
    rng = np.random.default_rng(5)
    data = []
    for n in (700, 1300):
        data.append((rng.uniform(0, 3, (n, 2)), np.sort(rng.uniform(0, 1e5, n))))
    small = [np.array([[1.0, 0.9, 1e-4, 0.5, 0.8, 0.9]]), np.array([[1.1, 0.8, 1e-4, 0.6, 0.7, 1.0]])]
    huge = [np.array([[1.0, 0.9, 1.0, 0.5, 0.8, 0.9]]), np.array([[1.1, 0.8, 1.0, 0.6, 0.7, 1.0]])]
    b = make_batch(data, small)
    b.set_graph(True)
    b.log_likelihood()
    for params in (huge, small, huge):
        b.set_parameters(params)
        got = b.log_likelihood()
        for d, (locs, times) in enumerate(data):
            ref = oracle.total(locs, times, params[d][0])
            assert rel_err(got[0][d], ref["ll"]) <= 1e-12


def test_batch_fp32_bulk_error():
    data = uneven_datasets(4, 1000, 5000, seed=41)
    params = [synth.jitter_params(2, seed=950 + d) for d in range(len(data))]
    b = make_batch(data, params, flags=_lib.OPENCL | _lib.FLOAT)
    ll, g = b.log_likelihood()
    lls = b.split(ll)
    for d, (locs, times) in enumerate(data):
        for k in range(2):
            ref = oracle.total(locs, times, params[d][k], precision="double")
            assert rel_err(lls[d][k], ref["ll"]) <= 1e-6


def test_batch_errors():
    b = Batch(2, [10, 20])
    with pytest.raises(_lib.HphError) as ex:
        b.set_data(2, np.zeros((10, 2)), np.zeros(10))
    assert ex.value.code == _lib.HPH_ERR_INVALID_ARGUMENT
    with pytest.raises(_lib.HphError) as ex:
        b.set_data(0, np.zeros((9, 2)), np.zeros(9))
    assert ex.value.code == _lib.HPH_ERR_INVALID_ARGUMENT
    b.set_parameters([synth.jitter_params(1), synth.jitter_params(1)])
    with pytest.raises(_lib.HphError) as ex:
        b.log_likelihood()
    assert ex.value.code == _lib.HPH_ERR_NOT_INITIALIZED
    with pytest.raises(_lib.HphError):
        Batch(4, [10])
    with pytest.raises(_lib.HphError) as ex:
        b.last_split(2)
    assert ex.value.code == _lib.HPH_ERR_INVALID_ARGUMENT
    assert not b.last_split(0)  # nothing evaluated yet


def test_batch_autotune_picks_a_decomposition_for_the_shape():
    """hph_batch_autotune times the work decompositions on the table and keeps the fastest for its
    shape: the batch (and a second batch of the same shape) then report the tuned configuration,
    results agree with the built-in decomposition's to rounding, and hph_batch_set_config
    overrides the tuned choice."""
    data = uneven_datasets(4, 2000, 12000, seed=41)
    params = [synth.jitter_params(2, seed=820 + d) for d in range(len(data))]
    b = make_batch(data, params)
    assert b.config() == (2, 8192, 0)
    ll0, g0 = b.log_likelihood()
    ms = b.autotune(repeats=1)
    assert ms > 0
    A, target, source = b.config()
    assert source == 1 and A in (1, 2) and target in (2048, 4096, 8192, 16384, 32768, 65536)
    b.set_graph(True)
    ll1, g1 = b.log_likelihood()
    assert np.all(rel_err(ll1, ll0) <= 1e-13)
    assert np.all(np.abs(g1 - g0) <= 1e-12 * np.maximum(np.abs(g0), 1.0))
    b2 = make_batch(data, params)
    assert b2.config() == (A, target, 1)
    b2.set_config(1, 1024)
    assert b2.config() == (1, 1024, 2)
    ll2, _ = b2.log_likelihood()
    assert np.all(rel_err(ll2, ll0) <= 1e-13)
    b.close()
    b2.close()
