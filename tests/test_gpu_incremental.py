"""Re-evaluation from cached per-event moments (the sampler's proposals change sigma, omega, theta
or mu0 only, R/Untitled.R:254-422).

Every result of an engine that walks a chain of such proposals is compared with a fresh engine's
full evaluation at the same parameters: ll to 1e-14 relative, gradient component k to 1e-13 of
max(|g_k|, 1e-3 max|g|). Background sums reused across the chain are the ones a full pass forms
for the unchanged tauX and tauT; excitation sums come from an excitation-only pass or, for theta
and mu0 moves, are reused within their skip bound (every rate moves by under 4e-22 relative).
A change of tauX, tauT or of the data and HPH_FLAG_ALL_PAIRS take the full pass. A probability
request at the cached tauX, tauT reuses the background sums and forms the excitation sums anew
(exact zeros skipped only); compared with a fresh engine's full probability pass to 1e-13 relative
(the excitation sums are added in another order) and with the oracle to 1e-12.
"""
import numpy as np
import pytest

import oracle
from conftest import rel_err
from hawkes_amd import Engine, _lib, sampler, synth

pytestmark = pytest.mark.gpu

CONFIGS = {"auto": None, "symmetric": (_lib.PAIR_KERNEL_SYMMETRIC, 2, 4096),
           "rowwise_narrow": (_lib.PAIR_KERNEL_ROWWISE, 2, 65536), "split": (_lib.PAIR_KERNEL_SPLIT, 2, 4096)}


def engine(locs, times, params, config=None, flags=_lib.OPENCL):
    locs = np.asarray(locs, dtype=np.float64)
    e = Engine(locs.shape[1], locs.shape[0], flags=flags)
    if config is not None:
        e.set_pair_config(*config)
    e.update_locations(locs)
    e.set_times(times)
    e.set_parameters(params)
    return e


def fresh(locs, times, params, config, flags, grad):
    e = engine(locs, times, params, config, flags)
    out = e.log_likelihood_and_gradient() if grad else (e.log_likelihood(), None)
    e.close()
    return out


def close(a, b, grad):
    assert rel_err(a[0], b[0]) <= 1e-14, (a[0], b[0])
    if grad:
        g, h = np.asarray(a[1]), np.asarray(b[1])
        assert np.all(np.abs(g - h) <= 1e-13 * np.maximum(np.abs(h), 1e-3 * np.abs(h).max())), (g, h)


def chain(p0, steps, seed):
    """MH-style moves: one of sigma, omega, theta, mu0 at a time; (parameters, moved index)."""
    rng = np.random.default_rng(seed)
    out, cur = [], np.array(p0, dtype=float)
    for _ in range(steps):
        idx = sampler.MH_INDICES[rng.integers(len(sampler.MH_INDICES))]
        cur = cur.copy()
        cur[idx] *= np.exp(0.2 * rng.normal())
        out.append((cur, idx))
    return out


@pytest.mark.parametrize("config", ["auto", "symmetric", "rowwise_narrow", "split"])
@pytest.mark.parametrize("grad", [False, True])
def test_chain_matches_full_evaluations(config, grad):
    locs, times = synth.cluster_process(3000, side=15.0, seed=12)
    p0 = synth.P0
    e = engine(locs, times, p0, CONFIGS[config])
    e.log_likelihood_and_gradient() if grad else e.log_likelihood()
    assert e.last_evaluation() == _lib.EVAL_FULL
    kinds = []
    for k, (p, idx) in enumerate(chain(p0, 16, 5 if grad else 6)):
        e.set_parameters(p)
        got = e.log_likelihood_and_gradient() if grad else (e.log_likelihood(), None)
        kind = e.last_evaluation()
        print(config, grad, k, idx, kind, got[0], flush=True)
        # sigma and omega moves need the excitation pass; theta and mu0 moves reuse both sums
        # unless the excitation skip point moved by more than ln 2 since they were formed
        if idx in (0, 3):
            assert kind == _lib.EVAL_EXCITATION
        else:
            assert kind in (_lib.EVAL_CACHED, _lib.EVAL_EXCITATION)
        kinds.append(kind)
        close(got, fresh(locs, times, p, CONFIGS[config], _lib.OPENCL, grad), grad)
    assert _lib.EVAL_CACHED in kinds
    e.close()


def test_other_changes_take_the_full_pass():
    locs, times = synth.cluster_process(2500, side=10.0, seed=13)
    p = synth.P0.copy()
    e = engine(locs, times, p)
    e.log_likelihood_and_gradient()
    # tauT move
    p2 = p.copy()
    p2[2] *= 1.3
    e.set_parameters(p2)
    close(e.log_likelihood_and_gradient(), fresh(locs, times, p2, None, _lib.OPENCL, True), True)
    assert e.last_evaluation() == _lib.EVAL_FULL
    # a single location moves
    l2 = locs.copy()
    l2[100] += 0.5
    e.update_locations(l2[100], index=100)
    close(e.log_likelihood_and_gradient(), fresh(l2, times, p2, None, _lib.OPENCL, True), True)
    assert e.last_evaluation() == _lib.EVAL_FULL
    # theta move after the data change, then an LL-only sigma move, then the gradient again
    p3 = p2.copy()
    p3[4] *= 0.8
    e.set_parameters(p3)
    close(e.log_likelihood_and_gradient(), fresh(l2, times, p3, None, _lib.OPENCL, True), True)
    assert e.last_evaluation() == _lib.EVAL_CACHED
    p4 = p3.copy()
    p4[0] *= 1.1
    e.set_parameters(p4)
    close((e.log_likelihood(), None), fresh(l2, times, p4, None, _lib.OPENCL, False), False)
    assert e.last_evaluation() == _lib.EVAL_EXCITATION
    # the LL-only call kept no gradient moments: the gradient takes the full pass
    e.set_parameters(p4)
    close(e.log_likelihood_and_gradient(), fresh(l2, times, p4, None, _lib.OPENCL, True), True)
    assert e.last_evaluation() == _lib.EVAL_FULL
    # probabilities after a chain step equal a fresh engine's (from the cached background sums
    # here, a full probability pass there)
    ref = engine(l2, times, p4)
    np.testing.assert_allclose(e.probs_self_excite(), ref.probs_self_excite(), rtol=1e-13, atol=1e-300)
    assert e.last_evaluation() == _lib.EVAL_EXCITATION
    ref.close()
    e.close()


@pytest.mark.parametrize("config", ["auto", "symmetric", "split"])
@pytest.mark.parametrize("flags", [_lib.OPENCL, _lib.OPENCL | _lib.FLOAT])
def test_probabilities_from_cached_sums(config, flags):
    """A probability request after a likelihood call at the same tauX, tauT: background sums from
    the cache, excitation sums from the band kernel (exact zeros skipped only); the cache, the
    log-likelihood and 1/lambda are left as they were."""
    locs, times = synth.cluster_process(3000, side=np.sqrt(3000 / 100) * (3.0 if config == "split" else 1.0),
                                        seed=31)
    p = np.array([1.05, 0.6, 1 / 320.0, 0.95, 0.9, 1.1])
    e = engine(locs, times, p, CONFIGS[config], flags)
    ll1 = e.log_likelihood_and_gradient()
    probs = e.probs_self_excite()
    assert e.last_evaluation() == _lib.EVAL_EXCITATION
    ref = engine(locs, times, p, CONFIGS[config], flags)
    full = ref.probs_self_excite()  # no cache: the full probability pass
    assert ref.last_evaluation() == _lib.EVAL_FULL
    ref.close()
    np.testing.assert_allclose(probs, full, rtol=1e-13, atol=1e-300)
    if flags == _lib.OPENCL:
        o = oracle.total(locs, times, p, probs=True)
        np.testing.assert_allclose(probs, o["probs"], rtol=1e-12, atol=1e-300)
    # the likelihood cache is untouched: the same call re-evaluates to the same bits
    e.set_parameters(p)
    ll2 = e.log_likelihood_and_gradient()
    assert e.last_evaluation() == _lib.EVAL_CACHED
    assert ll1[0] == ll2[0] and np.array_equal(ll1[1], ll2[1])
    # after a sampler move the request uses the moved sigma, omega
    q = p.copy()
    q[0] *= 1.1
    q[3] *= 0.9
    e.set_parameters(q)
    e.log_likelihood()
    pq = e.probs_self_excite()
    ref = engine(locs, times, q, CONFIGS[config], flags)
    np.testing.assert_allclose(pq, ref.probs_self_excite(), rtol=1e-13, atol=1e-300)
    ref.close()
    e.close()


def test_fp32_bulk_and_all_pairs():
    locs, times = synth.cluster_process(2500, side=10.0, seed=14)
    for flags in (_lib.OPENCL | _lib.FLOAT, _lib.OPENCL | _lib.ALL_PAIRS):
        e = engine(locs, times, synth.P0, flags=flags)
        e.log_likelihood_and_gradient()
        for p, idx in chain(synth.P0, 6, 9):
            e.set_parameters(p)
            close(e.log_likelihood_and_gradient(), fresh(locs, times, p, None, flags, True), True)
            assert (e.last_evaluation() == _lib.EVAL_FULL) == bool(flags & _lib.ALL_PAIRS)
        e.close()


def test_large_chain_and_location_gradient():
    locs, times, p0 = synth.workload("cluster_b35k")
    e = engine(locs, times, p0)
    e.log_likelihood()
    for p, idx in chain(p0, 4, 21):
        e.set_parameters(p)
        close((e.log_likelihood(), None), fresh(locs, times, p, None, _lib.OPENCL, False), False)
        assert e.last_evaluation() != _lib.EVAL_FULL
    # the location gradient needs natural-form rates from a full pass: equal to a fresh engine's
    g = e.location_gradient()
    ref = engine(locs, times, p)
    g_ref = ref.location_gradient()
    ref.close()
    e.close()
    assert np.all(np.abs(g - g_ref) <= 1e-13 * np.maximum(np.abs(g_ref), 1e-3 * np.abs(g_ref).max()))


@pytest.mark.parametrize("config", ["auto", "symmetric", "rowwise_narrow", "split"])
@pytest.mark.parametrize("flags", [_lib.OPENCL, _lib.OPENCL | _lib.FLOAT], ids=["fp64", "fp32_bulk"])
def test_shared_tau_batch_matches_full_evaluations(config, flags):
    """A batch whose vectors share tauX and tauT forms the background sums once (a full pass for
    vector 0, or the single-vector cache) and one excitation pass per vector: each result equals a
    fresh engine's full evaluation (ll 1e-14, gradient 1e-13 as above), with and without the
    gradient, and the single-vector cache it refilled serves a later single call."""
    import oracle
    locs, times = synth.cluster_process(3000, side=15.0, seed=15)
    P = synth.jitter_params(6, seed=16)
    P[:, 1:3] = synth.P0[1:3]
    e = engine(locs, times, synth.P0, CONFIGS[config], flags)
    e.set_parameters_batch(P)
    for grad in (True, False, True):
        r = e.log_likelihood_batch(gradient=grad)
        assert e.last_evaluation() == _lib.EVAL_EXCITATION
        ll, g = (r if grad else (r, None))
        for k in range(len(P)):
            close((ll[k], g[k] if grad else None), fresh(locs, times, P[k], CONFIGS[config], flags, grad), grad)
    if not flags & _lib.FLOAT:
        ref = oracle.total(locs, times, P[2])
        assert rel_err(ll[2], ref["ll"]) <= 1e-12
        err = np.abs(g[2] - ref["grad"]) / np.maximum(np.abs(ref["grad"]), 1e-3 * ref["scale"][1:])
        assert err.max() <= 1e-12, err
    # the cache now holds vector 0's sums: a single call at P[0] reuses all of them
    e.set_parameters(P[0])
    close(e.log_likelihood_and_gradient(), fresh(locs, times, P[0], CONFIGS[config], flags, True), True)
    assert e.last_evaluation() == _lib.EVAL_CACHED
    # a batch with one tauX apart takes the K-vector pass
    Q = P.copy()
    Q[3, 1] *= 1.01
    e.set_parameters_batch(Q)
    ll, g = e.log_likelihood_batch()
    assert e.last_evaluation() == _lib.EVAL_FULL
    close((ll[3], g[3]), fresh(locs, times, Q[3], CONFIGS[config], flags, True), True)
    e.close()
