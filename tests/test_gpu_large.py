"""N ~ 1e6 (cluster process with 700,000 background events over 43,800 h on 100 x 100 km): checks
through size-independent properties, where the oracle would take hours.

  * translation invariance: shifting every location leaves ll and the gradient unchanged
    (1e-12 relative; the pair terms depend on differences only);
  * permutation invariance: reordering the events while keeping the last one (T = last element)
    leaves them unchanged (1e-12 relative);
  * the gradient agrees with central differences of the GPU log-likelihood (h = 1e-6 relative);
  * probabilities lie in [0, 1], the earliest event's is 0;
  * FP32-bulk agrees with FP64 to 1e-6 (ll, gradient, probabilities).
"""
import numpy as np
import pytest

from hawkes_amd import Engine, _lib, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def large():
    locs, times, p = synth.workload("cluster_b700k")
    print("N =", len(times), flush=True)
    return locs, times, p


def engine(locs, times, p, flags=_lib.OPENCL):
    e = Engine(2, len(times), flags=flags)
    e.update_locations(locs)
    e.set_times(times)
    e.set_parameters(p)
    return e


def grad_close(a, b, tol):
    scale = np.abs(b).max()
    return np.all(np.abs(a - b) <= tol * np.maximum(np.abs(b), 1e-3 * scale))


def test_translation_and_permutation_invariance(large):
    locs, times, p = large
    e = engine(locs, times, p)
    ll, g = e.log_likelihood_and_gradient()
    e.update_locations(locs + np.array([123.5, -77.25]))
    ll_t, g_t = e.log_likelihood_and_gradient()
    assert abs(ll_t - ll) <= 1e-12 * abs(ll)
    assert grad_close(g_t, g, 1e-12)
    rng = np.random.default_rng(11)
    perm = np.concatenate([rng.permutation(len(times) - 1), [len(times) - 1]])
    e.update_locations(locs[perm])
    e.set_times(times[perm])
    ll_p, g_p = e.log_likelihood_and_gradient()
    assert abs(ll_p - ll) <= 1e-12 * abs(ll)
    assert grad_close(g_p, g, 1e-12)
    e.close()


def test_gradient_against_central_differences(large):
    locs, times, p = large
    e = engine(locs, times, p)
    _, g = e.log_likelihood_and_gradient()
    for k in (0, 2, 3):
        h = 1e-6 * p[k]
        up, dn = np.array(p, dtype=float), np.array(p, dtype=float)
        up[k] += h
        dn[k] -= h
        e.set_parameters(up)
        lu = e.log_likelihood()
        e.set_parameters(dn)
        ld = e.log_likelihood()
        fd = (lu - ld) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(abs(g[k]), 1e-3 * np.abs(g).max()), (k, fd, g[k])
    e.close()


def test_probabilities_and_fp32_bulk(large):
    locs, times, p = large
    e = engine(locs, times, p)
    ll, g = e.log_likelihood_and_gradient()
    probs = e.probs_self_excite()
    e.close()
    assert np.all((probs >= 0.0) & (probs <= 1.0))
    assert probs[np.argmin(times)] == 0.0
    f = engine(locs, times, p, flags=_lib.OPENCL | _lib.FLOAT)
    ll32, g32 = f.log_likelihood_and_gradient()
    probs32 = f.probs_self_excite()
    f.close()
    assert abs(ll32 - ll) <= 1e-6 * abs(ll)
    assert grad_close(g32, g, 1e-6)
    assert np.max(np.abs(probs32 - probs)) <= 1e-6
