"""The MCMC drivers on the HIP engine reproduce the chains they run on the oracle.

The same seeded adaptive Metropolis-Hastings run (the random-scan sampler of
R/Untitled.R:254-422: sigmaXprec, omega, theta, mu0 proposed one at a time, tauX and tauT fixed)
and the same seeded HMC run, once with the HIP engine and once with the CPU oracle standing in
for the engine (conftest.OracleEngine). The proposals come from the random stream alone, so the
chains are identical exactly when every accept/reject decision agrees: the samples and adapted
radii must be equal bit for bit, and the log-likelihood values along the chain equal to 1e-12.
On the engine, the MH moves go through the re-evaluation from cached per-event sums (sigma and
omega moves: cached background sums + excitation pass; theta and mu0 moves: cached sums only),
which the test asserts it exercised. HMC trajectories use the exact gradient; their end points
agree to 1e-9 relative (ten leapfrog steps amplify 1e-12 differences of the gradient).
"""
import numpy as np
import pytest

from conftest import OracleEngine, rel_err
from hawkes_amd import Engine, _lib, sampler, synth

pytestmark = pytest.mark.gpu


class CountingEngine:
    """The HIP engine, recording how each single-vector call was formed."""

    def __init__(self, locs, times):
        self.e = Engine(2, len(times))
        self.e.update_locations(locs)
        self.e.set_times(times)
        self.kinds = []

    def set_parameters(self, p):
        self.e.set_parameters(p)

    def log_likelihood(self):
        ll = self.e.log_likelihood()
        self.kinds.append(self.e.last_evaluation())
        return ll

    def log_likelihood_and_gradient(self):
        r = self.e.log_likelihood_and_gradient()
        self.kinds.append(self.e.last_evaluation())
        return r


@pytest.fixture(scope="module")
def data():
    return synth.cluster_process(700, side=np.sqrt(1000 / 100), seed=661)


def test_adaptive_mh_chain_matches_oracle_chain(data):
    locs, times = data
    gpu, cpu = CountingEngine(locs, times), OracleEngine(locs, times)
    runs = [sampler.adaptive_mh(eng, 50, synth.P0, rng=np.random.default_rng(2024)) for eng in (gpu, cpu)]
    g, c = runs
    print("N=%d acceptance gpu %.2f cpu %.2f; evaluation kinds %s" % (
        len(times), g["acceptance"], c["acceptance"], np.bincount(gpu.kinds, minlength=3)), flush=True)
    assert 0.0 < c["acceptance"] < 1.0
    assert np.array_equal(g["samples"], c["samples"])
    assert np.array_equal(g["radii"], c["radii"]) and g["acceptance"] == c["acceptance"]
    assert np.all(rel_err(g["target"], c["target"]) <= 1e-12)
    # both re-evaluation kinds ran (random scan over sigma, omega, theta, mu0)
    assert _lib.EVAL_EXCITATION in gpu.kinds and _lib.EVAL_CACHED in gpu.kinds


def test_hmc_chain_matches_oracle_chain(data):
    locs, times = data
    gpu, cpu = CountingEngine(locs, times), OracleEngine(locs, times)
    start = np.array([1.05, 0.6, 1 / 320.0, 0.95, 0.9, 1.1])
    runs = [sampler.hmc(eng, 3, start, step_size=2e-3, n_leapfrog=10, rng=np.random.default_rng(2025), adapt=False)
            for eng in (gpu, cpu)]
    g, c = runs
    print("HMC acceptance gpu %.2f cpu %.2f" % (g["acceptance"], c["acceptance"]), flush=True)
    assert g["acceptance"] == c["acceptance"] and c["acceptance"] > 0.0
    np.testing.assert_allclose(g["samples"], c["samples"], rtol=1e-9, atol=0)
    np.testing.assert_allclose(g["target"], c["target"], rtol=1e-9, atol=0)
    # every HMC step changes all six parameters: full passes only
    assert all(k == _lib.EVAL_FULL for k in gpu.kinds)
