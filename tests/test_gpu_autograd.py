"""torch.autograd wrapper and MCMC drivers on the HIP engine."""
import numpy as np
import pytest
import torch

import oracle
from hawkes_amd import Engine, sampler, synth
from hawkes_amd.autograd import loglik, loglik_batch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def data():
    return synth.cluster_process(1400, seed=51)


def engine_for(locs, times):
    e = Engine(2, len(times))
    e.update_locations(locs)
    e.set_times(times)
    return e


def test_autograd_gradient_matches_oracle(data):
    locs, times = data
    e = engine_for(locs, times)
    p = torch.tensor(synth.P0, dtype=torch.float64, requires_grad=True)
    ll = loglik(p, e)
    ll.backward()
    ref = oracle.total(locs, times, synth.P0)
    assert abs(ll.item() - ref["ll"]) <= 1e-12 * abs(ref["ll"])
    err = np.abs(p.grad.numpy() - ref["grad"]) / np.maximum(np.abs(ref["grad"]), 1e-3 * ref["scale"][1:])
    assert err.max() <= 1e-12


def test_gradcheck(data):
    locs, times = data
    e = engine_for(locs, times)
    p = torch.tensor([1.1, 0.7, 0.004, 0.9, 0.8, 1.2], dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(lambda q: loglik(q, e), (p,), eps=1e-6, atol=1e-6, rtol=1e-5)


def test_batched_autograd(data):
    locs, times = data
    e = engine_for(locs, times)
    P = torch.tensor(synth.jitter_params(5), dtype=torch.float64, requires_grad=True)
    ll = loglik_batch(P, e)
    (ll * torch.arange(1, 6, dtype=torch.float64)).sum().backward()
    for k in range(5):
        ref = oracle.total(locs, times, P[k].detach().numpy())
        assert abs(ll[k].item() - ref["ll"]) <= 1e-12 * abs(ref["ll"])
        np.testing.assert_allclose(P.grad[k].numpy(), (k + 1) * ref["grad"], rtol=1e-10,
                                   atol=1e-10 * np.max(np.abs(ref["grad"])))


def test_mh_and_hmc_on_engine(data):
    locs, times = data
    e = engine_for(locs, times)
    mh = sampler.adaptive_mh(e, 200, synth.P0, rng=np.random.default_rng(1))
    assert 0.0 < mh["acceptance"] < 1.0
    start = mh["samples"][-1]
    # leapfrog with a small step conserves H = U + |m|^2 / 2 (second-order integrator)
    theta = np.log(start)
    m0 = np.random.default_rng(9).normal(size=6)
    u0, du0 = sampler._u_and_grad(e, theta)
    th, m, u1, _ = sampler.leapfrog(e, theta, m0, 1e-5, 10, (u0, du0))
    assert abs((u1 + 0.5 * m @ m) - (u0 + 0.5 * m0 @ m0)) < 1e-3
    h = sampler.hmc(e, 30, start, step_size=1e-3, n_leapfrog=10, burn_in=15, rng=np.random.default_rng(2))
    assert h["acceptance"] > 0.0 and np.all(np.isfinite(h["target"]))
    chains = sampler.hmc_chains(e, synth.jitter_params(8), 3, step_size=0.005, n_leapfrog=5,
                                rng=np.random.default_rng(3))
    assert chains["samples"].shape == (3, 8, 6)
