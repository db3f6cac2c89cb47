"""MCMC driver logic on the CPU, with the oracle standing in for the engine (test only)."""
import numpy as np
import pytest
from scipy.stats import halfnorm, truncnorm

from conftest import OracleEngine
from hawkes_amd import sampler, synth


@pytest.fixture(scope="module")
def small_engine():
    locs, times = synth.cluster_process(60, side=3.0, horizon=500.0, seed=17)
    return OracleEngine(locs, times)


def test_log_prior_is_half_normal():
    p = np.array([0.7, 1.3, 0.02, 2.0, 0.4, 0.9])
    expect = sum(halfnorm.logpdf(p[k], scale=sampler.PRIOR_SD[k]) for k in range(6))
    assert sampler.log_prior(p) == pytest.approx(expect, rel=1e-13)
    assert sampler.log_prior(-p) == -np.inf
    h = 1e-6
    for k in range(6):
        q = p.copy()
        q[k] += h
        r = p.copy()
        r[k] -= h
        fd = (sampler.log_prior(q) - sampler.log_prior(r)) / (2 * h)
        assert fd == pytest.approx(sampler.grad_log_prior(p)[k], rel=1e-6, abs=1e-9)


def test_truncnorm_density_and_draws():
    for mean, sd in ((0.5, 2.0), (3.0, 0.5), (-1.0, 1.0)):
        a = (0 - mean) / sd
        for x in (0.01, 0.5, 2.0, 7.0):
            assert sampler._truncnorm_logpdf(x, mean, sd) == pytest.approx(
                truncnorm.logpdf(x, a, np.inf, loc=mean, scale=sd), rel=1e-12, abs=1e-12)
        rng = np.random.default_rng(1)
        xs = np.array([sampler._truncnorm_draw(rng, mean, sd) for _ in range(4000)])
        assert xs.min() >= 0
        assert abs(xs.mean() - truncnorm.mean(a, np.inf, loc=mean, scale=sd)) < 0.06 * sd


def test_potential_gradient_in_log_space(small_engine):
    theta = np.log(np.array([1.1, 0.8, 0.01, 0.9, 0.5, 0.7]))
    u, du = sampler._u_and_grad(small_engine, theta)
    h = 1e-6
    for k in range(6):
        tp = theta.copy()
        tp[k] += h
        tm = theta.copy()
        tm[k] -= h
        fd = (sampler._u_and_grad(small_engine, tp)[0] - sampler._u_and_grad(small_engine, tm)[0]) / (2 * h)
        assert fd == pytest.approx(du[k], rel=1e-5, abs=1e-6)


def test_leapfrog_is_reversible_and_conserves_energy(small_engine):
    theta = np.log(np.array([1.1, 0.8, 0.01, 0.9, 0.5, 0.7]))
    m = np.array([0.3, -0.2, 0.1, 0.4, -0.5, 0.2])
    u0, _ = sampler._u_and_grad(small_engine, theta)
    th1, m1, u1, _ = sampler.leapfrog(small_engine, theta, m, 1e-3, 10)
    th2, m2, _, _ = sampler.leapfrog(small_engine, th1, -m1, 1e-3, 10)
    np.testing.assert_allclose(th2, theta, atol=1e-10)
    np.testing.assert_allclose(-m2, m, atol=1e-9)
    h0 = u0 + 0.5 * m @ m
    h1 = u1 + 0.5 * m1 @ m1
    assert abs(h1 - h0) < 1e-3


def test_adaptive_mh_runs_and_adapts(small_engine):
    out = sampler.adaptive_mh(small_engine, 120, [1.0, 1 / 1.6, 1 / 336.0, 1.0, 1.0, 1.0], rng=np.random.default_rng(3))
    assert out["samples"].shape == (120, 6)
    assert 0.0 < out["acceptance"] < 1.0
    assert np.all(out["samples"] > 0)
    assert not np.allclose(out["radii"], 2.0)
    # parameters outside the random scan (tauX, tauT) never move
    assert np.all(out["samples"][:, 1] == 1 / 1.6) and np.all(out["samples"][:, 2] == 1 / 336.0)


def test_hmc_runs(small_engine):
    out = sampler.hmc(small_engine, 8, [1.0, 0.8, 0.01, 1.0, 0.5, 0.7], step_size=0.01, n_leapfrog=5,
                      rng=np.random.default_rng(4), adapt=False)
    assert out["samples"].shape == (8, 6)
    assert out["acceptance"] > 0.5


def test_hmc_chains_batched(small_engine):
    P = np.array([[1.0, 0.8, 0.01, 1.0, 0.5, 0.7], [0.9, 0.7, 0.012, 1.1, 0.6, 0.8]])
    out = sampler.hmc_chains(small_engine, P, 3, step_size=0.01, n_leapfrog=3, rng=np.random.default_rng(5))
    assert out["samples"].shape == (3, 2, 6)
    assert np.all(out["acceptance"] >= 0)
