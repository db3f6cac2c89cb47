"""MCMC drivers over the engine: the reference's adaptive Metropolis-Hastings sampler and an HMC
option on the exact gradient.

* ``adaptive_mh`` restates R/Untitled.R:254-422 (``sampler``): random scan over parameters
  sigmaXprec, omega, theta, mu0 (R's indices 1, 4:6); truncated-normal proposals at 0 with
  per-parameter radii; Hastings correction with the truncated-normal densities; radii adapted
  towards acceptance 0.44 with the ratio clipped to [0.5, 2] after a batch of SampBound
  proposals, SampBound <- ceiling(SampBound^1.1). Prior (``Potential``, R/Untitled.R:225-231):
  half-normal sd 10 for the first five parameters, half-normal sd 1 for mu0.
* ``hmc`` samples theta = log(params) with leapfrog trajectories on U(theta) = -loglik - log
  prior - sum(theta) (the Jacobian of the log transform), using the engine's gradient.
* ``hmc_chains`` runs K chains whose leapfrog steps are one batched launch of K gradients.

Random numbers come from numpy's Generator, so paths differ from R's; the algorithms do not.
"""
import time

import numpy as np
from scipy.special import log_ndtr, ndtr, ndtri

PRIOR_SD = np.array([10.0, 10.0, 10.0, 10.0, 10.0, 1.0])
MH_INDICES = (0, 3, 4, 5)
LOG_HALF_NORMAL_CONST = np.log(2.0) - 0.5 * np.log(2.0 * np.pi)


def log_prior(p):
    """Sum of log half-normal densities (truncnorm at 0, mean 0), sd PRIOR_SD; -inf off support."""
    p = np.asarray(p, dtype=np.float64)
    if np.any(p < 0):
        return -np.inf
    z = p / PRIOR_SD
    return float(np.sum(LOG_HALF_NORMAL_CONST - np.log(PRIOR_SD) - 0.5 * z * z))


def grad_log_prior(p):
    return -np.asarray(p, dtype=np.float64) / PRIOR_SD ** 2


def potential(engine, p):
    """-loglik - log prior (R/Untitled.R:225-231)."""
    lp = log_prior(p)
    if not np.isfinite(lp):
        return np.inf
    engine.set_parameters(p)
    return -engine.log_likelihood() - lp


def _truncnorm_logpdf(x, mean, sd):
    """log density of N(mean, sd^2) truncated to [0, inf) at x >= 0."""
    z = (x - mean) / sd
    return -0.5 * z * z - 0.5 * np.log(2.0 * np.pi) - np.log(sd) - log_ndtr(mean / sd)


def _truncnorm_draw(rng, mean, sd):
    lo = ndtr(-mean / sd)
    u = lo + rng.random() * (1.0 - lo)
    return mean + sd * ndtri(min(max(u, 1e-300), 1.0 - 1e-16))


def adaptive_mh(engine, n_iter, params, radius=2.0, burn_in=0, rng=None, indices=MH_INDICES):
    rng = np.random.default_rng(666) if rng is None else rng
    cur = np.array(params, dtype=np.float64)
    cur_u = potential(engine, cur)
    radii = np.full(6, float(radius))
    acc_batch = np.zeros(6)
    bound = np.full(6, 5.0)
    count = np.zeros(6)
    samples, target = [], []
    accepted = 0
    t0 = time.perf_counter()
    for it in range(1, n_iter + 1):
        idx = indices[rng.integers(len(indices))]
        prop = cur.copy()
        former = prop[idx]
        prop[idx] = _truncnorm_draw(rng, former, radii[idx])
        prop_u = potential(engine, prop)
        cur_h = cur_u - _truncnorm_logpdf(prop[idx], former, radii[idx])
        prop_h = prop_u - _truncnorm_logpdf(former, prop[idx], radii[idx])
        if -prop_h + cur_h > min(0.0, np.log(rng.random())):
            cur, cur_u = prop, prop_u
            accepted += 1
            acc_batch[idx] += 1
        count[idx] += 1
        if count[idx] == bound[idx]:
            ratio = min(max(acc_batch[idx] / bound[idx] / 0.44, 0.5), 2.0)
            radii[idx] *= ratio
            count[idx] = 0
            bound[idx] = np.ceil(bound[idx] ** 1.1)
            acc_batch[idx] = 0
        if it > burn_in:
            samples.append(cur.copy())
            target.append(cur_u)
    return {"samples": np.array(samples), "target": np.array(target), "radii": radii,
            "acceptance": accepted / n_iter, "time": time.perf_counter() - t0}


def _u_and_grad(engine, theta):
    """U(theta) = -loglik(e^theta) - log prior(e^theta) - sum(theta) and its gradient; U = inf
    (a rejected point) where the parameters are not finite."""
    with np.errstate(over="ignore"):
        p = np.exp(theta)
    if not np.all(np.isfinite(p)):
        return np.inf, np.zeros(6)
    engine.set_parameters(p)
    ll, g = engine.log_likelihood_and_gradient()
    u = -ll - log_prior(p) - float(np.sum(theta))
    du = -(g + grad_log_prior(p)) * p - 1.0
    return u, du


def leapfrog(engine, theta, momentum, step_size, n_steps, u_grad=None):
    """n_steps leapfrog steps (n_steps + 1 gradient evaluations when u_grad is not supplied).
    Returns (theta, momentum, U, dU) at the end point."""
    u, du = _u_and_grad(engine, theta) if u_grad is None else u_grad
    th = np.array(theta, dtype=np.float64)
    m = np.array(momentum, dtype=np.float64) - 0.5 * step_size * du
    for s in range(n_steps):
        th = th + step_size * m
        u, du = _u_and_grad(engine, th)
        if not np.isfinite(u):
            break
        m = m - (step_size if s < n_steps - 1 else 0.5 * step_size) * du
    return th, m, u, du


def hmc(engine, n_iter, params, step_size=0.01, n_leapfrog=20, burn_in=0, rng=None, adapt=True,
        target_accept=0.8):
    rng = np.random.default_rng(667) if rng is None else rng
    theta = np.log(np.asarray(params, dtype=np.float64))
    u, du = _u_and_grad(engine, theta)
    samples, target = [], []
    accepted = 0
    window = 0
    t0 = time.perf_counter()
    for it in range(1, n_iter + 1):
        m0 = rng.normal(size=6)
        th, m, u_new, du_new = leapfrog(engine, theta, m0, step_size, n_leapfrog, (u, du))
        h0 = u + 0.5 * m0 @ m0
        h1 = u_new + 0.5 * m @ m
        if np.isfinite(h1) and np.log(rng.random()) < h0 - h1:
            theta, u, du = th, u_new, du_new
            accepted += 1
            window += 1
        if adapt and it <= burn_in and it % 5 == 0:
            # the reference's (commented) step-size rule: ratio of the 5-step acceptance to the
            # target, clipped to [0.5, 2] (R/Untitled.R:401-408)
            step_size *= min(max((window / 5) / target_accept, 0.5), 2.0)
            window = 0
        if it > burn_in:
            samples.append(np.exp(theta))
            target.append(u)
    return {"samples": np.array(samples), "target": np.array(target), "step_size": step_size,
            "acceptance": accepted / n_iter, "time": time.perf_counter() - t0}


def _u_and_grad_batch(engine, thetas):
    with np.errstate(over="ignore"):
        p = np.exp(thetas)
    bad = ~np.all(np.isfinite(p), axis=1)
    p = np.where(bad[:, None], 1.0, p)
    engine.set_parameters_batch(p)
    ll, g = engine.log_likelihood_batch(gradient=True)
    lp = np.array([log_prior(q) for q in p])
    u = -ll - lp - thetas.sum(axis=1)
    du = -(g + grad_log_prior(p)) * p - 1.0
    u[bad] = np.inf
    du[bad] = 0.0
    return u, du


def hmc_chains(engine, params, n_iter, step_size=0.01, n_leapfrog=20, rng=None):
    """K chains (params: K x 6); each leapfrog step evaluates the K gradients in one launch."""
    rng = np.random.default_rng(668) if rng is None else rng
    theta = np.log(np.asarray(params, dtype=np.float64))
    K = theta.shape[0]
    u, du = _u_and_grad_batch(engine, theta)
    samples = []
    accepted = np.zeros(K)
    t0 = time.perf_counter()
    for _ in range(n_iter):
        m0 = rng.normal(size=(K, 6))
        th = theta.copy()
        m = m0 - 0.5 * step_size * du
        u_new, du_new = u, du
        for s in range(n_leapfrog):
            th = th + step_size * m
            u_new, du_new = _u_and_grad_batch(engine, th)
            m = m - (step_size if s < n_leapfrog - 1 else 0.5 * step_size) * du_new
        h0 = u + 0.5 * np.sum(m0 * m0, axis=1)
        h1 = u_new + 0.5 * np.sum(m * m, axis=1)
        ok = np.isfinite(h1) & (np.log(rng.random(K)) < h0 - h1)
        theta[ok], u[ok], du[ok] = th[ok], u_new[ok], du_new[ok]
        accepted += ok
        samples.append(np.exp(theta))
    return {"samples": np.array(samples), "acceptance": accepted / n_iter, "time": time.perf_counter() - t0}
