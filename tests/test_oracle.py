"""Pins the CPU oracle (oracle/) before it is trusted as the checker of the HIP path.

Each oracle function is checked against something a plausible mistake would fail:
  * golden vectors: 50-digit mpmath evaluation of the independent R statement
    (R/log_lik_and_grad.R:2-93, written in the sd parametrisation with full normalisers);
  * closed forms: N = 1, theta = 0;
  * invariants: spatial translation, joint permutation keeping the last event, scaling;
  * gradient: central finite differences, and the R analytic gradients for mu0, tau_x,
    tau_t (R/log_lik_and_grad.R:176-247) via the chain rule d/dprec = -prec^-2 d/dsd,
    theta and h at omega = 1 (where the R formulas are valid);
  * the CPU baseline flavours (serial, AVX, threaded) against the oracle;
  * location gradient: mpmath.diff golden vectors, translation and rotation invariance,
    central finite differences.
"""
import math

import numpy as np
import pytest
from scipy.special import erfc
from scipy.stats import norm

import oracle
from conftest import rel_err


def test_golden_loglik(golden):
    for c in golden:
        o = oracle.total(c["locations"], c["times"], c["params"], probs=True)
        assert rel_err(o["ll"], c["ll"]) < 1e-14, c["name"]
        rows = oracle.rows(c["locations"], c["times"], c["params"])
        # R's rows carry the full normaliser; the oracle adds (D-1) log(1/sqrt(2 pi)) once per row at the end
        shift = (c["D"] - 1) * math.log(1 / math.sqrt(2 * math.pi))
        np.testing.assert_allclose(rows[:, 7] + shift, c["ll_rows"], rtol=1e-13, atol=1e-14, err_msg=c["name"])


def test_golden_gradient(golden):
    for c in golden:
        o = oracle.total(c["locations"], c["times"], c["params"])
        g = np.array(c["grad"])
        # componentwise error against the magnitude of the summed terms
        err = np.abs(o["grad"] - g) / np.maximum(np.maximum(np.abs(g), 1e-3 * o["scale"][1:]), 1e-300)
        assert err.max() < 1e-12, (c["name"], err)


def test_golden_probs(golden):
    for c in golden:
        o = oracle.total(c["locations"], c["times"], c["params"], probs=True)
        np.testing.assert_allclose(o["probs"], c["probs"], rtol=1e-13, atol=1e-300, err_msg=c["name"])


def test_double_oracle_matches_long_double(golden):
    for c in golden:
        a = oracle.total(c["locations"], c["times"], c["params"], precision="long")
        b = oracle.total(c["locations"], c["times"], c["params"], precision="double")
        assert rel_err(b["ll"], a["ll"]) < 1e-14
        assert np.all(np.abs(b["grad"] - a["grad"]) <= 1e-13 * a["scale"][1:])


def test_single_event_closed_form():
    # N = 1: lambda = mu0 tauX^D tauT / (2 pi) (only the j = i background term), no excitation,
    # integral: theta (e^0 - 1) = 0, mu0 (Phi(0) - Phi(-tauT t0)).
    p = np.array([1.3, 0.7, 0.2, 0.9, 0.6, 1.1])
    t0 = 3.5
    for D in (1, 2, 3):
        loc = np.zeros((1, D)) + 0.25
        o = oracle.total(loc, [t0], p, probs=True)
        lam = p[5] * p[1] ** D * p[2] / (2 * math.pi)
        expect = math.log(lam) - p[5] * (0.5 - norm.cdf(-p[2] * t0)) + (D - 1) * math.log(1 / math.sqrt(2 * math.pi))
        assert rel_err(o["ll"], expect) < 1e-15
        assert o["probs"][0] == 0.0


def test_two_events_closed_form():
    x = np.array([[0.0, 0.0], [0.3, 0.4]])
    t = np.array([1.0, 2.5])
    s, tx, tt, w, th, m = 1.3, 0.7, 0.2, 0.9, 0.6, 1.1
    p = np.array([s, tx, tt, w, th, m])
    r2, dt = 0.25, 1.5
    phi = lambda z: math.exp(-z * z / 2) / math.sqrt(2 * math.pi)
    b01 = phi(tx * 0.5) * phi(tt * dt)
    lam0 = m * tx ** 2 * tt * (phi(0) ** 2 + b01)
    lam1 = m * tx ** 2 * tt * (phi(0) ** 2 + b01) + s ** 2 * th * w * math.exp(-w * dt) * phi(s * 0.5)
    T = 2.5
    integ = sum(th * (math.exp(-w * (T - ti)) - 1) - m * (norm.cdf(tt * (T - ti)) - norm.cdf(-tt * ti)) for ti in t)
    expect = math.log(lam0) + math.log(lam1) + integ + 2 * math.log(1 / math.sqrt(2 * math.pi))
    o = oracle.total(x, t, p)
    assert rel_err(o["ll"], expect) < 1e-14
    assert math.isclose(r2, 0.25)


def test_theta_zero_has_no_excitation():
    rng = np.random.default_rng(3)
    x = rng.uniform(0, 2, (30, 2))
    t = np.sort(rng.uniform(0, 50, 30))
    p = np.array([1.0, 0.8, 0.1, 0.7, 0.0, 1.0])
    o = oracle.total(x, t, p, probs=True)
    assert np.all(o["probs"] == 0.0)
    # theta = 0: ll is the background-only likelihood, independent of sigma and omega
    q = p.copy()
    q[0], q[3] = 3.0, 5.0
    assert oracle.total(x, t, q)["ll"] == pytest.approx(o["ll"], rel=1e-15)


def test_equal_times_do_not_excite_each_other():
    x = np.array([[0.0, 0.0], [0.1, 0.0], [0.2, 0.1]])
    t = np.array([1.0, 1.0, 1.0])
    o = oracle.total(x, t, [1.0, 1.0, 0.5, 1.0, 1.0, 1.0], probs=True)
    assert np.all(o["probs"] == 0.0)


def _random_data(n=40, seed=11):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 3, (n, 2))
    t = np.sort(rng.uniform(0, 60, n))
    p = np.array([1.2, 0.9, 0.08, 0.6, 0.7, 0.9])
    return x, t, p


def test_translation_invariance():
    x, t, p = _random_data()
    a = oracle.total(x, t, p)
    b = oracle.total(x + np.array([123.25, -7.5]), t, p)
    assert rel_err(b["ll"], a["ll"]) < 1e-13


def test_permutation_invariance_keeping_last():
    x, t, p = _random_data()
    rng = np.random.default_rng(5)
    perm = np.concatenate([rng.permutation(len(t) - 1), [len(t) - 1]])
    a = oracle.total(x, t, p, probs=True)
    b = oracle.total(x[perm], t[perm], p, probs=True)
    assert rel_err(b["ll"], a["ll"]) < 1e-14
    np.testing.assert_allclose(b["probs"], a["probs"][perm], rtol=1e-14)


def test_spatial_scaling():
    # x -> a x with sigma, tauX -> sigma/a, tauX/a shifts ll by -N D log a.
    x, t, p = _random_data()
    a = 2.5
    q = p.copy()
    q[0] /= a
    q[1] /= a
    A = oracle.total(x, t, p)["ll"]
    B = oracle.total(a * x, t, q)["ll"]
    assert rel_err(B, A - len(t) * 2 * math.log(a)) < 1e-13


def test_gradient_central_differences():
    x, t, p = _random_data(n=30)
    g = oracle.total(x, t, p)["grad"]
    for k in range(6):
        h = 1e-6 * abs(p[k])
        up, dn = p.copy(), p.copy()
        up[k] += h
        dn[k] -= h
        fd = (oracle.total(x, t, up)["ll"] - oracle.total(x, t, dn)["ll"]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])), (k, fd, g[k])


def test_golden_location_gradient(golden):
    """d ll / d x against mpmath.diff of the R statement at 50 digits, both oracle precisions."""
    for c in golden:
        gold = np.array(c["loc_grad"])
        for precision, tol in (("long", 1e-14), ("double", 1e-12)):
            g, sc = oracle.location_gradient(c["locations"], c["times"], c["params"], precision=precision)
            err = np.abs(g - gold) / np.maximum(np.maximum(np.abs(gold), 1e-3 * sc), 1e-300)
            assert err.max() <= tol, (c["name"], precision, err.max())


def test_location_gradient_invariants():
    """Translation invariance makes the location gradients sum to zero over events; rotation
    invariance (D = 2) makes sum_k (x_k g_ky - y_k g_kx) zero; a mistake in the pair weights
    (w_k + w_j, or the direction of the excitation term) breaks both."""
    x, t, p = _random_data(n=60)
    g, sc = oracle.location_gradient(x, t, p)
    assert np.all(np.abs(g.sum(axis=0)) <= 1e-13 * sc.sum(axis=0))
    torque = np.sum(x[:, 0] * g[:, 1] - x[:, 1] * g[:, 0])
    assert abs(torque) <= 1e-12 * np.sum(np.abs(x) * sc)


def test_location_gradient_central_differences():
    x, t, p = _random_data(n=25, seed=12)
    g, _ = oracle.location_gradient(x, t, p)
    for k in (0, 7, 24):
        for d in range(2):
            h = 1e-6
            up, dn = x.copy(), x.copy()
            up[k, d] += h
            dn[k, d] -= h
            fd = (oracle.total(up, t, p)["ll"] - oracle.total(dn, t, p)["ll"]) / (2 * h)
            assert abs(fd - g[k, d]) <= 1e-6 * max(1.0, abs(g[k, d])), (k, d, fd, g[k, d])


def _r_rates(x, t, sd):
    """bg and se terms of R/log_lik_and_grad.R in sd parametrisation (numpy)."""
    h, tau_x, tau_t, omega, theta, mu0 = sd
    n, D = x.shape
    r2 = ((x[:, None, :] - x[None, :, :]) ** 2).sum(-1)
    dt = t[:, None] - t[None, :]
    kx_b = (2 * np.pi * tau_x ** 2) ** (-D / 2) * np.exp(-r2 / (2 * tau_x ** 2))
    kt_b = norm.pdf(t[None, :], loc=t[:, None], scale=tau_t)
    kx_s = (2 * np.pi * h ** 2) ** (-D / 2) * np.exp(-r2 / (2 * h ** 2))
    less = (t[None, :] < t[:, None])
    bg = mu0 * (kx_b * kt_b).sum(1)
    se = theta * (less * kx_s * omega * np.exp(-omega * dt)).sum(1)
    return r2, dt, kx_b, kt_b, kx_s, less, bg + se


def test_r_gradient_formulas_chain_rule():
    # R's mu_0_grad, tau_x_grad, tau_t_grad (sd parametrisation) -> precision parametrisation.
    x, t, p = _random_data(n=25)
    sd = [1 / p[0], 1 / p[1], 1 / p[2], p[3], p[4], p[5]]
    h, tau_x, tau_t, omega, theta, mu0 = sd
    n, D = x.shape
    r2, dt, kx_b, kt_b, kx_s, less, rate = _r_rates(x, t, sd)
    tn = t[-1]
    g_mu0 = ((kx_b * kt_b).sum(1) / rate).sum() - (norm.cdf((tn - t) / tau_t) - norm.cdf(-t / tau_t)).sum()
    g_taux = (mu0 / rate * (kx_b * (r2 / tau_x ** 3 - D / tau_x) * kt_b).sum(1)).sum()
    g_taut = (mu0 / rate * (kx_b * kt_b * (dt ** 2 / tau_t ** 3 - 1 / tau_t)).sum(1)).sum()
    other = norm.pdf(t, loc=tn, scale=tau_t) * (tn - t) + norm.pdf(t, loc=0, scale=tau_t) * t
    g_taut += mu0 / tau_t * other.sum()
    g = oracle.total(x, t, p)["grad"]
    assert rel_err(g[5], g_mu0) < 1e-12
    assert rel_err(g[1], -g_taux / p[1] ** 2) < 1e-12
    assert rel_err(g[2], -g_taut / p[2] ** 2) < 1e-12


def test_r_theta_and_h_gradients_at_omega_one():
    x, t, p = _random_data(n=25)
    p = p.copy()
    p[3] = 1.0
    sd = [1 / p[0], 1 / p[1], 1 / p[2], p[3], p[4], p[5]]
    h, tau_x, tau_t, omega, theta, mu0 = sd
    n, D = x.shape
    r2, dt, kx_b, kt_b, kx_s, less, rate = _r_rates(x, t, sd)
    tn = t[-1]
    lam = (less * kx_s * np.exp(-omega * dt)).sum(1)
    g_theta = ((np.exp(-omega * (tn - t)) - 1) / omega + lam / rate).sum()
    lam_h = (less * kx_s * (r2 / h ** 3 - D / h) * np.exp(-omega * dt)).sum(1)
    g_h = theta * (lam_h / rate).sum()
    g = oracle.total(x, t, p)["grad"]
    assert rel_err(g[4], g_theta) < 1e-12
    assert rel_err(g[0], -g_h / p[0] ** 2) < 1e-12


@pytest.mark.parametrize("flavour", [oracle.SERIAL, oracle.AVX, oracle.THREADED])
def test_cpu_baselines_match_oracle(flavour):
    rng = np.random.default_rng(8)
    x = rng.uniform(0, 10, (203, 2))
    t = np.sort(rng.uniform(0, 8760, 203))
    p = np.array([1.0, 1 / 1.6, 1 / 336.0, 1.0, 1.0, 1.0])
    o = oracle.total(x, t, p)
    r = oracle.cpu_baseline(flavour, x, t, p)
    assert rel_err(r[0], o["ll"]) < 1e-12
    assert np.all(np.abs(r[1:] - o["grad"]) <= 1e-11 * o["scale"][1:])
    r1 = oracle.cpu_baseline(flavour, x, t, p, grad=False)
    assert rel_err(r1[0], o["ll"]) < 1e-12


def test_oracle_rows_partial_range():
    x, t, p = _random_data(n=50)
    full = oracle.rows(x, t, p)
    part = oracle.rows(x, t, p, 10, 20)
    np.testing.assert_array_equal(part, full[10:20])


def test_phi_via_erfc():
    # Phi(z) = erfc(-z/sqrt 2)/2 as in src/CDF.h:583-592
    z = np.linspace(-8, 8, 33)
    np.testing.assert_allclose(0.5 * erfc(-z / np.sqrt(2)), norm.cdf(z), rtol=1e-13, atol=1e-300)
