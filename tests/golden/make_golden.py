"""Writes the golden fixtures in this folder (``golden_*.json``).

The values come from a 50-digit mpmath evaluation of the *independent* naive statement of
the model in R/log_lik_and_grad.R:2-93 (sd parametrisation h = 1/sigmaXprec,
tau_x = 1/tauXprec, tau_t = 1/tauTprec, full multivariate-normal normalisers via the
dmvnorm/dnorm/pnorm definitions), restated here from the formulas:

  bg_rate(i)   = mu0 * sum_j N_D(x_j; x_i, tau_x^2 I) * N(t_j; t_i, tau_t^2)      (:2-22)
  se_rate(i)   = theta * sum_j [t_j < t_i] N_D(x_j; x_i, h^2 I) * omega e^{-omega(t_i-t_j)}
                                                                              (:24-49)
  integral     = mu0 * sum_i (Phi((t_n - t_i)/tau_t) - Phi(-t_i/tau_t))
                 - theta * sum_i (e^{-omega(t_n - t_i)} - 1),   t_n = last element  (:61-77)
  log_lik      = sum_i log(bg_rate(i) + se_rate(i)) - integral                  (:79-93)
  prob_se(i)   = se_rate(i) / (bg_rate(i) + se_rate(i))                         (:51-55)

The gradient with respect to the reference's parameters (sigmaXprec, tauXprec, tauTprec,
omega, theta, mu0) is mpmath.diff of log_lik at 50 digits, and so is the location gradient
d log_lik / d x_kd ("loc_grad", N x D). No reference code is run.

Usage: python tests/golden/make_golden.py   (rewrites the fixtures; deterministic)
"""
import json
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 50
HERE = os.path.dirname(os.path.abspath(__file__))


def _normal_pdf_d(xj, xi, sd):
    """Density of N_D(xi, sd^2 I) at xj."""
    D = len(xi)
    r2 = mp.fsum((mp.mpf(a) - mp.mpf(b)) ** 2 for a, b in zip(xj, xi))
    return (2 * mp.pi * sd ** 2) ** (-mp.mpf(D) / 2) * mp.exp(-r2 / (2 * sd ** 2))


def _phi_cdf(z):
    return mp.ncdf(z)


def rates(locs, times, prec):
    sigma, tauX, tauT, omega, theta, mu0 = prec
    h, tau_x, tau_t = 1 / sigma, 1 / tauX, 1 / tauT
    N = len(times)
    bg, se = [], []
    for i in range(N):
        ti = mp.mpf(times[i])
        b = mp.fsum(_normal_pdf_d(locs[j], locs[i], tau_x) *
                    mp.npdf(mp.mpf(times[j]), ti, tau_t) for j in range(N))
        s = mp.fsum(_normal_pdf_d(locs[j], locs[i], h) * omega *
                    mp.exp(-omega * (ti - mp.mpf(times[j])))
                    for j in range(N) if mp.mpf(times[j]) < ti)
        bg.append(mu0 * b)
        se.append(theta * s)
    return bg, se


def log_lik(locs, times, prec, per_row=False):
    sigma, tauX, tauT, omega, theta, mu0 = prec
    tau_t = 1 / tauT
    bg, se = rates(locs, times, prec)
    tn = mp.mpf(times[-1])
    rows = []
    for i in range(len(times)):
        ti = mp.mpf(times[i])
        rows.append(mp.log(bg[i] + se[i])
                    - mu0 * (_phi_cdf((tn - ti) / tau_t) - _phi_cdf(-ti / tau_t))
                    + theta * (mp.exp(-omega * (tn - ti)) - 1))
    total = mp.fsum(rows)
    if per_row:
        return total, rows, [s / (b + s) for b, s in zip(bg, se)]
    return total


def gradient(locs, times, prec):
    g = []
    for k in range(6):
        def f(v, k=k):
            q = list(prec)
            q[k] = v
            return log_lik(locs, times, q)
        g.append(mp.diff(f, prec[k]))
    return g


def location_gradient(locs, times, prec):
    out = []
    for i in range(len(locs)):
        row = []
        for d in range(len(locs[i])):
            def f(v, i=i, d=d):
                L = [[mp.mpf(a) for a in r] for r in locs]
                L[i][d] = v
                return log_lik(L, times, prec)
            row.append(mp.diff(f, mp.mpf(locs[i][d])))
        out.append(row)
    return out


def case(name, locs, times, params, note):
    locs = np.asarray(locs, dtype=np.float64)
    if locs.ndim == 1:
        locs = locs[:, None]
    times = np.asarray(times, dtype=np.float64)
    prec = [mp.mpf(float(p)) for p in params]
    ll, rows, probs = log_lik(locs.tolist(), times.tolist(), prec, per_row=True)
    g = gradient(locs.tolist(), times.tolist(), prec)
    lg = location_gradient(locs.tolist(), times.tolist(), prec)
    return {
        "name": name,
        "note": note,
        "D": int(locs.shape[1]),
        "N": int(locs.shape[0]),
        "locations": locs.tolist(),
        "times": times.tolist(),
        "params": [float(p) for p in params],
        "ll": float(ll),
        "ll_rows": [float(r) for r in rows],
        "grad": [float(x) for x in g],
        "probs": [float(p) for p in probs],
        "loc_grad": [[float(v) for v in row] for row in lg],
        "ll_str": mp.nstr(ll, 30),
    }


def main():
    rng = np.random.default_rng(20240666)
    P0 = [1.0, 1 / 1.6, 1 / 336.0, 1.0, 1.0, 1.0]
    cases = []
    cases.append(case("n1", [[0.3, -1.2]], [5.0], P0, "single event: closed form"))
    cases.append(case("n2", [[0.0, 0.0], [0.4, -0.3]], [1.0, 2.5], [1.3, 0.7, 0.2, 0.9, 0.6, 1.1],
                      "two events, all parameters away from 1"))
    # the reference's R smoke-test shape: N=10, D=2, normal locations, times 1..10, Exp(1) params
    cases.append(case("rtest10", rng.normal(size=(10, 2)), np.arange(1, 11, dtype=float),
                      rng.exponential(size=6), "shape of R/Untitled.R:87-115 test()"))
    # omega != 1 and theta != 1 (R's theta/h/omega gradients disagree here)
    cases.append(case("omega25", rng.uniform(0, 3, size=(9, 2)), np.sort(rng.uniform(0, 20, 9)),
                      [0.8, 1.4, 0.15, 2.5, 0.4, 0.7], "omega = 2.5"))
    # ties in time: strict dt > 0 mask (src/NewHawkes.hpp:540)
    t = np.array([0.0, 1.0, 1.0, 1.0, 2.5, 2.5, 4.0, 7.0])
    cases.append(case("ties", rng.uniform(0, 2, size=(8, 2)), t, [1.1, 0.9, 0.5, 1.2, 0.8, 0.6],
                      "repeated event times"))
    # unsorted times: T is the last element, not the maximum
    cases.append(case("unsorted", rng.uniform(0, 2, size=(7, 2)),
                      np.array([3.0, 0.5, 6.0, 2.0, 4.5, 1.0, 5.0]), [1.0, 0.8, 0.3, 0.7, 0.9, 1.2],
                      "unsorted times; T = t[N-1] = 5 < max"))
    cases.append(case("d1", rng.uniform(0, 4, size=(8, 1)), np.sort(rng.uniform(0, 30, 8)),
                      [0.9, 1.3, 0.1, 0.6, 0.7, 0.5], "D = 1"))
    cases.append(case("d3", rng.uniform(0, 4, size=(8, 3)), np.sort(rng.uniform(0, 30, 8)),
                      [0.9, 1.3, 0.1, 0.6, 0.7, 0.5], "D = 3"))
    # benchmark-like scales (P0, hours and km)
    cases.append(case("p0_16", rng.uniform(0, 10, size=(16, 2)), np.sort(rng.uniform(0, 8760, 16)),
                      P0, "config-1 scales, N = 16"))
    clustered_t = np.sort(np.concatenate([rng.uniform(0, 200, 6), 50 + rng.exponential(5, 6)]))
    cases.append(case("clustered12", rng.normal(0, 0.3, size=(12, 2)), clustered_t,
                      [1.0, 1 / 1.6, 1 / 336.0, 0.2, 0.5, 0.05], "clustered events, strong excitation"))
    path = os.path.join(HERE, "golden_cases.json")
    with open(path, "w") as f:
        json.dump({"generator": "tests/golden/make_golden.py", "dps": 50, "cases": cases}, f, indent=1)
    print("wrote", path, len(cases), "cases")


if __name__ == "__main__":
    main()
