"""hph_get_event_sums: the six per-event sums behind a log-likelihood-and-gradient evaluation,
against the oracle's per-row sums (oracle.rows, long double).

The engine's sums carry no normalising constants (B_ij = exp(-(tauX^2 r^2 + tauT^2 dt^2) / 2),
S_ij = exp(-omega dt - sigma^2 r^2 / 2)); the oracle's are pdf-form (B / 2 pi, S / sqrt(2 pi)).
Tolerances: the background sums contain each event's own pair (B_ii = 1) and their cut terms are
below e^-(45 + ln N) of the sum, so they are compared at 1e-12 relative. The excitation sums are
cut relative to the rate (terms below e^-X_S |cB|), so they are compared in the rate's units:
|cS dS| <= 1e-12 lambda_i, with the moments scaled as the gradient uses them (sigma^2 S r^2,
omega S dt). Checked after a full pass of each kind and after a sampler move re-evaluated from
cached sums.
"""
import numpy as np
import pytest

import oracle
from hawkes_amd import Engine, HphError, _lib, synth

pytestmark = pytest.mark.gpu

TWO_PI = 2.0 * np.pi
SQRT_TWO_PI = np.sqrt(2.0 * np.pi)


@pytest.fixture(scope="module")
def data():
    return synth.cluster_process(2500, side=np.sqrt(2500 / 100), seed=4242)


def check(sums, locs, times, params):
    ref = oracle.rows(locs, times, params)
    sigma, tauX, tauT, omega, theta, mu0 = params
    D = locs.shape[1]
    cS = sigma ** D * theta * omega
    lam = ref[:, 6]
    b = sums[:, 0:3] / TWO_PI
    s = sums[:, 3:6] / SQRT_TWO_PI
    err_b = np.max(np.abs(b - ref[:, 0:3]) / np.abs(ref[:, 0:3]).clip(1e-300))
    scale = np.array([1.0, sigma ** 2, omega])
    err_s = np.max(np.abs(cS * (s - ref[:, 3:6]) * scale) / lam[:, None])
    print("background rel %.2e excitation (rate units) %.2e" % (err_b, err_s), flush=True)
    assert err_b <= 1e-12
    assert err_s <= 1e-12


@pytest.mark.parametrize("kind", [_lib.PAIR_KERNEL_ROWWISE, _lib.PAIR_KERNEL_SYMMETRIC, _lib.PAIR_KERNEL_SPLIT])
def test_event_sums_against_oracle_rows(data, kind):
    locs, times = data
    p = np.array([1.05, 0.6, 1 / 320.0, 0.95, 0.9, 1.1])
    e = Engine(2, len(times))
    e.set_pair_kernel(kind)
    e.update_locations(locs)
    e.set_times(times)
    e.set_parameters(p)
    with pytest.raises(HphError):
        e.event_sums()  # nothing evaluated yet
    e.log_likelihood_and_gradient()
    assert e.last_pass()[0] == kind
    check(e.event_sums(), locs, times, p)
    # a sigma / omega move re-evaluated from the cached background sums
    q = p.copy()
    q[0] *= 1.07
    q[3] *= 0.93
    e.set_parameters(q)
    e.log_likelihood_and_gradient()
    assert e.last_evaluation() == _lib.EVAL_EXCITATION
    check(e.event_sums(), locs, times, q)
    # a parameter change without an evaluation: no sums of the current parameters
    e.set_parameters(p)
    with pytest.raises(HphError):
        e.event_sums()
    e.close()
