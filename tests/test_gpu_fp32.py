"""FP32-bulk mode (HPH_FLAG_FLOAT) against the FP64 oracle.

The background exponential (the bulk of the pair work) is one v_exp_f32 on the FP64-formed
exponent rounded to FP32; distances, time differences, the excitation term, every accumulation,
log, Phi and the exp integral terms and the gradient algebra stay FP64. Tolerances (stated per
result; measured errors are printed):
  * log-likelihood: relative error <= 1e-6;
  * gradient component k: |g - g_oracle| <= 1e-6 * max(|g_oracle|, 1e-3 * S_k), S_k = sum of
    |per-row terms|;
  * probabilities: |p - p_oracle| <= 1e-6 * max(p_oracle, 1e-3).
"""
import numpy as np
import pytest

import oracle
from conftest import rel_err
from hawkes_amd import Engine, _lib, synth

pytestmark = pytest.mark.gpu

TOL = 1e-6

CONFIGS = {"rowwise": (_lib.PAIR_KERNEL_ROWWISE, 2, 4096), "symmetric": (_lib.PAIR_KERNEL_SYMMETRIC, 2, 4096),
           "symmetric_a1": (_lib.PAIR_KERNEL_SYMMETRIC, 1, 1024),
           "rowwise_narrow": (_lib.PAIR_KERNEL_ROWWISE, 2, 65536), "split": (_lib.PAIR_KERNEL_SPLIT, 2, 4096)}


def fp32_engine(locs, times, params, kernel):
    e = Engine(locs.shape[1], locs.shape[0], flags=_lib.OPENCL | _lib.FLOAT)
    e.set_pair_config(*CONFIGS[kernel])
    e.update_locations(locs)
    e.set_times(times)
    e.set_parameters(params)
    return e


def errors(e, ref):
    ll, g = e.log_likelihood_and_gradient()
    p = e.probs_self_excite()
    ll_err = rel_err(ll, ref["ll"])
    g_err = np.abs(g - ref["grad"]) / np.maximum(np.abs(ref["grad"]), 1e-3 * ref["scale"][1:])
    p_err = np.abs(p - ref["probs"]) / np.maximum(ref["probs"], 1e-3)
    return float(ll_err), float(g_err.max()), float(p_err.max())


@pytest.mark.parametrize("kernel", ["rowwise", "symmetric", "symmetric_a1", "rowwise_narrow", "split"])
def test_fp32_cluster(kernel):
    locs, times = synth.cluster_process(3500, seed=123)
    ref = oracle.total(locs, times, synth.P0, probs=True)
    e = fp32_engine(locs, times, synth.P0, kernel)
    ll_err, g_err, p_err = errors(e, ref)
    print("fp32 %s N=%d: ll %.2e grad %.2e probs %.2e" % (kernel, len(times), ll_err, g_err, p_err))
    assert ll_err <= TOL and g_err <= TOL and p_err <= TOL


@pytest.mark.parametrize("D", [1, 3])
def test_fp32_dimensions(D):
    rng = np.random.default_rng(9 + D)
    n = 2000
    locs = rng.uniform(0, 5, (n, D))
    times = np.sort(rng.uniform(0, 600, n))
    p = np.array([0.9, 0.7, 0.05, 0.4, 0.8, 0.6])
    ref = oracle.total(locs, times, p, probs=True)
    for kernel in ("rowwise", "symmetric"):
        e = fp32_engine(locs, times, p, kernel)
        ll_err, g_err, p_err = errors(e, ref)
        print("fp32 D=%d %s: ll %.2e grad %.2e probs %.2e" % (D, kernel, ll_err, g_err, p_err))
        assert ll_err <= TOL and g_err <= TOL and p_err <= TOL


def test_fp32_config2_size():
    """N ~ 49.5k cluster process: FP32-bulk against the compensated FP64 oracle."""
    locs, times, p = synth.workload("cluster_b35k")
    ref = oracle.total(locs, times, p, precision="double", probs=True)
    e = fp32_engine(locs, times, p, "symmetric")
    ll_err, g_err, p_err = errors(e, ref)
    print("fp32 N=%d: ll %.2e grad %.2e probs %.2e" % (len(times), ll_err, g_err, p_err))
    assert ll_err <= TOL and g_err <= TOL and p_err <= TOL


@pytest.mark.parametrize("kernel", ["rowwise", "symmetric", "split"])
def test_fp32_tile_skip_matches_all_pairs(kernel):
    """FP32-bulk with and without tile skipping (skipped background terms are below
    e^-(45 + ln N) relative to each event's sums): equal to within rounding."""
    locs, times = synth.cluster_process(6000, side=20.0, horizon=43_800.0, seed=41)
    out = []
    for flags in (_lib.OPENCL | _lib.FLOAT, _lib.OPENCL | _lib.FLOAT | _lib.ALL_PAIRS):
        e = Engine(2, len(times), flags=flags)
        e.set_pair_config(*CONFIGS[kernel])
        e.update_locations(locs)
        e.set_times(times)
        e.set_parameters(synth.P0)
        out.append(e.log_likelihood_and_gradient())
    assert abs(out[0][0] - out[1][0]) <= 1e-15 * abs(out[1][0])
    np.testing.assert_allclose(out[0][1], out[1][1], rtol=1e-13, atol=1e-15 * np.abs(out[1][1]).max())
