"""Parity of the HIP engine (through the C ABI) with the CPU oracle and the golden vectors.

Tolerances (FP64 mode):
  * log-likelihood: relative error <= 1e-12;
  * gradient component k: |g - g_oracle| <= 1e-12 * max(|g_oracle|, 1e-3 * S_k), where S_k is
    the oracle's sum of |per-row terms| of that component (gradients cancel near the MLE);
  * self-excitation probabilities: relative error <= 1e-12 per event (exact 0 for events with
    no earlier event).
"""
import math

import numpy as np
import pytest

import oracle
from conftest import rel_err
from hawkes_amd import Engine, HphError, _lib, synth

pytestmark = pytest.mark.gpu

LL_TOL = 1e-12
GRAD_TOL = 1e-12
PROB_TOL = 1e-12


KERNELS = {"auto": (_lib.PAIR_KERNEL_AUTO, 2, 4096), "rowwise": (_lib.PAIR_KERNEL_ROWWISE, 2, 4096),
           "symmetric": (_lib.PAIR_KERNEL_SYMMETRIC, 2, 4096),
           "symmetric_a1": (_lib.PAIR_KERNEL_SYMMETRIC, 1, 1024),
           "rowwise_small_grid": (_lib.PAIR_KERNEL_ROWWISE, 2, 64),
           "rowwise_narrow": (_lib.PAIR_KERNEL_ROWWISE, 2, 65536),
           "split": (_lib.PAIR_KERNEL_SPLIT, 2, 4096), "split_a1": (_lib.PAIR_KERNEL_SPLIT, 1, 1024)}


def make_engine(locs, times, params, kernel="auto", flags=_lib.OPENCL):
    locs = np.asarray(locs, dtype=np.float64)
    if locs.ndim == 1:
        locs = locs[:, None]
    e = Engine(locs.shape[1], locs.shape[0], flags=flags)
    e.set_pair_config(*KERNELS[kernel])
    e.update_locations(locs)
    e.set_times(times)
    e.set_parameters(params)
    return e


def grad_err(g, ref):
    g, go, scale = np.asarray(g), ref["grad"], ref["scale"][1:]
    return np.abs(g - go) / np.maximum(np.maximum(np.abs(go), 1e-3 * scale), 1e-300)


def assert_matches_oracle(locs, times, params, probs=True, precision="long", name="", kernel="auto"):
    ref = oracle.total(locs, times, params, precision=precision, probs=probs)
    e = make_engine(locs, times, params, kernel=kernel)
    ll, g = e.log_likelihood_and_gradient()
    assert rel_err(ll, ref["ll"]) <= LL_TOL, (name, ll, ref["ll"])
    assert grad_err(g, ref).max() <= GRAD_TOL, (name, grad_err(g, ref))
    assert e.log_likelihood() == ll
    if probs:
        p = e.probs_self_excite()
        np.testing.assert_allclose(p, ref["probs"], rtol=PROB_TOL, atol=0, err_msg=name)
    e.close()
    return ref


@pytest.mark.parametrize("kernel", ["rowwise", "symmetric", "split"])
def test_golden_vectors(golden, kernel):
    for c in golden:
        if kernel in ("symmetric", "split") and c["D"] > 3:
            continue
        print("golden case", c["name"], kernel, flush=True)
        e = make_engine(c["locations"], c["times"], c["params"], kernel=kernel)
        ll, g = e.log_likelihood_and_gradient()
        ref = oracle.total(c["locations"], c["times"], c["params"])
        assert rel_err(ll, c["ll"]) <= LL_TOL, c["name"]
        gold = np.array(c["grad"])
        err = np.abs(g - gold) / np.maximum(np.maximum(np.abs(gold), 1e-3 * ref["scale"][1:]), 1e-300)
        assert err.max() <= GRAD_TOL, (c["name"], err)
        np.testing.assert_allclose(e.probs_self_excite(), c["probs"], rtol=PROB_TOL, atol=0, err_msg=c["name"])
        e.close()


@pytest.mark.parametrize("kernel", ["rowwise", "symmetric", "split"])
def test_uniform_n1000_config(kernel):
    locs, times, p = synth.workload("uniform_n1000")
    assert_matches_oracle(locs, times, p, name="uniform_n1000", kernel=kernel)


@pytest.mark.parametrize("kernel", ["rowwise", "symmetric", "symmetric_a1", "rowwise_small_grid", "rowwise_narrow",
                                    "split", "split_a1"])
def test_cluster_process_small(kernel):
    locs, times = synth.cluster_process(2800, side=np.sqrt(4000 / 100), seed=777)
    assert_matches_oracle(locs, times, synth.P0, name="cluster_4k", kernel=kernel)


@pytest.mark.parametrize("D,kernel", [(1, "rowwise"), (3, "rowwise"), (5, "rowwise"), (8, "rowwise"),
                                      (1, "symmetric"), (3, "symmetric"), (1, "rowwise_narrow"),
                                      (5, "rowwise_narrow"), (1, "split"), (2, "split"), (3, "split")])
def test_dimensions(D, kernel):
    rng = np.random.default_rng(100 + D)
    n = 1700
    locs = rng.uniform(0, 5, (n, D))
    times = np.sort(rng.uniform(0, 500, n))
    p = np.array([0.9, 0.7, 0.05, 0.4, 0.8, 0.6])
    assert_matches_oracle(locs, times, p, name="D%d" % D, kernel=kernel)


@pytest.mark.parametrize("kernel", ["rowwise", "symmetric", "rowwise_narrow", "split"])
def test_unsorted_times_and_ties(kernel):
    rng = np.random.default_rng(21)
    n = 1900
    locs = rng.uniform(0, 4, (n, 2))
    times = np.round(rng.uniform(0, 300, n))  # many ties, not sorted; T = last element
    p = np.array([1.1, 0.8, 0.02, 0.3, 0.9, 0.7])
    assert_matches_oracle(locs, times, p, name="unsorted_ties", kernel=kernel)


@pytest.mark.parametrize("kernel", ["rowwise", "symmetric", "symmetric_a1", "rowwise_narrow", "split", "split_a1"])
def test_row_counts_around_block_sizes(kernel):
    rng = np.random.default_rng(4)
    for n in (1, 2, 63, 64, 65, 255, 256, 257, 511, 512, 513, 1023, 1025, 1537):
        locs = rng.uniform(0, 3, (n, 2))
        times = np.sort(rng.uniform(0, 100, n))
        assert_matches_oracle(locs, times, [1.0, 0.9, 0.1, 0.5, 0.6, 0.9], name="n%d" % n, kernel=kernel)


@pytest.mark.parametrize("kernel", ["rowwise", "symmetric", "rowwise_narrow", "split"])
def test_batch_equals_single_calls_bitwise(kernel):
    locs, times = synth.cluster_process(1500, seed=31)
    P = synth.jitter_params(8)
    e = make_engine(locs, times, P[0], kernel=kernel)
    e.set_parameters_batch(P)
    ll_b, g_b = e.log_likelihood_batch()
    for k in range(8):
        e.set_parameters(P[k])
        ll, g = e.log_likelihood_and_gradient()
        assert ll == ll_b[k]
        assert np.array_equal(g, g_b[k])
    ll_only = e.log_likelihood_batch(gradient=False)
    assert np.array_equal(ll_only, ll_b)


@pytest.mark.parametrize("name,params", [
    ("p0", synth.P0),
    ("small_omega_large_tauT", [1.0, 1 / 1.6, 1 / 3.0, 0.05, 1.0, 1.0]),
    ("large_omega", [1.0, 1 / 1.6, 1 / 336.0, 40.0, 1.0, 1.0]),
])
@pytest.mark.parametrize("kernel", ["rowwise", "symmetric", "split"])
def test_tile_skip_matches_all_pairs(name, params, kernel):
    """Tiles are skipped where every background term is below e^-(45 + ln N) (each event's sums
    change by < 4e-18 relative; its own pair contributes 1) or, in likelihood and gradient passes,
    every excitation term is below e^-X_S (rates change by < 2e-22 relative; passes that return
    probabilities skip exact zeros only, or terms below 2^-60 of each event's own excitation sum).
    Results equal the every-pair evaluation
    (HPH_FLAG_ALL_PAIRS) to within rounding: ll and probabilities to 1e-15 relative, gradient
    component k to 1e-15 of its summation scale S_k (the oracle's sum of |per-event terms|)."""
    locs, times = synth.cluster_process(6000, side=20.0, horizon=43_800.0, seed=41)
    scale = oracle.total(locs, times, params, precision="double")["scale"][1:]
    out = []
    for flags in (_lib.OPENCL, _lib.OPENCL | _lib.ALL_PAIRS):
        e = make_engine(locs, times, params, kernel=kernel, flags=flags)
        probs = e.probs_self_excite()  # first: a full probability pass on both engines
        ll, g = e.log_likelihood_and_gradient()
        if flags == _lib.OPENCL:
            # a later request at the cached sums: background sums of the likelihood pass, excitation
            # sums of the band kernel (another summation order and, for omega >= 0, the table
            # exponential against a natural-form pass's exp: a term with exponent x carries
            # ~|x| 2^-53 of rounding in either, up to 8e-14 near the exact-zero point |x| ~ 745,
            # where the tiny probabilities of the large-omega case are formed)
            e.set_parameters(params)
            e.log_likelihood_and_gradient()
            np.testing.assert_allclose(e.probs_self_excite(), probs, rtol=4e-13, atol=0, err_msg=name)
            assert e.last_evaluation() == _lib.EVAL_EXCITATION
        out.append((ll, g, probs))
        e.close()
    assert rel_err(out[0][0], out[1][0]) <= 1e-15, name
    assert np.all(np.abs(out[0][1] - out[1][1]) <= 1e-15 * scale), (name, out[0][1] - out[1][1])
    np.testing.assert_allclose(out[0][2], out[1][2], rtol=1e-15, atol=0, err_msg=name)


@pytest.mark.parametrize("kernel", ["rowwise", "symmetric", "split"])
def test_probs_from_likelihood_pass_equal_fresh_pass(kernel):
    locs, times = synth.cluster_process(3000, seed=19)
    e = make_engine(locs, times, synth.P0, kernel=kernel)
    e.log_likelihood_and_gradient()
    cached = e.probs_self_excite()
    e.set_parameters(synth.P0)  # invalidates: the next request forms the probabilities again
    fresh = e.probs_self_excite()
    assert np.array_equal(cached, fresh)
    P = synth.jitter_params(3)
    e.set_parameters_batch(P)
    e.log_likelihood_batch()  # batches leave the single-vector probabilities alone
    assert np.array_equal(e.probs_self_excite(), fresh)
    e.close()


@pytest.mark.parametrize("kernel", ["rowwise", "symmetric", "split"])
def test_huge_exponents_use_the_clamped_exp(kernel):
    """Exponents far beyond 1e9 (tauT^2 dt^2 / 2 up to 5e9 here): the engine must select the
    clamped exp (the clamp-free shifter form is exact only below ~1.4e9)."""
    rng = np.random.default_rng(77)
    n = 1500
    locs = rng.uniform(0, 3, (n, 2))
    times = np.sort(rng.uniform(0, 1e5, n))
    p = np.array([1.0, 0.9, 1.0, 0.5, 0.8, 0.9])
    assert_matches_oracle(locs, times, p, name="huge", kernel=kernel)


@pytest.mark.parametrize("kernel", ["symmetric", "split"])
def test_wide_region_far_offset_locations(kernel):
    """Locations over 3000 km, offset by 5e6 (projected coordinates), with a 1.6 km background
    kernel: tauX x diagonal is past the bound of the base-2 form's scaled coordinates, so the
    natural form must run; every pair's distance is formed from the original coordinates."""
    rng = np.random.default_rng(79)
    n = 3000
    locs = 5e6 + rng.uniform(0, 3000.0, (n, 2))
    locs[: n // 3] = locs[n // 3: 2 * (n // 3)] + rng.normal(0, 0.5, (n // 3, 2))
    times = np.sort(rng.uniform(0, 8760.0, n))
    assert_matches_oracle(locs, times, synth.P0, probs=False, name="wide", kernel=kernel)
    # isolated events have subnormal probabilities (excitation terms near e^-745): those agree to
    # the subnormal spacing, every other one to PROB_TOL
    e = make_engine(locs, times, synth.P0, kernel=kernel)
    p = e.probs_self_excite()
    e.close()
    p_ref = oracle.total(locs, times, synth.P0, probs=True)["probs"]
    np.testing.assert_allclose(p, p_ref, rtol=PROB_TOL, atol=1e-300)


@pytest.mark.parametrize("kernel", ["symmetric", "symmetric_a1", "rowwise", "split"])
def test_near_tied_times_keep_the_strict_mask(kernel):
    """Pairs of events one ulp apart in time excite each other (dt > 0), exact ties do not. Scaled
    times would merge the one-ulp pairs, so the engine must keep the original times for them; the
    oracle decides the mask on the original times."""
    rng = np.random.default_rng(78)
    n = 5000
    locs = rng.uniform(0, 4, (n, 2))
    times = np.sort(rng.uniform(0, 8760.0, n))
    for i in range(10, n - 1, 97):
        times[i + 1] = np.nextafter(times[i], np.inf)
    for i in range(50, n - 1, 211):
        times[i + 1] = times[i]
    times = np.sort(times)
    assert np.sum(np.diff(times) == 0) > 0 and np.sum((np.diff(times) > 0) & (np.diff(times) < 1e-11)) > 0
    assert_matches_oracle(locs, times, synth.P0, name="near-ties", kernel=kernel)


@pytest.mark.parametrize("name,params", [
    ("theta0", [1.0, 0.8, 0.05, 0.5, 0.0, 0.9]),
    ("omega0", [1.0, 0.8, 0.05, 0.0, 0.7, 0.9]),
    ("sigma0", [0.0, 0.8, 0.05, 0.5, 0.7, 0.9]),
    ("negative_times", [1.1, 0.7, 0.04, 0.6, 0.8, 0.7]),
])
@pytest.mark.parametrize("kernel", ["rowwise_narrow", "symmetric", "split"])
def test_boundary_parameters_and_negative_times(name, params, kernel):
    """theta = 0, omega = 0 or sigma = 0 switch the excitation term off (its constant is zero) or
    make it time-independent; times may be negative (Phi(-tauT t_i) > 1/2)."""
    rng = np.random.default_rng(81)
    n = 2600
    locs = rng.uniform(0, 3, (n, 2))
    times = np.sort(rng.uniform(-50, 50, n))
    assert_matches_oracle(locs, times, params, name=name, kernel=kernel)


@pytest.mark.parametrize("kernel", ["rowwise", "symmetric", "split"])
def test_negative_theta_takes_no_excitation_cut(kernel):
    """theta < 0 (accepted by setParameters, unphysical): lambda_i = cB sum B - |cS| sum S is a
    difference, so the floor lambda_i >= |cB| that bounds the negligible-excitation cut does not
    hold; with cB and cS of different signs the passes skip exactly zero terms only. Results
    match the oracle (1e-12) and the every-pair evaluation to rounding (1e-15)."""
    locs, times = synth.cluster_process(3000, side=np.sqrt(4300 / 100), seed=91)
    p = np.array([1.0, 1 / 1.6, 1 / 336.0, 1.0, -3e-3, 1.0])
    ref = assert_matches_oracle(locs, times, p, name="theta<0", kernel=kernel)
    out = []
    for flags in (_lib.OPENCL, _lib.OPENCL | _lib.ALL_PAIRS):
        e = make_engine(locs, times, p, kernel=kernel, flags=flags)
        out.append(e.log_likelihood_and_gradient())
        e.close()
    assert rel_err(out[0][0], out[1][0]) <= 1e-15
    assert np.all(np.abs(out[0][1] - out[1][1]) <= 1e-15 * ref["scale"][1:])


@pytest.mark.parametrize("kernel", ["rowwise", "symmetric", "split"])
def test_negative_omega_takes_the_natural_form(kernel):
    """omega < 0 (accepted by setParameters, unphysical; with theta < 0 too, so that theta omega
    > 0 keeps every rate positive): excitation terms grow with the time gap (exponents of either
    sign), which the base-2 form's table exponential does not take; the passes run the natural
    form, and no excitation tile is skipped (the cut needs omega >= 0). Results match the oracle
    (1e-12)."""
    locs, times = synth.cluster_process(3000, side=np.sqrt(4300 / 100), seed=92)
    p = np.array([1.0, 1 / 1.6, 1 / 336.0, -2e-4, -1.0, 1.0])
    assert_matches_oracle(locs, times, p, name="omega<0", kernel=kernel)
    if kernel != "rowwise":
        for params, base2 in ((p, False), (np.abs(p), True)):  # omega > 0: the base-2 form
            e = make_engine(locs, times, params, kernel=kernel)
            e.log_likelihood_and_gradient()
            assert e.last_pass()[1] is base2
            e.close()


def test_repeat_calls_are_bitwise_identical():
    locs, times = synth.cluster_process(3000, seed=9)
    e = make_engine(locs, times, synth.P0)
    a = e.log_likelihood_and_gradient()
    pa = e.probs_self_excite()
    for _ in range(3):
        e.set_parameters(synth.P0)  # invalidates the cached result
        b = e.log_likelihood_and_gradient()
        assert a[0] == b[0] and np.array_equal(a[1], b[1])
        assert np.array_equal(pa, e.probs_self_excite())


def test_reused_queue_order_leaves_results_unchanged():
    """Full passes at tauT steps of 1-2%, then a 40% jump, then a step back: every call orders its
    work queue on the device for its parameters (items of one cost class in any order), and each
    result is bitwise equal to a fresh engine's first call at the same parameters, and to the
    oracle at 1e-12."""
    locs, times = synth.cluster_process(6000, side=8.0, seed=23)
    base = np.array(synth.P0, dtype=float)
    e = make_engine(locs, times, base, kernel="symmetric")
    for i, scale in enumerate((1.0, 1.01, 1.02, 0.99, 1.4, 1.41, 1.0)):
        p = base.copy()
        p[2] *= scale
        e.set_parameters(p)
        ll, g = e.log_likelihood_and_gradient()
        assert e.last_evaluation() == _lib.EVAL_FULL
        f = make_engine(locs, times, p, kernel="symmetric")
        llf, gf = f.log_likelihood_and_gradient()
        assert ll == llf and np.array_equal(g, gf), (scale, ll, llf)
        if i in (0, 4):
            ref = oracle.total(locs, times, p, precision="double")
            assert rel_err(ll, ref["ll"]) <= LL_TOL and grad_err(g, ref).max() <= GRAD_TOL
        print("tauT x %.2f: bitwise equal to a fresh engine" % scale, flush=True)


@pytest.fixture(scope="module")
def config2_reference():
    locs, times, p = synth.workload("cluster_b35k")
    return locs, times, p, oracle.total(locs, times, p, precision="double")


@pytest.mark.parametrize("kernel", ["rowwise", "symmetric"])
def test_config2_full_parity_and_finite_differences(config2_reference, kernel):
    """Cluster process, 35,000 background events (N ~ 49.5k): FP64 ll + 6-gradient against the
    compensated double oracle, and the gradient against central finite differences of the GPU
    log-likelihood (h = 1e-6 relative)."""
    locs, times, p, ref = config2_reference
    print("config2 N =", len(times), kernel, flush=True)
    e = make_engine(locs, times, p, kernel=kernel)
    ll, g = e.log_likelihood_and_gradient()
    assert rel_err(ll, ref["ll"]) <= LL_TOL
    assert grad_err(g, ref).max() <= GRAD_TOL
    for k in range(6):
        h = 1e-6 * p[k]
        up, dn = p.copy(), p.copy()
        up[k] += h
        dn[k] -= h
        e.set_parameters(up)
        fu = e.log_likelihood()
        e.set_parameters(dn)
        fd = (fu - e.log_likelihood()) / (2 * h)
        assert abs(fd - g[k]) <= 1e-5 * max(abs(g[k]), 1e-3 * ref["scale"][1 + k]), (k, fd, g[k])


@pytest.mark.parametrize("kernel", ["rowwise", "symmetric", "symmetric_a1", "split"])
def test_full_size_rows_against_oracle(kernel):
    """N ~ 1e5 (70,000 background events): per-event probabilities of sampled rows against
    the oracle's row sums over all N columns, and translation invariance of the total."""
    locs, times, p = synth.workload("cluster_b70k")
    n = len(times)
    e = make_engine(locs, times, p, kernel=kernel)
    probs = e.probs_self_excite()
    ll, g = e.log_likelihood_and_gradient()
    for lo in (0, 511, n // 3, n - 600, n - 64):
        rows = oracle.rows(locs, times, p, lo, lo + 64, precision="double")
        np.testing.assert_allclose(probs[lo:lo + 64], rows[:, 8], rtol=PROB_TOL, atol=0)
    assert probs[0] == 0.0 and np.all((probs >= 0) & (probs <= 1))
    e.update_locations(locs + np.array([1000.0, -250.0]))
    ll2 = e.log_likelihood()
    assert rel_err(ll2, ll) <= 1e-12
    e.close()


def test_kernels_agree_at_full_size():
    locs, times, p = synth.workload("cluster_b70k")
    res = []
    for kernel in ("rowwise", "symmetric", "split"):
        e = make_engine(locs, times, p, kernel=kernel)
        res.append(e.log_likelihood_and_gradient())
        e.close()
    for r in res[1:]:
        assert rel_err(r[0], res[0][0]) <= 1e-13
        assert np.all(np.abs(r[1] - res[0][1]) <= 1e-12 * np.maximum(np.abs(res[0][1]), 1.0))


def test_autotune_picks_a_candidate_and_keeps_parity(tmp_path, monkeypatch):
    locs, times = synth.cluster_process(6000, seed=77)
    p = synth.P0
    ref = oracle.total(locs, times, p, precision="double")
    e = make_engine(locs, times, p)
    ms = e.autotune(repeats=1)
    assert ms > 0
    kind, a, target = e.tuned_config()
    assert kind in (_lib.PAIR_KERNEL_ROWWISE, _lib.PAIR_KERNEL_SYMMETRIC) and a in (1, 2) and target >= 64
    e.set_parameters(p)
    ll, g = e.log_likelihood_and_gradient()
    assert rel_err(ll, ref["ll"]) <= LL_TOL
    assert grad_err(g, ref).max() <= GRAD_TOL
    # a second engine of the same shape starts from the tuned configuration
    e2 = make_engine(locs, times, p)
    assert e2.tuned_config() == (kind, a, target)


def test_single_location_update_and_state():
    locs, times = synth.cluster_process(900, seed=12)
    p = synth.P0
    e = make_engine(locs, times, p)
    ll0 = e.log_likelihood()
    e.store_state()
    moved = locs.copy()
    moved[17] += np.array([0.3, -0.2])
    e.update_locations(moved[17], index=17)
    ll1 = e.log_likelihood()
    assert rel_err(ll1, oracle.total(moved, times, p)["ll"]) <= LL_TOL
    e.restore_state()
    assert e.log_likelihood() == ll0
    e.set_parameters(p)
    assert rel_err(e.log_likelihood(), ll0) <= 1e-15
    e.accept_state()
    assert e.internal_dimension() == 2


def test_error_codes():
    e = Engine(2, 10)
    with pytest.raises(HphError) as ex:
        e.log_likelihood()
    assert ex.value.code == _lib.HPH_ERR_NOT_INITIALIZED
    with pytest.raises(HphError) as ex:
        e.set_times(np.zeros(9))
    assert ex.value.code == _lib.HPH_ERR_INVALID_ARGUMENT
    with pytest.raises(HphError) as ex:
        e.set_parameters(np.zeros(5))
    assert ex.value.code == _lib.HPH_ERR_INVALID_ARGUMENT
    with pytest.raises(HphError) as ex:
        e.set_parameters([1, 1, 1, 1, float("nan"), 1])
    assert ex.value.code == _lib.HPH_ERR_INVALID_ARGUMENT
    with pytest.raises(HphError) as ex:
        e.update_locations(np.zeros(3), index=10)
    assert ex.value.code == _lib.HPH_ERR_INVALID_ARGUMENT
    e.close()


def test_r_style_api():
    import hawkes_amd as hph
    locs, times, p = synth.workload("uniform_n1000")
    eng = hph.createEngine(2, 1000, gpu=1)
    with pytest.raises(RuntimeError):
        hph.getLogLikelihood(eng)
    eng = hph.updateLocations(eng, locs)
    eng = hph.setTimesData(eng, times)
    eng = hph.setParameters(eng, p)
    ref = oracle.total(locs, times, p)
    assert rel_err(hph.getLogLikelihood(eng), ref["ll"]) <= LL_TOL
    assert grad_err(hph.getLogLikeGradient(eng), ref).max() <= GRAD_TOL
    assert math.isfinite(hph.getProbsSelfExcite(eng).sum())
