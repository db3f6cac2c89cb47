"""Split passes (background terms on the culled order with bounding-box skips, then the excitation
terms) on data spread over a domain much wider than the background kernel's reach.

  * against the oracle (FP64: ll 1e-12 relative, gradient 1e-12 of each component's scale,
    probabilities 1e-12 relative or 1e-300 absolute for subnormal ones) in D = 1, 2, 3, with
    many windows of the culled order;
  * against the every-pair evaluation (HPH_FLAG_ALL_PAIRS) to 1e-15: every skipped group pair
    holds only background terms below e^-(45 + ln N);
  * at N ~ 1.5e5 on 100 km: the built-in choice picks the split passes, which agree with the
    symmetric pass (ll 1e-13, gradient 1e-12), FP32-bulk within 1e-6 of FP64, K-batches equal
    to single calls to rounding;
  * at N ~ 1e6, a K = 8 batch stays on the split passes and equals its single calls; a K = 16
    batch, whose partial sums do not fit at once, runs as two launches of 8.
"""
import numpy as np
import pytest

import oracle
from conftest import rel_err
from hawkes_amd import Engine, _lib, synth

pytestmark = pytest.mark.gpu

SPLIT = (_lib.PAIR_KERNEL_SPLIT, 2, 4096)


def engine(locs, times, params, config=SPLIT, flags=_lib.OPENCL):
    locs = np.asarray(locs, dtype=np.float64)
    if locs.ndim == 1:
        locs = locs[:, None]
    e = Engine(locs.shape[1], locs.shape[0], flags=flags)
    if config is not None:
        e.set_pair_config(*config)
    e.update_locations(locs)
    e.set_times(times)
    e.set_parameters(params)
    return e


def spread(n, D, side, horizon, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, side, (n, D)), np.sort(rng.uniform(0.0, horizon, n))


@pytest.mark.parametrize("D,side", [(1, 3000.0), (2, 400.0), (3, 120.0)])
def test_spread_data_against_oracle(D, side):
    locs, times = spread(4500, D, side, 8760.0, 300 + D)
    p = synth.P0
    ref = oracle.total(locs, times, p, probs=True)
    e = engine(locs, times, p)
    ll, g = e.log_likelihood_and_gradient()
    print("D", D, "ll", ll, flush=True)
    assert rel_err(ll, ref["ll"]) <= 1e-12
    scale = ref["scale"][1:]
    assert np.all(np.abs(g - ref["grad"]) <= 1e-12 * np.maximum(np.abs(ref["grad"]), 1e-3 * scale))
    # events far from every earlier one have probabilities in the subnormal range (down to
    # ~1e-321 here), where doubles hold a few bits: those agree to 1e-300 absolute
    np.testing.assert_allclose(e.probs_self_excite(), ref["probs"], rtol=1e-12, atol=1e-300)
    e.close()


@pytest.mark.parametrize("A", [1, 2])
def test_skipped_groups_match_every_pair(A):
    locs, times = synth.cluster_process(7000, side=250.0, horizon=20_000.0, seed=81)
    p = synth.P0
    scale = oracle.total(locs, times, p, precision="double")["scale"][1:]
    out = []
    for flags in (_lib.OPENCL, _lib.OPENCL | _lib.ALL_PAIRS):
        e = engine(locs, times, p, config=(_lib.PAIR_KERNEL_SPLIT, A, 2048), flags=flags)
        ll, g = e.log_likelihood_and_gradient()
        out.append((ll, g, e.probs_self_excite()))
        e.close()
    assert rel_err(out[0][0], out[1][0]) <= 1e-15
    assert np.all(np.abs(out[0][1] - out[1][1]) <= 1e-15 * scale)
    np.testing.assert_allclose(out[0][2], out[1][2], rtol=1e-15, atol=1e-300)


@pytest.mark.parametrize("A", [1, 2])
def test_skipped_supertiles_against_every_pair_and_oracle(A):
    """Target grid 64: chunks of at least 4,096 events, so the culled pass tests whole supertiles
    (4,096 events) against a block's box and skips the far ones with one test. N ~ 20k over 250 km:
    most supertiles are far. Every-pair evaluation (1e-15: ll, gradient of its scale,
    probabilities) and the double oracle (1e-12)."""
    locs, times = synth.cluster_process(14000, side=250.0, horizon=20_000.0, seed=83)
    p = synth.P0
    ref = oracle.total(locs, times, p, precision="double", probs=True)
    scale = ref["scale"][1:]
    out = []
    for flags in (_lib.OPENCL, _lib.OPENCL | _lib.ALL_PAIRS):
        e = engine(locs, times, p, config=(_lib.PAIR_KERNEL_SPLIT, A, 64), flags=flags)
        ll, g = e.log_likelihood_and_gradient()
        assert e.last_pass()[0] == _lib.PAIR_KERNEL_SPLIT
        out.append((ll, g, e.probs_self_excite()))
        e.close()
    print("N=%d ll %r" % (len(times), out[0][0]), flush=True)
    assert rel_err(out[0][0], out[1][0]) <= 1e-15
    assert np.all(np.abs(out[0][1] - out[1][1]) <= 1e-15 * scale)
    np.testing.assert_allclose(out[0][2], out[1][2], rtol=1e-15, atol=1e-300)
    assert rel_err(out[0][0], ref["ll"]) <= 1e-12
    assert np.all(np.abs(out[0][1] - ref["grad"]) <= 1e-12 * np.maximum(np.abs(ref["grad"]), 1e-3 * scale))
    np.testing.assert_allclose(out[0][2], ref["probs"], rtol=1e-12, atol=1e-300)


@pytest.fixture(scope="module")
def wide():
    locs, times = synth.cluster_process(105_000, side=100.0, horizon=17_520.0, seed=91)
    print("N =", len(times), flush=True)
    return locs, times


def test_built_in_choice_and_agreement_with_symmetric(wide):
    locs, times = wide
    p = synth.P0
    auto = engine(locs, times, p, config=None)
    ll_a, g_a = auto.log_likelihood_and_gradient()
    assert auto.tuned_config()[0] == _lib.PAIR_KERNEL_SPLIT
    probs_a = auto.probs_self_excite()
    auto.close()
    sym = engine(locs, times, p, config=(_lib.PAIR_KERNEL_SYMMETRIC, 2, 4096))
    ll_s, g_s = sym.log_likelihood_and_gradient()
    probs_s = sym.probs_self_excite()
    sym.close()
    assert rel_err(ll_a, ll_s) <= 1e-13
    assert np.all(np.abs(g_a - g_s) <= 1e-12 * np.maximum(np.abs(g_s), 1e-3 * np.abs(g_s).max()))
    np.testing.assert_allclose(probs_a, probs_s, rtol=1e-12, atol=1e-300)
    # the benchmark's 20 km workload stays on the time-sorted passes
    l2, t2, p2 = synth.workload("cluster_b70k")
    e2 = engine(l2, t2, p2, config=None)
    e2.log_likelihood()
    assert e2.tuned_config()[0] != _lib.PAIR_KERNEL_SPLIT
    e2.close()


def test_fp32_bulk_and_batches(wide):
    locs, times = wide
    P = synth.jitter_params(4, seed=5)
    e = engine(locs, times, P[0])
    e.set_parameters_batch(P)
    ll_b, g_b = e.log_likelihood_batch()
    for k in range(4):
        # the batch's work items split the columns differently (more vectors, shorter chunks):
        # equal to rounding
        e.set_parameters(P[k])
        ll, g = e.log_likelihood_and_gradient()
        assert rel_err(ll, ll_b[k]) <= 1e-14
        assert np.all(np.abs(g - g_b[k]) <= 1e-13 * np.maximum(np.abs(g), 1e-3 * np.abs(g).max()))
    e.set_parameters(P[0])
    ll64, g64 = e.log_likelihood_and_gradient()
    p64 = e.probs_self_excite()
    e.close()
    f = engine(locs, times, P[0], flags=_lib.OPENCL | _lib.FLOAT)
    ll32, g32 = f.log_likelihood_and_gradient()
    p32 = f.probs_self_excite()
    f.close()
    assert rel_err(ll32, ll64) <= 1e-6
    assert np.all(np.abs(g32 - g64) <= 1e-6 * np.maximum(np.abs(g64), 1e-3 * np.abs(g64).max()))
    assert np.max(np.abs(p32 - p64)) <= 1e-6


def test_k8_batch_at_n1e6_stays_on_the_split_passes():
    """K = 8 vectors of the N ~ 1e6 workload (cluster_b700k, the bench's large leg) in one batch:
    the column partials of the background pass (~23 GB per vector, allocated whole) fit the
    device, so the batch runs the split passes, and each vector equals its single call to
    rounding (measured bitwise)."""
    locs, times, p = synth.workload("cluster_b700k")
    P = synth.jitter_params(8, seed=668)
    print("N=%d K=8 batch on the split passes" % len(times), flush=True)
    e = engine(locs, times, p, config=None)
    e.set_parameters_batch(P)
    ll_b, g_b = e.log_likelihood_batch()
    assert e.last_pass()[0] == _lib.PAIR_KERNEL_SPLIT
    for k in range(8):
        e.set_parameters(P[k])
        ll, g = e.log_likelihood_and_gradient()
        assert e.last_pass()[0] == _lib.PAIR_KERNEL_SPLIT
        assert rel_err(ll, ll_b[k]) <= 1e-14
        assert np.all(np.abs(g - g_b[k]) <= 1e-13 * np.maximum(np.abs(g), 1e-3 * np.abs(g).max()))
    e.close()


def test_k16_batch_at_n1e6_runs_as_launches_that_fit():
    """K = 16 vectors at N ~ 1e6: the background pass's column partials of all 16 (~375 GB) do
    not fit the device, so the batch runs as consecutive launches of 8 vectors on the split
    passes, each bitwise equal to a batch of those 8 vectors."""
    locs, times, p = synth.workload("cluster_b700k")
    P = synth.jitter_params(16, seed=670)
    print("N=%d K=16 batch" % len(times), flush=True)
    e = engine(locs, times, p, config=None)
    e.set_parameters_batch(P)
    ll, g = e.log_likelihood_batch()
    assert e.last_pass()[0] == _lib.PAIR_KERNEL_SPLIT and e.vectors_per_launch() == 8
    for half in (slice(0, 8), slice(8, 16)):
        e.set_parameters_batch(P[half])
        ll8, g8 = e.log_likelihood_batch()
        assert e.vectors_per_launch() == 8
        assert np.array_equal(ll8, ll[half]) and np.array_equal(g8, g[half])
    e.close()


@pytest.mark.parametrize("case", ["one_location", "one_time", "two_clusters"])
def test_degenerate_boxes_against_oracle(case):
    """Boxes that collapse along an axis (every event at one location, or at one time) or split
    the events into two far groups: the skips still hold only negligible terms."""
    rng = np.random.default_rng(77)
    n = 3000
    if case == "one_location":
        locs = np.zeros((n, 2)) + 3.0
        times = np.sort(rng.uniform(0.0, 20_000.0, n))
    elif case == "one_time":
        locs = rng.uniform(0.0, 300.0, (n, 2))
        times = np.full(n, 5.0)
    else:
        locs = np.concatenate([rng.normal(0.0, 1.0, (n // 2, 2)), rng.normal(500.0, 1.0, (n - n // 2, 2))])
        times = np.sort(rng.uniform(0.0, 5_000.0, n))
    p = synth.P0
    ref = oracle.total(locs, times, p, probs=True)
    e = engine(locs, times, p)
    ll, g = e.log_likelihood_and_gradient()
    assert rel_err(ll, ref["ll"]) <= 1e-12, case
    scale = ref["scale"][1:]
    assert np.all(np.abs(g - ref["grad"]) <= 1e-12 * np.maximum(np.abs(ref["grad"]), 1e-3 * scale)), case
    np.testing.assert_allclose(e.probs_self_excite(), ref["probs"], rtol=1e-12, atol=1e-300, err_msg=case)
    e.close()
