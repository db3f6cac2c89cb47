"""Full-size parity: the K = 64 batched-proposal workload (cluster process, 70,000 background
events, N = 100,433) at its own shape, and the uneven dataset batch (32 datasets, N 2,127 to
198,853, 8 vectors each).

References: the oracle's long-double values in tests/golden/full_size_oracle.json (written by
tests/golden/make_full_size.py from the same synthetic inputs; a checksum of each input is
checked here), and the engine's own single-vector calls.

Tolerances: against the oracle, ll 1e-12 relative and gradient component k within
1e-12 * max(|g_k|, 1e-3 S_k) (S_k = the oracle's sum of |per-event terms|); batched against single
calls (same kernels, a different work decomposition and so another summation order), ll 1e-14
and gradient within the FP64 parity bar, 1e-12 * max(|g_k|, 1e-3 S_k) with S_k of P0: the
gradient components are sums of cancelling per-event terms, and in the base-2 form the tauX
component is formed as a difference of two moment sums, so their rounding (measured ~3e-15 S_k
between decompositions) is relative to S_k, not to |g_k|.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import ROOT, rel_err
from hawkes_amd import Engine, _lib, synth
from hawkes_amd.batch import Batch

pytestmark = pytest.mark.gpu

LL_TOL, GRAD_TOL = 1e-12, 1e-12


def checksum(locs, times):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(locs, dtype=np.float64).tobytes())
    h.update(np.ascontiguousarray(times, dtype=np.float64).tobytes())
    return h.hexdigest()[:32]


@pytest.fixture(scope="module")
def fixture():
    with open(os.path.join(ROOT, "tests", "golden", "full_size_oracle.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def workload(fixture):
    locs, times, p0 = synth.workload("cluster_b70k")
    for c in fixture["cluster_b70k"]:
        assert c["checksum"] == checksum(locs, times) and c["N"] == len(times)
    return locs, times, p0


def single_err(gb, g, scale):
    """Gradient difference of a batched and a single call, relative to max(|g_k|, 1e-3 S_k)."""
    return float(np.max(np.abs(np.asarray(gb) - g) / np.maximum(np.abs(g), 1e-3 * np.asarray(scale))))


def oracle_err(ll, g, ref):
    go, scale = np.asarray(ref["grad"]), np.asarray(ref["scale"][1:])
    return rel_err(ll, ref["ll"]), np.abs(np.asarray(g) - go) / np.maximum(np.abs(go), 1e-3 * scale)


def test_p0_against_oracle(fixture, workload):
    locs, times, p0 = workload
    ref = fixture["cluster_b70k"][0]
    assert ref["vector"] == -1 and np.array_equal(ref["params"], p0)
    for kind in (_lib.PAIR_KERNEL_AUTO, _lib.PAIR_KERNEL_ROWWISE):
        e = Engine(2, len(times))
        e.set_pair_config(kind, 2, 4096)
        e.update_locations(locs)
        e.set_times(times)
        e.set_parameters(p0)
        ll, g = e.log_likelihood_and_gradient()
        passed = e.last_pass()
        e.close()
        el, eg = oracle_err(ll, g, ref)
        print("P0 pass %s: ll %.2e grad %.2e" % (passed, el, eg.max()), flush=True)
        assert el <= LL_TOL and eg.max() <= GRAD_TOL, (passed, el, eg)


def test_k64_batch_at_full_size(fixture, workload):
    """The benchmark's K = 64 batch (synth.jitter_params(64, seed=668)) in one launch: every vector
    against its own single-vector call, and vectors 0, 31 and 63 against the oracle."""
    locs, times, _ = workload
    P = synth.jitter_params(64, seed=668)
    e = Engine(2, len(times))
    e.update_locations(locs)
    e.set_times(times)
    e.set_parameters_batch(P)
    ll_b, g_b = e.log_likelihood_batch()
    print("batch done", flush=True)
    for c in fixture["cluster_b70k"][1:]:
        k = c["vector"]
        assert np.array_equal(c["params"], P[k])
        el, eg = oracle_err(ll_b[k], g_b[k], c)
        print("k=%d vs oracle: ll %.2e grad %.2e" % (k, el, eg.max()), flush=True)
        assert el <= LL_TOL and eg.max() <= GRAD_TOL, (k, el, eg)
    worst_ll, worst_g = 0.0, 0.0
    for k in range(64):
        e.set_parameters(P[k])
        ll, g = e.log_likelihood_and_gradient()
        assert e.last_evaluation() == _lib.EVAL_FULL
        worst_ll = max(worst_ll, rel_err(ll_b[k], ll))
        worst_g = max(worst_g, single_err(g_b[k], g, fixture["cluster_b70k"][0]["scale"][1:]))
    e.close()
    print("batch vs single calls: ll %.2e grad %.2e" % (worst_ll, worst_g), flush=True)
    assert worst_ll <= 1e-14 and worst_g <= 1e-12


def test_shared_tau_batch_at_full_size(fixture, workload):
    """64 vectors sharing tauX and tauT (sigma, omega, theta, mu0 jittered: the benchmark's
    shared-tau batch): one background pass and 64 excitation passes. Vectors 0 and 63 against the
    oracle, every vector against a full single-vector evaluation (ll 1e-14, gradient 1e-12 as
    above)."""
    locs, times, _ = workload
    S = synth.jitter_params(64, seed=671)
    S[:, 1:3] = synth.P0[1:3]
    e = Engine(2, len(times))
    e.update_locations(locs)
    e.set_times(times)
    e.set_parameters_batch(S)
    ll_b, g_b = e.log_likelihood_batch()
    assert e.last_evaluation() == _lib.EVAL_EXCITATION
    e.close()
    for c in fixture["cluster_b70k_shared_tau"]:
        k = c["vector"]
        assert np.array_equal(c["params"], S[k])
        el, eg = oracle_err(ll_b[k], g_b[k], c)
        print("shared-tau k=%d vs oracle: ll %.2e grad %.2e" % (k, el, eg.max()), flush=True)
        assert el <= LL_TOL and eg.max() <= GRAD_TOL, (k, el, eg)
    f = Engine(2, len(times))
    f.set_times(times)
    worst_ll, worst_g = 0.0, 0.0
    for k in range(64):
        f.update_locations(locs)  # a data change: the next call is a full pass
        f.set_parameters(S[k])
        ll, g = f.log_likelihood_and_gradient()
        assert f.last_evaluation() == _lib.EVAL_FULL
        worst_ll = max(worst_ll, rel_err(ll_b[k], ll))
        worst_g = max(worst_g, single_err(g_b[k], g, fixture["cluster_b70k"][0]["scale"][1:]))
    f.close()
    print("shared-tau batch vs full single calls: ll %.2e grad %.2e" % (worst_ll, worst_g), flush=True)
    assert worst_ll <= 1e-14 and worst_g <= 1e-12


@pytest.fixture(scope="module")
def uneven():
    data = synth.uneven_datasets(32)
    params = [synth.jitter_params(8, seed=669 + d) for d in range(32)]
    return data, params


@pytest.mark.parametrize("flags", [_lib.OPENCL, _lib.OPENCL | _lib.FLOAT], ids=["fp64", "fp32_bulk"])
def test_uneven_dataset_batch_at_full_size(fixture, uneven, flags):
    """The 32-dataset batch (one graph-replayed launch sequence): vector 0 of every dataset against
    the oracle (1e-12 in FP64, 1e-6 for ll and gradient in FP32-bulk mode), and every vector
    against a single engine's FP64 call (FP64: ll 1e-13, gradient 1e-12 of max(|g_k|, 1e-3 S_k)
    with S_k of the dataset's vector 0 -- the single engines of the smaller datasets take the
    row-wise pass in the natural form, the batch the symmetric pass in the base-2 form; FP32-bulk:
    1e-6)."""
    data, params = uneven
    refs = {c["dataset"]: c for c in fixture["uneven_datasets"]}
    assert sorted(refs) == list(range(32))
    b = Batch(2, [len(t) for _, t in data], flags=flags)
    for d, (l, t) in enumerate(data):
        b.set_data(d, l, t)
    b.set_parameters(params)
    b.set_graph(True)
    ll_b, g_b = b.log_likelihood()
    ll_r, g_r = b.log_likelihood()  # a replay
    assert np.array_equal(ll_b, ll_r) and np.array_equal(g_b, g_r)
    split = [d for d in range(32) if b.last_split(d)]
    b.close()
    fp32 = bool(flags & _lib.FLOAT)
    tol_o = 1e-6 if fp32 else 1e-12
    tol_ll, tol_g = (1e-6, 1e-6) if fp32 else (1e-13, 1e-12)
    worst = {"oracle_ll": 0.0, "oracle_g": 0.0, "single_ll": 0.0, "single_g": 0.0}
    row = 0
    for d, (l, t) in enumerate(data):
        c = refs[d]
        assert c["checksum"] == checksum(l, t) and np.array_equal(c["params"], params[d][0])
        el, eg = oracle_err(ll_b[row], g_b[row], c)
        worst["oracle_ll"] = max(worst["oracle_ll"], el)
        worst["oracle_g"] = max(worst["oracle_g"], float(eg.max()))
        assert el <= tol_o and eg.max() <= tol_o, (d, el, eg)
        e = Engine(2, len(t))
        e.update_locations(l)
        e.set_times(t)
        for k in range(8):
            e.set_parameters(params[d][k])
            ll, g = e.log_likelihood_and_gradient()
            sl = rel_err(ll_b[row + k], ll)
            sg = single_err(g_b[row + k], g, c["scale"][1:])
            worst["single_ll"] = max(worst["single_ll"], sl)
            worst["single_g"] = max(worst["single_g"], sg)
            assert sl <= tol_ll and sg <= tol_g, (d, k, sl, sg)
        e.close()
        row += 8
    print("uneven batch (%s, split datasets %s): %s" % ("fp32" if fp32 else "fp64", split, worst), flush=True)
