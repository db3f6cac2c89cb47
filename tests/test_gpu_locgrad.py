"""Location gradient d loglik / d locations (hph_get_location_gradient) against the oracle,
through both passes (row-wise, and symmetric with 1 or 2 rows per lane for D <= 3).

Tolerance: component (k, d) within 1e-12 * max(|g_kd|, 1e-3 * S_kd), where S_kd is the
oracle's sum of |pair terms| of that component (location gradients cancel between neighbours).
"""
import numpy as np
import pytest

import oracle
from hawkes_amd import Engine, _lib, synth

pytestmark = pytest.mark.gpu

TOL = 1e-12
KERNELS = {"rowwise": (_lib.PAIR_KERNEL_ROWWISE, 2, 4096), "symmetric": (_lib.PAIR_KERNEL_SYMMETRIC, 2, 4096),
           "symmetric_a1": (_lib.PAIR_KERNEL_SYMMETRIC, 1, 1024),
           "rowwise_narrow": (_lib.PAIR_KERNEL_ROWWISE, 2, 65536), "split": (_lib.PAIR_KERNEL_SPLIT, 2, 4096)}


def engine(locs, times, params, flags=_lib.OPENCL, kernel="rowwise"):
    locs = np.asarray(locs, dtype=np.float64)
    if locs.ndim == 1:
        locs = locs[:, None]
    e = Engine(locs.shape[1], locs.shape[0], flags=flags)
    e.set_pair_config(*KERNELS[kernel])
    e.update_locations(locs)
    e.set_times(times)
    e.set_parameters(params)
    return e


def err(g, ref, scale):
    return np.abs(g - ref) / np.maximum(np.maximum(np.abs(ref), 1e-3 * scale), 1e-300)


def check(locs, times, params, name, flags=_lib.OPENCL, kernel="rowwise"):
    ref, scale = oracle.location_gradient(locs, times, params)
    e = engine(locs, times, params, flags, kernel)
    g = e.location_gradient()
    e.close()
    assert g.shape == ref.shape
    assert err(g, ref, scale).max() <= TOL, (name, err(g, ref, scale).max())
    return g


@pytest.mark.parametrize("kernel", ["rowwise", "symmetric", "split"])
def test_golden_vectors(golden, kernel):
    for c in golden:
        if kernel != "rowwise" and c["D"] > 3:
            continue
        print("golden case", c["name"], kernel, flush=True)
        e = engine(c["locations"], c["times"], c["params"], kernel=kernel)
        g = e.location_gradient()
        e.close()
        _, scale = oracle.location_gradient(c["locations"], c["times"], c["params"])
        gold = np.array(c["loc_grad"])
        assert err(g, gold, scale).max() <= TOL, c["name"]


@pytest.mark.parametrize("D,kernel", [(1, "rowwise"), (2, "rowwise"), (3, "rowwise"), (5, "rowwise"),
                                      (8, "rowwise"), (1, "symmetric"), (2, "symmetric"), (3, "symmetric"),
                                      (2, "symmetric_a1")])
def test_dimensions(D, kernel):
    rng = np.random.default_rng(300 + D)
    n = 1300
    locs = rng.uniform(0, 4, (n, D))
    times = np.sort(rng.uniform(0, 400, n))
    check(locs, times, [0.9, 0.8, 0.05, 0.4, 0.8, 0.6], "D%d" % D, kernel=kernel)


@pytest.mark.parametrize("kernel", ["rowwise", "symmetric", "symmetric_a1"])
def test_cluster_process_and_fp32_engine(kernel):
    """Cluster data (strong excitation); the FP32-bulk engine computes this gradient in FP64."""
    locs, times = synth.cluster_process(2500, side=5.0, seed=303)
    g = check(locs, times, synth.P0, "cluster", kernel=kernel)
    g32 = check(locs, times, synth.P0, "cluster_fp32_engine", flags=_lib.OPENCL | _lib.FLOAT, kernel=kernel)
    assert np.all(np.abs(g - g32) <= 1e-13 * np.maximum(np.abs(g), np.abs(g).max()))


@pytest.mark.parametrize("kernel", ["rowwise", "symmetric"])
def test_unsorted_times_and_ties(kernel):
    rng = np.random.default_rng(304)
    n = 1700
    locs = rng.uniform(0, 3, (n, 2))
    times = np.round(rng.uniform(0, 200, n))
    check(locs, times, [1.1, 0.8, 0.02, 0.3, 0.9, 0.7], "ties", kernel=kernel)


@pytest.mark.parametrize("kernel", ["rowwise", "symmetric", "symmetric_a1", "rowwise_narrow"])
def test_row_counts_around_block_sizes(kernel):
    rng = np.random.default_rng(305)
    for n in (1, 2, 63, 255, 256, 257, 511, 512, 513, 1023, 1025, 1537):
        locs = rng.uniform(0, 3, (n, 2))
        times = np.sort(rng.uniform(0, 50, n))
        check(locs, times, [1.0, 0.9, 0.1, 0.5, 0.6, 0.9], "n%d" % n, kernel=kernel)


@pytest.mark.parametrize("kernel", ["rowwise", "symmetric"])
def test_skip_matches_all_pairs_bitwise(kernel):
    """Tiles are skipped only where every term is exactly +0.0: same bits as every pair formed."""
    locs, times = synth.cluster_process(3000, side=20.0, horizon=43_800.0, seed=306)
    p = [1.0, 1 / 1.6, 1 / 30.0, 0.2, 0.8, 0.9]
    a = engine(locs, times, p, kernel=kernel).location_gradient()
    b = engine(locs, times, p, _lib.OPENCL | _lib.ALL_PAIRS, kernel=kernel).location_gradient()
    assert np.array_equal(a, b)


@pytest.mark.parametrize("kernel", ["rowwise", "symmetric", "symmetric_a1"])
def test_clustered_events_above_the_symmetric_threshold(kernel):
    """N = 5766 clustered events (the size where the engine picks the symmetric pass itself)."""
    locs, times = synth.cluster_process(4000, side=8.0, seed=308)
    check(locs, times, synth.P0, "cluster_5766", kernel=kernel)


@pytest.mark.parametrize("kernel", ["rowwise", "symmetric"])
def test_finite_differences_of_the_gpu_loglik(kernel):
    """Central differences of the GPU log-likelihood under single-location updates
    (updateLocations(i, x, D)) agree with the analytic location gradient; the cached rates are
    invalidated by every update."""
    locs, times = synth.cluster_process(1200, side=4.0, seed=307)
    e = engine(locs, times, synth.P0, kernel=kernel)
    g = e.location_gradient()
    for k in (0, 17, len(times) // 2, len(times) - 1):
        for d in range(2):
            h = 1e-6
            x = locs[k].copy()
            x[d] += h
            e.update_locations(x, index=k)
            up = e.log_likelihood()
            x[d] -= 2 * h
            e.update_locations(x, index=k)
            dn = e.log_likelihood()
            e.update_locations(locs[k], index=k)
            fd = (up - dn) / (2 * h)
            assert abs(fd - g[k, d]) <= 1e-6 * max(1.0, abs(g[k, d])), (k, d, fd, g[k, d])
    assert np.array_equal(e.location_gradient(), g)
    e.close()


@pytest.mark.parametrize("kernel", ["rowwise", "symmetric"])
def test_full_size_invariants(kernel):
    """N = 100,433 (the benchmark workload): translation invariance makes the gradients sum to
    zero over events, rotation invariance makes sum_k (x_k g_ky - y_k g_kx) zero. Both hold to the
    rounding of the per-event sums."""
    locs, times, p = synth.workload("cluster_b70k")
    print("N =", len(times), kernel, flush=True)
    e = engine(locs, times, p, kernel=kernel)
    g = e.location_gradient()
    e.close()
    assert np.all(np.isfinite(g))
    mag = np.abs(g).sum(axis=0)
    assert np.all(np.abs(g.sum(axis=0)) <= 1e-10 * mag), (g.sum(axis=0), mag)
    torque = np.sum(locs[:, 0] * g[:, 1] - locs[:, 1] * g[:, 0])
    assert abs(torque) <= 1e-10 * np.sum(np.abs(locs) * np.abs(g)[:, ::-1]), torque


def test_config2_against_double_oracle():
    """Cluster process with 35,000 background events (N ~ 49.5k) against the double oracle."""
    locs, times, p = synth.workload("cluster_b35k")
    print("N =", len(times), flush=True)
    ref, scale = oracle.location_gradient(locs, times, p, precision="double")
    for kernel in ("rowwise", "symmetric"):
        e = engine(locs, times, p, kernel=kernel)
        g = e.location_gradient()
        e.close()
        assert err(g, ref, scale).max() <= TOL, kernel


def test_errors():
    """Wrong length and null pointer are rejected with HPH_ERR_INVALID_ARGUMENT."""
    import ctypes
    e = engine(np.zeros((4, 2)), np.arange(4.0), synth.P0)
    L = _lib.lib()
    bad = np.zeros(7)
    ptr = bad.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert L.hph_get_location_gradient(e._h, ptr, 7) == _lib.HPH_ERR_INVALID_ARGUMENT
    assert L.hph_get_location_gradient(e._h, None, 8) == _lib.HPH_ERR_INVALID_ARGUMENT
    e.close()
