"""Deterministic synthetic event data for the benchmark workloads.

All draws use numpy's PCG64 generator with fixed seeds (base 666, the seed convention of
src/benchmark.cpp:59 and R/Untitled.R:89), so every workload is reproducible.

* ``uniform_events``: N events, locations i.i.d. uniform on [0, L]^D, times i.i.d. uniform on
  [0, T) and sorted (the N = 1000 log-likelihood workload).
* ``cluster_process``: a simulated Hawkes cluster process. B background events with times
  uniform on [0, T) and locations uniform on [0, L]^2; every event (recursively) has
  Poisson(branching) children, each delayed by Exp(mean delay) and kept if before T, displaced
  by N(0, sd^2) per axis. Merged and sorted by time; the realised N is data-dependent.
* ``jitter_params``: K parameter vectors P0 * exp(sd * z), z ~ N(0, I6).
"""
import numpy as np

P0 = np.array([1.0, 1.0 / 1.6, 1.0 / 336.0, 1.0, 1.0, 1.0])


def uniform_events(n, side=10.0, horizon=8760.0, dim=2, seed=666):
    rng = np.random.Generator(np.random.PCG64(seed))
    locs = rng.uniform(0.0, side, size=(n, dim))
    times = np.sort(rng.uniform(0.0, horizon, size=n))
    return locs, times


def cluster_process(background, side=20.0, horizon=8760.0, branching=0.3, offspring_sd=0.1,
                    mean_delay=24.0, seed=666):
    rng = np.random.Generator(np.random.PCG64(seed))
    t = rng.uniform(0.0, horizon, size=background)
    xy = rng.uniform(0.0, side, size=(background, 2))
    all_t, all_xy = [t], [xy]
    parents_t, parents_xy = t, xy
    while parents_t.size:
        counts = rng.poisson(branching, size=parents_t.size)
        if counts.sum() == 0:
            break
        idx = np.repeat(np.arange(parents_t.size), counts)
        ct = parents_t[idx] + rng.exponential(mean_delay, size=idx.size)
        cxy = parents_xy[idx] + rng.normal(0.0, offspring_sd, size=(idx.size, 2))
        keep = ct < horizon
        ct, cxy = ct[keep], cxy[keep]
        all_t.append(ct)
        all_xy.append(cxy)
        parents_t, parents_xy = ct, cxy
    times = np.concatenate(all_t)
    locs = np.concatenate(all_xy)
    order = np.argsort(times, kind="stable")
    return locs[order], times[order]


def jitter_params(K, sd=0.1, base=P0, seed=667):
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.asarray(base)[None, :] * np.exp(sd * rng.normal(size=(K, 6)))


def causal_pairs(times):
    """Number of ordered pairs (i, j) with t_j < t_i."""
    t = np.sort(np.asarray(times))
    return int(np.searchsorted(t, t, side="left").sum())


def uneven_datasets(count=32, lo=2_000, hi=200_000, seed=666000):
    """Independent cluster-process datasets with N log-uniform in [lo, hi] (realised N varies),
    square side sqrt(N/100) km, background count 0.7 N; dataset d uses seed + d."""
    import math
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for d in range(count):
        n_target = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        locs, times = cluster_process(max(1, int(round(0.7 * n_target))), side=math.sqrt(n_target / 100.0),
                                      seed=seed + d)
        out.append((locs, times))
    return out


# Named workloads (sizes and generators of the benchmark configurations).
def workload(name):
    if name == "uniform_n1000":
        locs, times = uniform_events(1000)
        return locs, times, P0.copy()
    if name == "cluster_b35k":
        locs, times = cluster_process(35_000)
        return locs, times, P0.copy()
    if name == "cluster_b70k":
        locs, times = cluster_process(70_000)
        return locs, times, P0.copy()
    if name == "cluster_b700k":
        locs, times = cluster_process(700_000, side=100.0, horizon=43_800.0)
        return locs, times, P0.copy()
    raise KeyError(name)
