"""Benchmark: FP64 log-likelihood + 6-gradient pair-evaluations/s at N ~ 1e5 on one GPU, and the
K = 64 batched-proposal speedup over the plain per-proposal path.

Workload: simulated Hawkes cluster process, 70,000 background events (N = 100,433), times on
[0, 8760) h, locations on a 20 x 20 km square, branching ratio 0.3, offspring sd 0.1 km, mean
delay 24 h (hawkes_amd.synth.workload("cluster_b70k")). A step is one log-likelihood +
gradient evaluation through the engine API (N^2 pair-evaluations), with a fresh parameter
vector each step (log-normal jitter, sd 0.1, around (1, 1/1.6, 1/336, 1, 1, 1)). Inputs are
resident in device memory before timing starts.

Prints one JSON line. Options: --steps K --warmup W --gpus N (only 1 is supported)
--batch K64 (0 skips the batched legs: all six parameters jittered, and tauX, tauT shared)
--cpu-seconds S (0 skips the CPU baseline) --datasets D (0 skips the uneven dataset batch)
--large 0|1 (the N ~ 1e6 leg) --legs (extra legs: all_pairs, fp32, small, locgrad, mh = the
adaptive MH sampler on re-evaluated proposals) --config (a fixed pair-pass configuration instead
of the autotuned one) --only (run the listed legs alone, e.g. for a profile of one of them).
Each leg's timed region is a roctx range "bench:<leg>"; --trace-summary DIR [--bench-json FILE]
reads a `rocprofv3 --kernel-trace --marker-trace` of a run and prints its kernels by leg.

value counts the N^2 ordered pairs the reference forms per call; the engine forms each
background pair once for both events and skips tiles whose background terms are all below
e^-(45 + ln N) (each event's sums change by < 4e-18 relative), or whose excitation terms are all below e^-X_S (excitation_skip_arg; rates change
by < 2e-22 relative) (roofline.pairs_formed gives the pairs actually formed, all_pairs the rate
with no skipping).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "loglik+grad pair-evals/s at N=1e5 FP64; K=64 batched vs plain speedup"
FP64_VECTOR_PEAK_TFLOPS = 78.6
FP32_VECTOR_PEAK_TFLOPS = 157.3


def flop_model(n_background, n_excitation):
    """Algorithmic FLOP of one LL + gradient evaluation: 35 per background pair + 28 per
    excitation pair formed (exp = 20 FLOP, FMA = 2); 35 N^2 + 28 C when every ordered pair is
    formed, 35 N(N+1)/2 + 28 C when the background symmetry is used."""
    return 35.0 * n_background + 28.0 * n_excitation


ZERO_ARG = 746.0  # exp(-746) is exactly +0.0 in FP64


def background_skip_arg(n):
    """FP64 background skip point of the kernels: min(746, 45 + ln N) (terms below e^-(45 + ln N)
    change every event's sums by < 4e-18 relative)."""
    return min(ZERO_ARG, 45.0 + np.log(max(n, 1)))


def excitation_skip_arg(params, n, D=2):
    """Excitation skip point of the kernels in likelihood / gradient passes: min(746,
    50 + ln N + max(0, ln(|cS| / |cB|))), cB = mu0 tauX^D tauT / (2 pi), cS = theta omega sigma^D /
    sqrt(2 pi) (skipped terms change every event's rate by < 2e-22 relative)."""
    sigma, tauX, tauT, omega, theta, mu0 = params
    cBs, cSs = mu0 * tauX ** D * tauT, theta * omega * sigma ** D
    if cSs != 0 and (cBs > 0) != (cSs > 0):
        return ZERO_ARG
    cB = abs(cBs) / (2 * np.pi)
    cS = abs(cSs) / np.sqrt(2 * np.pi)
    if not cB > 0:
        return ZERO_ARG
    r = np.log(cS / cB) if cS > 0 else 0.0
    return min(ZERO_ARG, 50.0 + np.log(max(n, 1)) + max(0.0, r))


def pairs_formed(times, params, skip_zero, rows_per_wave=64, tile=256):
    """Background and excitation pairs the pair kernel forms for sorted times: per (wavefront of
    64 rows, tile of 256 columns), the tile's background is skipped where every term is
    negligible and its excitation where every term is exactly zero; the excitation term is
    formed only for tiles not wholly after the rows (all pairs of the tile in the masked
    diagonal case, as the kernel evaluates then masks)."""
    t = np.sort(np.asarray(times))
    n = t.size
    zero_b = background_skip_arg(n)
    zero_s = excitation_skip_arg(params, n)
    w0 = np.arange(0, n, rows_per_wave)
    w_lo, w_hi = t[w0], t[np.minimum(w0 + rows_per_wave - 1, n - 1)]
    rows = np.minimum(w0 + rows_per_wave, n) - w0
    j0 = np.arange(0, n, tile)
    t_f, t_l = t[j0], t[np.minimum(j0 + tile - 1, n - 1)]
    cols = np.minimum(j0 + tile, n) - j0
    before = t_l[None, :] < w_lo[:, None]
    after = t_f[None, :] >= w_hi[:, None]
    s_on = ~after
    b_on = np.ones_like(before)
    if skip_zero:
        gap = np.where(before, w_lo[:, None] - t_l[None, :],
                       np.where(t_f[None, :] > w_hi[:, None], t_f[None, :] - w_hi[:, None], 0.0))
        b_on = 0.5 * params[2] ** 2 * gap * gap < zero_b
        if params[3] >= 0:
            s_on &= ~(before & (params[3] * gap >= zero_s))
    # the kernel runs whole 64-lane wavefronts; count lanes of real rows only
    pair_counts = rows[:, None] * cols[None, :]
    return float((pair_counts * b_on).sum()), float((pair_counts * s_on).sum())


def pairs_formed_symmetric(times, params, skip_zero, waves=4, rows_per_lane=2, tile=256):
    """Background and excitation pairs the symmetric kernel forms. Each background pair off the
    diagonal blocks is formed once for both events (the N(N+1)/2 of the cost model); per
    (wavefront of 128 rows, earlier tile of 256 columns) the background is skipped where every
    term is negligible and the excitation where every term is exactly zero; the diagonal blocks
    (R x R, R = 512) are paired in both orders."""
    t = np.sort(np.asarray(times))
    n = t.size
    zero_b = background_skip_arg(n)
    zero_s = excitation_skip_arg(params, n)
    wrows = 64 * rows_per_lane
    R = waves * wrows
    nb = (n + R - 1) // R
    gw0 = np.arange(0, nb * R, wrows)
    live = gw0 < n
    gw0 = gw0[live]
    w_lo = t[gw0]
    rows = np.minimum(gw0 + wrows, n) - gw0
    block = gw0 // R
    j0 = np.arange(0, max((nb - 1) * R, 0), tile)
    t_l = t[np.minimum(j0 + tile - 1, n - 1)]
    earlier = j0[None, :] < (block * R)[:, None]
    before = t_l[None, :] < w_lo[:, None]
    gap = np.where(before, w_lo[:, None] - t_l[None, :], 0.0)
    b_on = earlier.copy()
    s_on = earlier.copy()
    if skip_zero:
        b_on &= 0.5 * params[2] ** 2 * gap * gap < zero_b
        if params[3] >= 0:
            s_on &= ~(before & (params[3] * gap >= zero_s))
    nB = float((rows[:, None] * tile * b_on).sum())
    nS = float((rows[:, None] * tile * s_on).sum())
    for P in range(nb):
        m = min(R, n - P * R)
        nB += m * m / 2.0  # ordered pairs of the diagonal block, counted once per unordered pair
        nS += m * m
    return nB, nS


def kernel_name(cfg):
    kind, rows_per_lane, target = cfg
    if kind == 2:
        return "sym_kernel<2,true,%d> (target grid %d)" % (rows_per_lane, target)
    if kind == 3:
        return "split passes: sym_kernel<2,true,%d> background on the culled order + excitation (target grid %d)" % (
            rows_per_lane, target)
    return "pair_kernel<2,true> (target grid %d)" % target


def load_traffic(kind, cfg):
    """HBM bytes per pair-kernel launch from the committed rocprofv3 --pmc summary, if present and
    measured on the pass that ran (kind: 'sym' or 'pair') at the configuration that ran (cfg:
    tuned_config(); the record lists one measurement per target grid)."""
    path = os.path.join(ROOT, "profiles", "pair_kernel_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    if not rec.get("kernel", "").startswith(kind + "_kernel"):
        return None
    for m in rec.get("measurements", []):
        if m.get("rows_per_lane") == cfg[1] and m.get("target_grid") == cfg[2]:
            return m.get("bytes_per_launch")
    return None


class _Ranges:
    """roctx ranges named "bench:<leg>" around each leg's timed region, so that a
    `rocprofv3 --kernel-trace --marker-trace` of a whole run assigns every kernel dispatch to the
    leg that timed it (see trace_summary). A no-op where the roctx library cannot be loaded."""

    def __init__(self):
        import ctypes
        self._lib = None
        for name in ("librocprofiler-sdk-roctx.so.1", "/opt/rocm/lib/librocprofiler-sdk-roctx.so.1", "libroctx64.so"):
            try:
                lib = ctypes.CDLL(name)
                lib.roctxRangePushA.argtypes = [ctypes.c_char_p]
                lib.roctxRangePushA.restype = ctypes.c_int
                lib.roctxRangePop.restype = ctypes.c_int
                self._lib = lib
                break
            except (OSError, AttributeError):
                continue

    def __call__(self, leg):
        ranges = self

        class _Range:
            def __enter__(self):
                if ranges._lib is not None:
                    ranges._lib.roctxRangePushA(("bench:" + leg).encode())

            def __exit__(self, *exc):
                if ranges._lib is not None:
                    ranges._lib.roctxRangePop()
        return _Range()


RANGE = _Ranges()


def trace_summary(trace_dir, bench_json):
    """Per-leg kernel statistics from a rocprofv3 kernel + marker trace of a bench.py run (the
    dispatches that started inside each "bench:<leg>" range), and the headline roofline fraction
    recomputed from the traced average duration of the headline's pair-pass kernel with the FLOP per
    launch that the same run printed (bench_json: that run's JSON line)."""
    import csv
    import glob
    ktrace = sorted(glob.glob(os.path.join(trace_dir, "**", "*kernel_trace.csv"), recursive=True))
    mtrace = sorted(glob.glob(os.path.join(trace_dir, "**", "*marker_api_trace.csv"), recursive=True))
    if not ktrace or not mtrace:
        sys.exit("trace_summary: need *kernel_trace.csv and *marker_api_trace.csv under " + trace_dir)
    ranges = []
    for path in mtrace:
        with open(path) as f:
            for r in csv.DictReader(f):
                label = next((v for v in r.values() if isinstance(v, str) and v.startswith("bench:")), None)
                if label is not None:
                    ranges.append((label[6:], int(r["Start_Timestamp"]), int(r["End_Timestamp"])))
    disp = []
    for path in ktrace:
        with open(path) as f:
            for r in csv.DictReader(f):
                disp.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]))
    disp.sort()
    out = {"trace": trace_dir, "legs": {}}
    for leg, s, e in ranges:
        per = {}
        for a, b, name in disp:
            if s <= a <= e:
                per.setdefault(name, []).append((b - a) / 1e3)
        kernels = sorted(per.items(), key=lambda kv: -sum(kv[1]))
        out["legs"][leg] = [{"kernel": name.split("(")[0], "calls": len(us), "avg_us": float(np.mean(us)),
                             "min_us": float(np.min(us)), "max_us": float(np.max(us)),
                             "total_ms": float(np.sum(us)) / 1e3} for name, us in kernels]
    if bench_json:
        with open(bench_json) as f:
            rec = json.loads(f.read().strip().splitlines()[-1])
        roof = rec.get("roofline") or {}
        head = out["legs"].get("headline")
        if head and roof.get("flop_per_launch"):
            k = head[0]
            traced = roof["flop_per_launch"] / (k["avg_us"] * 1e-6) / 1e12
            out["headline_roofline"] = {
                "kernel": k["kernel"], "calls": k["calls"], "traced_avg_us": k["avg_us"],
                "in_run_hip_event_avg_us": 1e3 * roof["kernel_ms"], "flop_per_launch": roof["flop_per_launch"],
                "traced_tflops": traced, "traced_frac": traced / FP64_VECTOR_PEAK_TFLOPS,
                "bench_frac": roof["frac"], "ms_per_step": rec.get("ms_per_step")}
    print(json.dumps(out, indent=1))


def host_cpu():
    """Model name and logical CPU count of this host (recorded beside every CPU number)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count()}


def cpu_baseline(locs, times, params, seconds):
    """The restated reference CPU engines on bounded row samples of the same workload (LL + 6-gradient,
    every row against all N columns): OpenMP rows + AVX2 on all host threads (the reported value,
    the reference's TBB/AVX engine), and single-core serial scalar and AVX2 4-wide FP64."""
    import oracle
    n = len(times)
    mid = n // 2
    t0 = time.perf_counter()
    rows = 64
    oracle.cpu_baseline(oracle.THREADED, locs, times, params, mid, mid + rows)
    rate = rows * n / max(time.perf_counter() - t0, 1e-6)
    rows = int(min(n - mid, max(64, seconds * rate / n)))
    t0 = time.perf_counter()
    oracle.cpu_baseline(oracle.THREADED, locs, times, params, mid, mid + rows)
    dt = time.perf_counter() - t0
    flavours = {}
    for name, flavour in (("serial", oracle.SERIAL), ("avx", oracle.AVX)):
        r = 8
        t1 = time.perf_counter()
        oracle.cpu_baseline(flavour, locs, times, params, mid, mid + r)
        d1 = time.perf_counter() - t1
        flavours[name] = {"value": r * n / d1, "cores": 1, "rows": r}
    flavours["threaded"] = {"value": rows * n / dt, "cores": oracle.num_threads(), "rows": rows}
    return {
        "value": rows * n / dt,
        "unit": "pair-evals/s",
        "cores": oracle.num_threads(),
        "kind": "port",
        "sample": "%d rows x all %d columns of the same workload (LL + 6-gradient), %.1f s, "
                  "OpenMP rows + AVX2 4-wide FP64 restating the reference's TBB/AVX engine" % (rows, n, dt),
        "flavours": flavours,
        "host": host_cpu(),
    }


def cpu_config1(seconds):
    """The reference's CPU-runnable case (N = 1000 uniform events, 100 repeated log-likelihood
    evaluations) on the three CPU flavours restated from the reference engines: serial scalar
    FP64 (1 core), AVX2 4-wide FP64 (1 core), OpenMP rows + AVX2 (all host threads). The number of
    evaluations is cut below 100 only when the budget `seconds` would not cover the serial one."""
    import oracle
    from hawkes_amd import synth
    locs, times, p0 = synth.workload("uniform_n1000")
    n = len(times)
    t0 = time.perf_counter()
    oracle.cpu_baseline(oracle.SERIAL, locs, times, p0, grad=False)
    t_serial = time.perf_counter() - t0
    evals = int(max(1, min(100, seconds / (2.0 * max(t_serial, 1e-6)))))
    out = {"N": n, "evals": evals, "unit": "pair-evals/s", "host": host_cpu()}
    for name, flavour, cores in (("serial", oracle.SERIAL, 1), ("avx", oracle.AVX, 1),
                                 ("threaded", oracle.THREADED, oracle.num_threads())):
        t0 = time.perf_counter()
        for _ in range(evals):
            oracle.cpu_baseline(flavour, locs, times, p0, grad=False)
        dt = time.perf_counter() - t0
        out[name] = {"value": evals * n * n / dt, "seconds": dt, "cores": cores}
    return out


def dataset_batch(count):
    """Uneven multi-dataset batch: `count` cluster-process datasets (N log-uniform in 2e3..2e5),
    8 jittered parameter vectors each, LL + gradient for all in one graph-replayed launch sequence
    (one time-sorted pass over the datasets, the split passes for those spread over several
    negligible-background distances: split_datasets), FP64 and FP32-bulk, each with the work
    decomposition the batch autotuner picks for the table. Errors: vector 0 of every dataset
    against the FP64 oracle (values committed in tests/golden/full_size_oracle.json by
    tests/golden/make_full_size.py for these same inputs: long double up to N = 50,000, double
    above), and every vector of the FP32-bulk batch against the FP64 batch."""
    from hawkes_amd import _lib, synth
    from hawkes_amd.batch import Batch
    data = synth.uneven_datasets(count)
    params = [synth.jitter_params(8, seed=669 + d) for d in range(count)]
    pairs = sum(8 * len(t) ** 2 for _, t in data)
    out = {"datasets": count, "vectors_per_dataset": 8, "N_min": min(len(t) for _, t in data),
           "N_max": max(len(t) for _, t in data), "pairs_per_launch": float(pairs)}
    refs = {}
    try:
        with open(os.path.join(ROOT, "tests", "golden", "full_size_oracle.json")) as f:
            refs = {c["dataset"]: c for c in json.load(f)["uneven_datasets"]} if count == 32 else {}
    except (OSError, ValueError, KeyError):
        refs = {}
    res = {}
    for name, flags in (("fp64", _lib.OPENCL), ("fp32_bulk", _lib.OPENCL | _lib.FLOAT)):
        b = Batch(2, [len(t) for _, t in data], flags=flags)
        for d, (l, t) in enumerate(data):
            b.set_data(d, l, t)
        b.set_parameters(params)
        tuned_ms = b.autotune(repeats=1)
        b.set_graph(True)
        b.log_likelihood()  # captures the graph
        reps = 2
        with RANGE("datasets_" + name):
            t0 = time.perf_counter()
            for _ in range(reps):
                ll, g = b.log_likelihood()
            wall = (time.perf_counter() - t0) / reps
        res[name] = (ll, g)
        A, target, source = b.config()
        rec = {"value": pairs / wall, "ms_per_launch": 1e3 * wall, "device_ms": b.timing(),
               "config": {"rows_per_lane": A, "target_items": target, "tuned": source == 1,
                          "autotune_best_device_ms": tuned_ms},
               "split_datasets": [{"dataset": d, "N": len(data[d][1])} for d in range(count) if b.last_split(d)]}
        if refs:
            errs = {}
            for d, c in sorted(refs.items()):
                row = 8 * d
                go, sc = np.asarray(c["grad"]), np.asarray(c["scale"][1:])
                errs[str(d)] = {"N": c["N"], "ll": abs(ll[row] - c["ll"]) / abs(c["ll"]),
                                "grad": float(np.max(np.abs(g[row] - go) / np.maximum(np.abs(go), 1e-3 * sc)))}
            rec["rel_err_vs_fp64_oracle_vector0"] = errs
            rec["rel_err_vs_fp64_oracle_max"] = {"ll": max(e["ll"] for e in errs.values()),
                                                 "grad": max(e["grad"] for e in errs.values())}
        out[name] = rec
        b.close()
    ll64, g64 = res["fp64"]
    ll32, g32 = res["fp32_bulk"]
    rel = np.abs(ll32 - ll64) / np.abs(ll64)
    starts = np.cumsum([0] + [8] * count)
    out["fp32_bulk"]["rel_err_ll_vs_fp64_batch_per_dataset_max"] = [
        float(rel[starts[d]:starts[d + 1]].max()) for d in range(count)]
    out["fp32_bulk"]["rel_err_ll_vs_fp64_batch_max"] = float(rel.max())
    return out


def large_dataset(n_leapfrog=20):
    """N ~ 1e6 cluster process (700k background events, 100 x 100 km, 43,800 h): FP32-bulk
    log-likelihood + gradient + self-excitation probabilities with their error against the FP64
    path on the same device, and one 20-step HMC leapfrog trajectory (21 gradient evaluations)
    timed end to end. The engine's built-in choice runs the split passes here (locations spread
    over ~8 times the negligible-background distance); the first, untimed call builds their
    culled order."""
    from hawkes_amd import Engine, _lib, sampler, synth
    locs, times, p = synth.workload("cluster_b700k")
    n = len(times)
    out = {"N": n, "pairs_per_call": float(n) * n}
    res = {}
    for name, flags in (("fp64", _lib.OPENCL), ("fp32_bulk", _lib.OPENCL | _lib.FLOAT)):
        e = Engine(2, n, flags=flags)
        e.update_locations(locs)
        e.set_times(times)
        # the untimed first calls (uploads, culled order, plans, the probability request's
        # buffers) run another tauX, so that the timed call is a full pass and not a
        # re-evaluation from the first call's cached sums
        pw = np.array(p, dtype=float)
        pw[1] *= 1.001
        e.set_parameters(pw)
        e.set_timing(True)
        e.log_likelihood_and_gradient()
        e.probs_self_excite()
        with RANGE("large_" + name):
            t0 = time.perf_counter()
            e.set_parameters(p)
            ll, g = e.log_likelihood_and_gradient()
            t_llg = time.perf_counter() - t0
        assert e.last_evaluation() == _lib.EVAL_FULL
        # probabilities: the excitation sums formed again with a cut relative to each event's own
        # sum (a likelihood pass's cut is bounded relative to the rates, not to tiny
        # probabilities), on the background sums the likelihood call above left in the cache
        with RANGE("large_probs_" + name):
            t0 = time.perf_counter()
            probs = e.probs_self_excite()
            t_probs = time.perf_counter() - t0
        res[name] = (ll, g, probs)
        out[name] = {"ms_loglik_grad": 1e3 * t_llg, "ms_probs_pass": 1e3 * t_probs, "value": float(n) * n / t_llg,
                     "pair_pass": kernel_name(e.tuned_config())}
        if name == "fp32_bulk":
            rng = np.random.default_rng(7)
            theta = np.log(p)
            m0 = rng.normal(size=6)
            t0 = time.perf_counter()
            u0, du0 = sampler._u_and_grad(e, theta)
            # step sized so the first momentum kick stays O(0.01): log-space gradients scale with N
            step = 0.01 / (1.0 + float(np.linalg.norm(du0)))
            th, m, u, du = sampler.leapfrog(e, theta, m0, step, n_leapfrog, (u0, du0))
            out["hmc_leapfrog"] = {"steps": n_leapfrog, "gradient_evaluations": n_leapfrog + 1,
                                   "seconds": time.perf_counter() - t0, "mode": "fp32_bulk", "step_size": step,
                                   "energy_error": float((u + 0.5 * m @ m) - (u0 + 0.5 * m0 @ m0))}
        e.close()
    (l64, g64, p64), (l32, g32, p32) = res["fp64"], res["fp32_bulk"]
    out["fp32_bulk"]["rel_err_ll_vs_fp64"] = abs(l32 - l64) / abs(l64)
    out["fp32_bulk"]["rel_err_grad_vs_fp64"] = float(np.max(np.abs(g32 - g64) / np.maximum(np.abs(g64), 1e-3 * np.max(np.abs(g64)))))
    out["fp32_bulk"]["err_probs_vs_fp64"] = float(np.max(np.abs(p32 - p64) / np.maximum(p64, 1e-3)))
    return out


def mh_sampler(n_iter=300):
    """The reference's adaptive Metropolis-Hastings sampler (random scan over sigma, omega, theta,
    mu0; tauX and tauT fixed, R/Untitled.R:254-422) on the N ~ 5e4 cluster workload: ms per
    iteration (one log-likelihood per proposal) against a full log-likelihood evaluation. The
    engine re-evaluates a proposal from the cached per-event sums: theta and mu0 moves reuse both
    sums, sigma and omega moves reuse the background sums and run the excitation pass."""
    from hawkes_amd import Engine, _lib, sampler, synth
    locs, times, p0 = synth.workload("cluster_b35k")
    n = len(times)
    e = Engine(2, n)
    e.update_locations(locs)
    e.set_times(times)
    e.set_parameters(p0)
    e.log_likelihood()
    # full evaluations: tauX moves force the whole pass
    full = []
    for k in range(5):
        p = p0.copy()
        p[1] *= 1.0 + 0.01 * (k + 1)
        e.set_parameters(p)
        t0 = time.perf_counter()
        e.log_likelihood()
        full.append(time.perf_counter() - t0)
    e.set_parameters(p0)
    e.log_likelihood()
    kinds = {_lib.EVAL_FULL: 0, _lib.EVAL_EXCITATION: 0, _lib.EVAL_CACHED: 0}
    rng = np.random.default_rng(666)
    t0 = time.perf_counter()
    res = sampler.adaptive_mh(e, n_iter, p0, rng=rng)
    wall = time.perf_counter() - t0
    # the mix of evaluation kinds over a second run of the same length (not timed)
    cur = np.array(p0, dtype=float)
    for _ in range(100):
        idx = sampler.MH_INDICES[rng.integers(len(sampler.MH_INDICES))]
        cur[idx] *= float(np.exp(0.1 * rng.normal()))
        e.set_parameters(cur)
        e.log_likelihood()
        kinds[e.last_evaluation()] += 1
    e.close()
    ms_full = 1e3 * float(np.median(full))
    ms_iter = 1e3 * wall / n_iter
    return {"N": n, "iterations": n_iter, "ms_per_iteration": ms_iter, "ms_full_loglik": ms_full,
            "speedup_vs_full": ms_full / ms_iter, "acceptance": res["acceptance"],
            "evaluation_kinds_per_100_moves": {"full": kinds[_lib.EVAL_FULL],
                                               "excitation_pass": kinds[_lib.EVAL_EXCITATION],
                                               "cached_sums": kinds[_lib.EVAL_CACHED]}}


def small_n():
    """The N = 1000 workload (uniform on [0,10]^2 x [0, 8760) h): 100 repeated log-likelihood
    evaluations on the plain path, and K = 64 jittered proposals (LL + gradient) in one batched
    launch against 64 plain calls -- the regime where per-call overhead, not arithmetic, bounds
    the plain path."""
    from hawkes_amd import Engine, synth
    locs, times, p0 = synth.workload("uniform_n1000")
    n = len(times)
    e = Engine(2, n)
    e.update_locations(locs)
    e.set_times(times)
    e.set_parameters(p0)
    e.log_likelihood()
    P = synth.jitter_params(100, seed=670)
    t0 = time.perf_counter()
    for k in range(100):
        e.set_parameters(P[k])
        e.log_likelihood()
    t_ll = (time.perf_counter() - t0) / 100
    Q = P[:64]
    for k in range(64):  # warm the plain LL + gradient path
        e.set_parameters(Q[k])
        e.log_likelihood_and_gradient()
    t0 = time.perf_counter()
    for k in range(64):
        e.set_parameters(Q[k])
        e.log_likelihood_and_gradient()
    t_plain = time.perf_counter() - t0
    e.set_parameters_batch(Q)
    e.log_likelihood_batch()
    reps = 10
    t0 = time.perf_counter()
    for _ in range(reps):
        e.log_likelihood_batch()
    t_batch = (time.perf_counter() - t0) / reps
    e.close()
    return {"N": n, "loglik_ms_per_eval_100_evals": 1e3 * t_ll, "loglik_pair_evals_per_s": n * n / t_ll,
            "K64_plain_ms": 1e3 * t_plain, "K64_batched_ms": 1e3 * t_batch, "K64_speedup": t_plain / t_batch}


def batched_leg(e, K, reps, n, plain_value):
    """K = 64 jittered vectors (all six parameters, synth.jitter_params(K, seed=668)) evaluated in
    one launch (LL + gradient), against the plain per-proposal path (one full evaluation per
    vector: the headline's rate)."""
    from hawkes_amd import synth
    P = synth.jitter_params(K, seed=668)
    e.set_parameters_batch(P)
    e.log_likelihood_batch()
    with RANGE("batched"):
        tb = time.perf_counter()
        for _ in range(reps):
            e.log_likelihood_batch()
        bt = (time.perf_counter() - tb) / reps
    bvalue = K * n * n / bt
    out = {"K": K, "value": bvalue, "ms_per_batch": 1e3 * bt, "evaluation": e.last_evaluation()}
    if plain_value:
        out["speedup_vs_plain"] = bvalue / plain_value
    return out


def shared_tau_leg(e, K, reps, n, plain_value):
    """K vectors that share tauX and tauT (the reference sampler's scan set: sigma, omega, theta,
    mu0 jittered, R/Untitled.R:326; synth.jitter_params(K, seed=671) with tauX, tauT reset to P0):
    the batch forms the background sums once and runs one excitation pass per vector. Reported
    against the plain per-proposal path (one full evaluation per vector: the headline's rate) and
    against K single calls in sequence (each of which re-evaluates from the previous call's cached
    background sums)."""
    from hawkes_amd import _lib, synth
    P = synth.jitter_params(K, seed=671)
    P[:, 1:3] = synth.P0[1:3]
    e.set_parameters_batch(P)
    e.log_likelihood_batch()
    assert e.last_evaluation() == _lib.EVAL_EXCITATION
    with RANGE("shared_tau"):
        tb = time.perf_counter()
        for _ in range(reps):
            e.log_likelihood_batch()
        bt = (time.perf_counter() - tb) / reps
    e.set_parameters(P[-1])
    e.log_likelihood_and_gradient()
    with RANGE("shared_tau_single_calls"):
        ts = time.perf_counter()
        for k in range(K):
            e.set_parameters(P[k])
            e.log_likelihood_and_gradient()
        st = time.perf_counter() - ts
    bvalue = K * n * n / bt
    out = {"K": K, "value": bvalue, "ms_per_batch": 1e3 * bt, "ms_sequential_single_calls": 1e3 * st,
           "speedup_vs_sequential_single_calls": st / bt}
    if plain_value:
        out["speedup_vs_plain"] = bvalue / plain_value
    return out


LEGS = ("headline", "all_pairs", "fp32", "batched", "shared_tau", "datasets", "large", "small", "locgrad", "mh",
        "cpu")


def formed_pairs(kind, rows_per_lane, times, params, skip):
    """(background, excitation) pairs formed by the pass that ran, or None for the split passes
    (their bounding-box skips are not modelled here, so no FLOP count is claimed for them)."""
    from hawkes_amd import _lib
    if kind == _lib.PAIR_KERNEL_SYMMETRIC:
        return pairs_formed_symmetric(times, params, skip, rows_per_lane=rows_per_lane)
    if kind == _lib.PAIR_KERNEL_ROWWISE:
        return pairs_formed(times, params, skip)
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--workload", default="cluster_b70k")
    ap.add_argument("--datasets", type=int, default=32, help="uneven dataset batch size (0 skips the leg)")
    ap.add_argument("--large", type=int, default=1, help="1: N ~ 1e6 FP32-bulk + HMC leg, 0: skip")
    ap.add_argument("--legs", default="all_pairs,fp32,small,locgrad,mh",
                    help="comma list of the extra legs to run (all_pairs, fp32, small, locgrad, mh) or 'none'")
    ap.add_argument("--only", default="all",
                    help="comma list of legs to run, for profiling one of them alone: " + ", ".join(LEGS) +
                         " (default all: the headline and every leg the other options enable)")
    ap.add_argument("--trace-summary", default=None, metavar="DIR",
                    help="no run: summarise the rocprofv3 kernel + marker trace under DIR of a bench.py run by "
                         "leg (with --bench-json: that run's JSON line, for the headline roofline)")
    ap.add_argument("--bench-json", default=None)
    ap.add_argument("--config", default="auto",
                    help="pair-pass configuration: 'auto' (autotuned) or kind,rows_per_lane,target_grid with "
                         "kind sym or rowwise, e.g. sym,2,8192 (no autotuning; for profiling one kernel)")
    args = ap.parse_args()
    if args.trace_summary:
        trace_summary(args.trace_summary, args.bench_json)
        return
    legs = set() if args.legs == "none" else set(args.legs.split(","))
    only = None if args.only == "all" else set(args.only.split(","))
    if only is not None and not only <= set(LEGS):
        ap.error("--only takes legs among " + ", ".join(LEGS))

    def want(leg, enabled=True):
        return (only is None and enabled) or (only is not None and leg in only)

    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from hawkes_amd import Engine, _lib, synth

    locs, times, p0 = synth.workload(args.workload)
    n = len(times)
    causal = synth.causal_pairs(times)
    params = synth.jitter_params(max(args.steps, 3) + args.warmup)

    def engine(flags):
        e = Engine(2, n, flags=flags)
        e.update_locations(locs)
        e.set_times(times)
        e.set_parameters(p0)
        e.set_timing(True)
        if args.config != "auto":
            kind, rows_per_lane, target = args.config.split(",")
            e.set_pair_config(_lib.PAIR_KERNEL_SYMMETRIC if kind == "sym" else _lib.PAIR_KERNEL_ROWWISE,
                              int(rows_per_lane), int(target))
        e.log_likelihood_and_gradient()  # uploads the data; first launch
        if args.config == "auto":
            e.autotune(repeats=2)        # picks the pair-pass configuration for this shape
        return e

    def run(flags, steps, leg):
        e = engine(flags)
        for s in range(args.warmup):
            e.set_parameters(params[s])
            e.log_likelihood_and_gradient()
        pair_ms, fin_ms, results, kinds, passes = [], [], [], [], []
        with RANGE(leg):
            t0 = time.perf_counter()
            for s in range(steps):
                e.set_parameters(params[args.warmup + s])
                results.append(e.log_likelihood_and_gradient())
                a, b = e.timing()
                pair_ms.append(a)
                fin_ms.append(b)
                kinds.append(e.last_evaluation())
                passes.append(e.last_pass())
            wall = time.perf_counter() - t0
        # every step jitters all six parameters: each one is a full pass, none a re-evaluation
        assert all(k == _lib.EVAL_FULL for k in kinds), kinds
        assert len(set(passes)) == 1, passes
        run.results = results
        kind, base2 = passes[0]
        rows_per_lane = e.tuned_config()[1]
        skip = not (flags & _lib.ALL_PAIRS)
        formed = [formed_pairs(kind, rows_per_lane, times, params[args.warmup + s], skip) for s in range(steps)]
        return e, wall, pair_ms, fin_ms, formed

    out = {
        "metric": METRIC, "value": None, "unit": "pair-evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: Hawkes cluster process, 70k background events, seed 666",
        "config": {"workload": args.workload, "N": n, "causal_pairs": causal, "pairs_per_step": n * n,
                   "batch_K": args.batch},
    }

    e = None
    value = None
    fp64_results = None
    if want("headline"):
        e, wall, pair_ms, fin_ms, formed = run(_lib.OPENCL, args.steps, "headline")
        fp64_results = run.results
        out["ms_per_step"] = 1e3 * wall / args.steps
        value = n * n / (wall / args.steps)
        out["value"] = value
        kind, base2 = e.last_pass()
        pair_avg_ms = float(np.mean(pair_ms))
        every_pair_flop = flop_model(n * (n + 1) / 2.0, causal)
        roofline = {
            "bound": "valu",
            "achieved": None, "peak": FP64_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": None,
            "traffic": load_traffic("sym" if kind == _lib.PAIR_KERNEL_SYMMETRIC else "pair", e.tuned_config()),
            "traffic_source": "not measured in this run: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of the same "
                              "kernel, workload and configuration, calibrated by a plain copy, in "
                              "profiles/pair_kernel_traffic.json (null when no measurement matches)",
            "kernel": kernel_name(e.tuned_config()),
            "pass": {"kind": kind, "base2_form": base2},
            "kernel_ms": pair_avg_ms,
            "kernel_ms_min_max": [float(np.min(pair_ms)), float(np.max(pair_ms))],
            "kernel_ms_per_step": [round(float(m), 4) for m in pair_ms],
            "finalize_reduce_ms": float(np.mean(fin_ms)),
            "timing": "HIP events around the pair pass on the engine's stream, averaged over the timed steps",
            "every_pair_flop": every_pair_flop,
            "every_pair_tflops_effective": every_pair_flop / (pair_avg_ms * 1e-3) / 1e12,
        }
        if all(f is not None for f in formed):
            flops = [flop_model(b, x) for b, x in formed]
            # mean FLOP per launch over the mean launch duration: what a profile's average
            # duration of the same dispatches reproduces
            achieved = float(np.mean(flops)) / (pair_avg_ms * 1e-3) / 1e12
            nb = float(np.mean([b for b, _ in formed]))
            nx = float(np.mean([x for _, x in formed]))
            roofline.update({
                "achieved": achieved, "frac": achieved / FP64_VECTOR_PEAK_TFLOPS,
                "flop_per_launch": float(np.mean(flops)),
                "flop_model": "35 per background pair + 28 per excitation pair formed (exp = 20 FLOP, FMA = 2); "
                              "pairs counted from the pass's own tile decisions",
                "pairs_formed": {"background": nb, "excitation": nx},
                "formed_fraction": {"background_of_N(N+1)/2": nb / (n * (n + 1) / 2.0), "excitation_of_C": nx / causal},
            })
        else:
            roofline["note"] = "split passes: formed pairs not modelled, no FLOP rate claimed"
        out["roofline"] = roofline

    # Every pair formed (no tile skipping): the same results to rounding, more work.
    if want("all_pairs", "all_pairs" in legs):
        ea, wall_a, pair_a, _, formed_a = run(_lib.OPENCL | _lib.ALL_PAIRS, max(3, args.steps // 4), "all_pairs")
        steps_a = len(pair_a)
        rec = {"value": n * n / (wall_a / steps_a), "ms_per_step": 1e3 * wall_a / steps_a,
               "kernel_ms": float(np.mean(pair_a)), "kernel": kernel_name(ea.tuned_config())}
        if formed_a[0] is not None:
            rec["roofline_frac"] = flop_model(*formed_a[0]) / (float(np.mean(pair_a)) * 1e-3) / 1e12 / \
                FP64_VECTOR_PEAK_TFLOPS
        out["all_pairs"] = rec
        ea.close()

    # FP32-bulk mode: the background exponential in FP32, everything else FP64; error against the
    # FP64 results above (when they ran).
    if want("fp32", "fp32" in legs):
        ef, wall_f, pair_f, _, formed_f = run(_lib.OPENCL | _lib.FLOAT, args.steps, "fp32")
        rec = {"value": n * n / (wall_f / args.steps), "ms_per_step": 1e3 * wall_f / args.steps,
               "kernel_ms": float(np.mean(pair_f)), "kernel": kernel_name(ef.tuned_config())}
        if formed_f[0] is not None:
            # same cost model; the FP32-bulk pass is mixed (background exp in FP32, the rest FP64), so
            # its rate is stated against both vector peaks
            ach_f = float(np.mean([flop_model(b, x) / (ms * 1e-3) for (b, x), ms in zip(formed_f, pair_f)])) / 1e12
            rec.update({"achieved_tflops": ach_f, "frac_fp64_peak": ach_f / FP64_VECTOR_PEAK_TFLOPS,
                        "frac_fp32_peak": ach_f / FP32_VECTOR_PEAK_TFLOPS})
        if fp64_results is not None:
            rec["rel_err_ll_vs_fp64"] = max(abs(a[0] - b[0]) / abs(b[0]) for a, b in zip(run.results, fp64_results))
            rec["rel_err_grad_vs_fp64"] = max(
                float(np.max(np.abs(a[1] - b[1]) / np.maximum(np.abs(b[1]), 1e-3 * np.max(np.abs(b[1])))))
                for a, b in zip(run.results, fp64_results))
        out["fp32_bulk"] = rec
        ef.close()

    need_e = want("locgrad", "locgrad" in legs) or want("batched", args.batch > 0) or \
        want("shared_tau", args.batch > 0)
    if e is None and need_e:
        e = engine(_lib.OPENCL)

    # Location gradient d ll / d x (N x D): after a parameter change it runs a natural-form rate
    # pass (1/lambda) and the gradient pass ("cold"); at unchanged state the rates are reused.
    if want("locgrad", "locgrad" in legs):
        e.set_parameters(p0)
        e.location_gradient()
        reps = 3
        cold, cold_ms = [], []
        with RANGE("locgrad"):
            for r in range(reps):
                e.set_parameters(params[r])
                t0 = time.perf_counter()
                e.location_gradient()
                cold.append(time.perf_counter() - t0)
                cold_ms.append(e.timing())
        with RANGE("locgrad_rates_reused"):
            t0 = time.perf_counter()
            for _ in range(reps):
                e.location_gradient()
            warm = (time.perf_counter() - t0) / reps
        lt = float(np.mean(cold))
        out["location_gradient"] = {
            "ms_per_call": 1e3 * lt, "rate_pass_ms": float(np.mean([a for a, _ in cold_ms])),
            "gradient_pass_ms": float(np.mean([b for _, b in cold_ms])),
            "ms_per_call_rates_reused": 1e3 * warm, "pair_evals_per_s": n * n / lt}

    if want("batched", args.batch > 0):
        out["batched"] = batched_leg(e, args.batch, args.batch_steps, n, value)
    if want("shared_tau", args.batch > 0):
        out["shared_tau_batch"] = shared_tau_leg(e, args.batch, args.batch_steps, n, value)
    if e is not None:
        e.close()

    if want("datasets", args.datasets > 0):
        out["dataset_batch"] = dataset_batch(args.datasets)
    if want("large", bool(args.large)):
        out["large_n1e6"] = large_dataset()
    if want("small", "small" in legs):
        out["small_n1000"] = small_n()
    if want("mh", "mh" in legs):
        out["mh_sampler"] = mh_sampler()
    if want("cpu", args.cpu_seconds > 0):
        out["cpu_baseline"] = cpu_baseline(locs, times, p0, max(args.cpu_seconds, 1.0))
        out["cpu_config1"] = cpu_config1(max(args.cpu_seconds, 1.0))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
