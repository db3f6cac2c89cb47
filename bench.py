"""Benchmark: FP64 log-likelihood + 6-gradient pair-evaluations/s at N ~ 1e5 on one GPU, and the
K = 64 batched-proposal speedup over the plain per-proposal path.

Workload: simulated Hawkes cluster process, 70,000 background events (N = 100,433), times on
[0, 8760) h, locations on a 20 x 20 km square, branching ratio 0.3, offspring sd 0.1 km, mean
delay 24 h (hawkes_amd.synth.workload("cluster_b70k")). A step is one log-likelihood +
gradient evaluation through the engine API (N^2 pair-evaluations), with a fresh parameter
vector each step (log-normal jitter, sd 0.1, around (1, 1/1.6, 1/336, 1, 1, 1)). Inputs are
resident in device memory before timing starts.

Prints one JSON line. Options: --steps K --warmup W --gpus N (only 1 is supported)
--batch K64 (0 skips the batched leg) --cpu-seconds S (0 skips the CPU baseline).
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


def flop_model(n, causal):
    """Algorithmic FLOP of one LL + gradient evaluation: 35 N^2 + 28 C (exp = 20 FLOP, FMA = 2)."""
    return 35.0 * n * n + 28.0 * causal


def load_traffic():
    """HBM bytes per pair-kernel launch from the committed rocprofv3 --pmc summary, if present."""
    path = os.path.join(ROOT, "profiles", "pair_kernel_traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_baseline(locs, times, params, seconds):
    """Threaded + AVX2 restatement of the reference CPU engine on a bounded sample of rows."""
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
    return {
        "value": rows * n / dt,
        "unit": "pair-evals/s",
        "cores": oracle.num_threads(),
        "kind": "port",
        "sample": "%d rows x all %d columns of the same workload (LL + 6-gradient), %.1f s, "
                  "OpenMP rows + AVX2 4-wide FP64 restating the reference's TBB/AVX engine" % (rows, n, dt),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--workload", default="cluster_b70k")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from hawkes_amd import Engine, synth

    locs, times, p0 = synth.workload(args.workload)
    n = len(times)
    causal = synth.causal_pairs(times)
    params = synth.jitter_params(max(args.steps + args.warmup, 1))
    e = Engine(2, n)
    e.update_locations(locs)
    e.set_times(times)
    e.set_parameters(p0)
    e.set_timing(True)
    e.log_likelihood_and_gradient()  # uploads the data; first launch

    for s in range(args.warmup):
        e.set_parameters(params[s])
        e.log_likelihood_and_gradient()

    pair_ms, fin_ms = [], []
    t0 = time.perf_counter()
    for s in range(args.steps):
        e.set_parameters(params[args.warmup + s])
        e.log_likelihood_and_gradient()
        a, b = e.timing()
        pair_ms.append(a)
        fin_ms.append(b)
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / args.steps
    value = n * n / (wall / args.steps)

    pair_avg_ms = float(np.mean(pair_ms))
    flop = flop_model(n, causal)
    achieved = flop / (pair_avg_ms * 1e-3) / 1e12
    roofline = {
        "bound": "valu",
        "achieved": achieved,
        "peak": FP64_VECTOR_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": achieved / FP64_VECTOR_PEAK_TFLOPS,
        "traffic": load_traffic(),
        "kernel": "pair_kernel<2,true>",
        "kernel_ms": pair_avg_ms,
        "finalize_reduce_ms": float(np.mean(fin_ms)),
        "flop_per_launch": flop,
    }

    batched = None
    if args.batch > 0:
        P = synth.jitter_params(args.batch, seed=668)
        e.set_parameters_batch(P)
        e.log_likelihood_batch()
        tb = time.perf_counter()
        for _ in range(args.batch_steps):
            e.log_likelihood_batch()
        bt = (time.perf_counter() - tb) / args.batch_steps
        bvalue = args.batch * n * n / bt
        batched = {"K": args.batch, "value": bvalue, "ms_per_batch": 1e3 * bt,
                   "speedup_vs_plain": bvalue / value}
    e.close()

    cpu = cpu_baseline(locs, times, p0, args.cpu_seconds) if args.cpu_seconds > 0 else None

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "pair-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: Hawkes cluster process, 70k background events, seed 666",
        "config": {"workload": args.workload, "N": n, "causal_pairs": causal, "pairs_per_step": n * n,
                   "batch_K": args.batch},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched": batched,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
