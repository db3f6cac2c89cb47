"""Writes ``full_size_oracle.json``: the oracle's log-likelihood, gradient and summation scales at
the full sizes of the benchmark workloads, for GPU parity tests that would otherwise spend
minutes of host time per vector on the GPU machine.

Cases (all from ``hawkes_amd.synth``, so the GPU tests rebuild the same inputs bit for bit; a
checksum of every input is stored and checked there):

* ``cluster_b70k`` (N = 100,433, the K = 64 batched-proposal workload): P0 and vectors 0, 31
  and 63 of ``synth.jitter_params(64, seed=668)`` (the batch the benchmark times);
* the datasets of ``synth.uneven_datasets(32)`` with N <= 50,000 (the uneven dataset batch):
  vector 0 of each, ``synth.jitter_params(8, seed=669 + d)[0]``.

Values come from ``oracle.total`` in long double (64-bit mantissa, Neumaier sums), the
restatement of src/NewHawkes.hpp:520-547, 636-719 pinned by tests/test_oracle.py.

Usage: python tests/golden/make_full_size.py   (deterministic; takes tens of minutes on 8 cores)
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from hawkes_amd import synth  # noqa: E402


def checksum(locs, times):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(locs, dtype=np.float64).tobytes())
    h.update(np.ascontiguousarray(times, dtype=np.float64).tobytes())
    return h.hexdigest()[:32]


def entry(name, locs, times, params):
    t0 = time.time()
    r = oracle.total(locs, times, params, precision="long")
    print("%s N=%d %.0f s" % (name, len(times), time.time() - t0), flush=True)
    return {"name": name, "N": int(len(times)), "checksum": checksum(locs, times),
            "params": [float(v) for v in params], "ll": r["ll"], "grad": [float(v) for v in r["grad"]],
            "scale": [float(v) for v in r["scale"]]}


def main():
    out = {"cluster_b70k": [], "uneven_datasets": []}
    locs, times, p0 = synth.workload("cluster_b70k")
    P = synth.jitter_params(64, seed=668)
    out["cluster_b70k"].append(dict(entry("p0", locs, times, p0), vector=-1))
    for k in (0, 31, 63):
        out["cluster_b70k"].append(dict(entry("k%d" % k, locs, times, P[k]), vector=k))
    for d, (l, t) in enumerate(synth.uneven_datasets(32)):
        if len(t) <= 50_000:
            out["uneven_datasets"].append(dict(entry("dataset%d" % d, l, t, synth.jitter_params(8, seed=669 + d)[0]),
                                               dataset=d, vector=0))
    with open(os.path.join(HERE, "full_size_oracle.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
