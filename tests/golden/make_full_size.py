"""Writes ``full_size_oracle.json``: the oracle's log-likelihood, gradient and summation scales at
the full sizes of the benchmark workloads, for GPU parity tests that would otherwise spend
minutes of host time per vector on the GPU machine.

Cases (all from ``hawkes_amd.synth``, so the GPU tests rebuild the same inputs bit for bit; a
checksum of every input is stored and checked there):

* ``cluster_b70k`` (N = 100,433, the K = 64 batched-proposal workload): P0 and vectors 0, 31
  and 63 of ``synth.jitter_params(64, seed=668)`` (the batch the benchmark times);
* ``cluster_b70k_shared_tau``: vectors 0 and 63 of ``synth.jitter_params(64, seed=671)`` with
  tauX and tauT set to P0's (the batch of vectors sharing tauX and tauT);
* every dataset of ``synth.uneven_datasets(32)`` (the uneven dataset batch): vector 0 of each,
  ``synth.jitter_params(8, seed=669 + d)[0]``.

Values come from ``oracle.total``, the restatement of src/NewHawkes.hpp:520-547, 636-719 pinned by
tests/test_oracle.py: in long double (64-bit mantissa, Neumaier sums) up to N = 50,000 and for
the N = 100,433 workload, in double (compensated sums) for the larger datasets of the batch
(``precision`` of each entry).

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


def entry(name, locs, times, params, precision="long"):
    t0 = time.time()
    r = oracle.total(locs, times, params, precision=precision)
    print("%s N=%d %.0f s" % (name, len(times), time.time() - t0), flush=True)
    return {"name": name, "N": int(len(times)), "checksum": checksum(locs, times),
            "params": [float(v) for v in params], "ll": r["ll"], "grad": [float(v) for v in r["grad"]],
            "scale": [float(v) for v in r["scale"]], "precision": precision}


def main():
    path = os.path.join(HERE, "full_size_oracle.json")
    # entries already computed (a rerun after an interruption resumes; inputs are deterministic)
    done = {}
    for prev in (path, path + ".partial"):
        if os.path.exists(prev):
            with open(prev) as f:
                for group in json.load(f).values():
                    for c in group:
                        done[(c["name"], c["checksum"], tuple(c["params"]))] = c
    out = {"cluster_b70k": [], "cluster_b70k_shared_tau": [], "uneven_datasets": []}

    def add(group, name, locs, times, params, precision="long", **extra):
        key = (name, checksum(locs, times), tuple(float(v) for v in params))
        c = done.get(key) or dict(entry(name, locs, times, params, precision), **extra)
        c.setdefault("precision", "long")
        out[group].append(c)
        with open(path + ".partial", "w") as f:
            json.dump(out, f)

    locs, times, p0 = synth.workload("cluster_b70k")
    P = synth.jitter_params(64, seed=668)
    add("cluster_b70k", "p0", locs, times, p0, vector=-1)
    for k in (0, 31, 63):
        add("cluster_b70k", "k%d" % k, locs, times, P[k], vector=k)
    # vectors sharing tauX and tauT (sigma, omega, theta, mu0 jittered; the benchmark's shared-tau batch)
    S = synth.jitter_params(64, seed=671)
    S[:, 1:3] = synth.P0[1:3]
    for k in (0, 63):
        add("cluster_b70k_shared_tau", "shared_k%d" % k, locs, times, S[k], vector=k)
    for d, (l, t) in enumerate(synth.uneven_datasets(32)):
        add("uneven_datasets", "dataset%d" % d, l, t, synth.jitter_params(8, seed=669 + d)[0],
            precision="long" if len(t) <= 50_000 else "double", dataset=d, vector=0)
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    os.remove(path + ".partial")


if __name__ == "__main__":
    main()
