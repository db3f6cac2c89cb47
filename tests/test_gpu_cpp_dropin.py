"""The C++ drop-in caller (tests/cpp/dropin_main.cpp, built by __graft_entry__.build()) runs
the reference's call sequence through hph::factory and prints the log-likelihood, gradient and
probabilities: they equal the Python binding's results through the C ABI bit for bit (one
engine, one code path) and the oracle's to 1e-12."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import ROOT, rel_err
from hawkes_amd import Engine, synth

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "tests", "cpp", "dropin")


def test_cpp_caller_matches_python_binding_and_oracle():
    assert os.path.exists(BIN), "tests/cpp/dropin not built (run __graft_entry__.build())"
    locs, times = synth.cluster_process(1400, side=6.0, seed=44)
    p = synth.jitter_params(1, seed=45)[0]
    n = len(times)
    lines = ["2 %d" % n] + ["%.17g %.17g %.17g" % (locs[i, 0], locs[i, 1], times[i]) for i in range(n)]
    lines.append(" ".join("%.17g" % v for v in p))
    run = subprocess.run([BIN], input="\n".join(lines) + "\n", capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stderr
    out = run.stdout.split("\n")
    ll_cpp = float(out[0])
    g_cpp = np.array([float(v) for v in out[1].split()])
    probs_cpp = np.array([float(v) for v in out[2:2 + n]])
    e = Engine(2, n)
    e.update_locations(locs)
    e.set_times(times)
    e.set_parameters(p)
    ll, g = e.log_likelihood_and_gradient()
    probs = e.probs_self_excite()
    e.close()
    assert ll_cpp == ll and np.array_equal(g_cpp, g) and np.array_equal(probs_cpp, probs)
    ref = oracle.total(locs, times, p, probs=True)
    assert rel_err(ll_cpp, ref["ll"]) <= 1e-12
    np.testing.assert_allclose(probs_cpp, ref["probs"], rtol=1e-12, atol=1e-300)
