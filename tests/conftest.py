import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libhph.so")


def rel_err(a, b):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-300)


@pytest.fixture(scope="session")
def golden():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "golden_cases.json")) as f:
        return json.load(f)["cases"]


class OracleEngine:
    """Engine-shaped adapter over the CPU oracle (double precision): the reference chain that
    sampler tests drive, on the CPU and against the HIP engine."""

    def __init__(self, locs, times):
        self.locs, self.times = locs, times
        self.calls = 0

    def set_parameters(self, p):
        self.p = np.asarray(p, dtype=np.float64)

    def _total(self, p):
        import oracle
        self.calls += 1
        return oracle.total(self.locs, self.times, p, precision="double")

    def log_likelihood(self):
        return self._total(self.p)["ll"]

    def log_likelihood_and_gradient(self):
        r = self._total(self.p)
        return r["ll"], r["grad"]

    def set_parameters_batch(self, P):
        self.P = np.asarray(P, dtype=np.float64)

    def log_likelihood_batch(self, gradient=True):
        rs = [self._total(q) for q in self.P]
        ll = np.array([r["ll"] for r in rs])
        g = np.array([r["grad"] for r in rs])
        return (ll, g) if gradient else ll
