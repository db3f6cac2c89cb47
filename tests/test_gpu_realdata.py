"""Real data: the 2010 Washington D.C. gunfire events (5022, with ties) against the oracle, with
the README's starting values (1, 1/1.6, 1/(14*24), 1, 1, 1), and the README's sampler workflow."""
import numpy as np
import pytest

import oracle
from conftest import rel_err
from hawkes_amd import Engine, _lib, sampler
from test_realdata import load_dcdata

pytestmark = pytest.mark.gpu

README_PARAMS = np.array([1.0, 1 / 1.6, 1 / (14 * 24), 1.0, 1.0, 1.0])


@pytest.fixture(scope="module")
def dc():
    locs, times = load_dcdata()
    return locs, times, oracle.total(locs, times, README_PARAMS, probs=True)


@pytest.mark.parametrize("kind", [_lib.PAIR_KERNEL_ROWWISE, _lib.PAIR_KERNEL_SYMMETRIC])
def test_dcdata_fp64(dc, kind):
    locs, times, ref = dc
    e = Engine(2, len(times))
    e.set_pair_kernel(kind)
    e.update_locations(locs)
    e.set_times(times)
    e.set_parameters(README_PARAMS)
    ll, g = e.log_likelihood_and_gradient()
    assert rel_err(ll, ref["ll"]) <= 1e-12
    err = np.abs(g - ref["grad"]) / np.maximum(np.abs(ref["grad"]), 1e-3 * ref["scale"][1:])
    assert err.max() <= 1e-12
    np.testing.assert_allclose(e.probs_self_excite(), ref["probs"], rtol=1e-12, atol=0)


def test_dcdata_fp32_bulk(dc):
    locs, times, ref = dc
    e = Engine(2, len(times), flags=_lib.OPENCL | _lib.FLOAT)
    e.update_locations(locs)
    e.set_times(times)
    e.set_parameters(README_PARAMS)
    ll, g = e.log_likelihood_and_gradient()
    assert rel_err(ll, ref["ll"]) <= 1e-6
    err = np.abs(g - ref["grad"]) / np.maximum(np.abs(ref["grad"]), 1e-3 * ref["scale"][1:])
    assert err.max() <= 1e-6
    p = e.probs_self_excite()
    assert np.max(np.abs(p - ref["probs"]) / np.maximum(ref["probs"], 1e-3)) <= 1e-6


def test_readme_sampler_workflow(dc):
    # README.md: sampler(n_iter=10, locations=cbind(dcData$X, dcData$Y), times=dcData$Time)
    locs, times, _ = dc
    e = Engine(2, len(times))
    e.update_locations(locs)
    e.set_times(times)
    out = sampler.adaptive_mh(e, 10, README_PARAMS, rng=np.random.default_rng(666))
    assert out["samples"].shape == (10, 6)
    assert np.all(np.isfinite(out["target"]))
