"""The base-2 pair form of the symmetric pass against the FP64 oracle at its edges.

The base-2 form stages coordinates as x' = sx (x - x0) (sx = tauX sqrt(log2(e)/2), x0 = the row
block's first event), so the rounding of x' grows with sx x diagonal (diagonal = the locations'
bounding box); and its background exponent is a plain sum of squares, non-negative, with the
clamp-free 2^-y. The cases here: a large coordinate scale (sx x diagonal 23.5, against 15 for the
20 km cluster workload) and a moderate one, and coincident events (groups of four sharing one
location, two of each group also sharing a time: r = 0 and dt = 0 pairs, the strict dt > 0 mask).
Same tolerances as the other FP64 parity tests (1e-12 relative; the gradient against
max(|g|, 1e-3 x the sum of |per-row terms|)), and 1e-6 in FP32-bulk mode.
"""
import numpy as np
import pytest

import oracle
from conftest import rel_err
from hawkes_amd import Engine, _lib, synth

pytestmark = pytest.mark.gpu

SX_PER_TAUX = 0.84932180028801904272  # sqrt(log2(e) / 2)


def events_at_scale(n, sx_diag, seed):
    """n events, uniform on a square whose diagonal times sx is sx_diag (params P0), times uniform
    on [0, 8760] h and sorted."""
    rng = np.random.Generator(np.random.PCG64(seed))
    side = sx_diag / (synth.P0[1] * SX_PER_TAUX) / np.sqrt(2.0)
    locs = rng.uniform(0.0, side, size=(n, 2))
    times = np.sort(rng.uniform(0.0, 8760.0, size=n))
    return locs, times


def coincident_events(n, seed):
    """n events in groups of four sharing one location; two of each group also share a time."""
    rng = np.random.Generator(np.random.PCG64(seed))
    m = n // 4
    base = rng.uniform(0.0, 20.0, size=(m, 2))
    locs = np.repeat(base, 4, axis=0)
    t = rng.uniform(0.0, 8760.0, size=(m, 3))
    times = np.concatenate([t[:, :1], t], axis=1).reshape(-1)  # the first two of a group tie
    order = np.argsort(times, kind="stable")
    return locs[order], times[order]


def check(locs, times, params, rows_per_lane, flags, tol_ll, tol_g, tol_p):
    ref = oracle.total(locs, times, params, precision="double", probs=True)
    e = Engine(2, len(times), flags=flags)
    e.set_pair_config(_lib.PAIR_KERNEL_SYMMETRIC, rows_per_lane, 4096)
    e.update_locations(locs)
    e.set_times(times)
    e.set_parameters(params)
    ll, g = e.log_likelihood_and_gradient()
    gerr = np.abs(g - ref["grad"]) / np.maximum(np.abs(ref["grad"]), 1e-3 * ref["scale"][1:])
    p = e.probs_self_excite()
    perr = np.max(np.abs(p - ref["probs"]) / np.maximum(np.abs(ref["probs"]), 1e-300))
    print("N=%d A=%d flags=%d: ll %.2e grad %.2e probs %.2e" % (len(times), rows_per_lane, flags,
                                                                 rel_err(ll, ref["ll"]), gerr.max(), perr))
    assert rel_err(ll, ref["ll"]) <= tol_ll
    assert gerr.max() <= tol_g, gerr
    assert perr <= tol_p
    assert np.all((p >= 0.0) & (p <= 1.0))


@pytest.mark.parametrize("rows_per_lane", [1, 2])
@pytest.mark.parametrize("sx_diag", [23.5, 12.0])
def test_base2_form_at_large_coordinate_scale(rows_per_lane, sx_diag):
    locs, times = events_at_scale(16384, sx_diag, seed=666 + int(sx_diag))
    check(locs, times, synth.P0, rows_per_lane, _lib.OPENCL, 1e-12, 1e-12, 1e-12)


@pytest.mark.parametrize("rows_per_lane", [1, 2])
def test_base2_form_with_coincident_events(rows_per_lane):
    locs, times = coincident_events(16384, seed=668)
    check(locs, times, synth.P0, rows_per_lane, _lib.OPENCL, 1e-12, 1e-12, 1e-12)


def test_base2_form_fp32_bulk_at_large_coordinate_scale():
    locs, times = events_at_scale(16384, 23.5, seed=669)
    check(locs, times, synth.P0, 2, _lib.OPENCL | _lib.FLOAT, 1e-6, 1e-6, 1e-6)
