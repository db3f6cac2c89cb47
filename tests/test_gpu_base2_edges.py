"""The base-2 pair form of the symmetric pass against the FP64 oracle at its edges.

The base-2 form stages coordinates as x' = sx (x - x0) (sx = tauX sqrt(log2(e)/2), x0 = the row
block's first event), so the rounding of x' grows with sx x diagonal (diagonal = the locations'
bounding box); and its background exponent is a plain sum of squares, non-negative, with the
clamp-free 2^-y. The cases here: a large coordinate scale (sx x diagonal 23.5, against 15 for the
20 km cluster workload) and a moderate one, and coincident events (groups of four sharing one
location, two of each group also sharing a time: r = 0 and dt = 0 pairs, the strict dt > 0 mask).
Every case asserts, on the host, the conditions under which the engine selects the base-2 form
(base2_gate, the same tests as the engine's) and that the engine reports having run it
(Engine.last_pass). Same tolerances as the other FP64 parity tests (1e-12 relative; the gradient
against max(|g|, 1e-3 x the sum of |per-row terms|)), and 1e-6 in FP32-bulk mode.

The split passes' background pass runs on the culled order, whose blocks span a whole window of
consecutive events in time; the base-2 form is then also gated on st x (window time span) <= 300.
test_split_pass_wide_window_takes_the_natural_form builds data past that gate.

In FP64 the time-sorted pass's table exponential folds the power of two into the table entry when
every background exponent of the call is at most 700 (ktab_gate; Engine.exponent_in_table): every
FP64 base-2 case here asserts which variant ran, and
test_exponent_in_table_variant_is_bitwise_the_ldexp_variant compares the two on one vector.
"""
import numpy as np
import pytest

import oracle
from conftest import rel_err
from hawkes_amd import Engine, _lib, synth

pytestmark = pytest.mark.gpu

SX_PER_TAUX = 0.84932180028801904272  # sqrt(log2(e) / 2)
HORIZON = 6000.0  # h: 16,384 events keep |omega| x (span of 512 sorted events) well below 300


def cull_window(n):
    """Events per window of the split passes' culled order (N / 8 as a power of two in [1024, 32768])."""
    w = 1024
    while w < 32768 and 2 * w <= n // 8:
        w *= 2
    return w


def base2_gate(locs, times, p, split=False):
    """The engine's conditions for the base-2 form (fast_ok in the engine), restated: omega >= 0,
    distinct times at least 8 ulp apart, |omega| x (largest span of 512 sorted events) <= 300, every
    background exponent below 5.5e6 and excitation exponent below 1e9, sx x diagonal <= 300,
    sigma sqrt(log2(e)/2) x diagonal <= 300, and for the split passes st x (largest window time
    span) <= 300."""
    t = np.sort(np.asarray(times, dtype=float))
    n = t.size
    tabs = max(abs(t[0]), abs(t[-1]))
    ulp = np.spacing(tabs)
    g = np.diff(t)
    if np.any((g > 0) & (g < 8 * ulp)):
        return False
    block_span = max(t[min(b + 511, n - 1)] - t[b] for b in range(0, n, 512))
    W = cull_window(n)
    window_span = max(t[min(b + W - 1, n - 1)] - t[b] for b in range(0, n, W))
    r2max = float(np.sum((locs.max(axis=0) - locs.min(axis=0)) ** 2))
    diag = np.sqrt(r2max)
    span = t[-1] - t[0]
    sigma, tauX, tauT, omega = p[0], p[1], p[2], p[3]
    st, sx = abs(tauT) * SX_PER_TAUX, abs(tauX) * SX_PER_TAUX
    ok = omega >= 0.0 and abs(omega) * block_span <= 300.0
    ok &= 0.5 * tauX ** 2 * r2max + 0.5 * tauT ** 2 * span ** 2 < 5.5e6
    ok &= abs(omega) * span + 0.5 * sigma ** 2 * r2max < 1e9
    ok &= sx * diag <= 300.0 and abs(sigma) * SX_PER_TAUX * diag <= 300.0
    if split:
        ok &= st * window_span <= 300.0
    return bool(ok)


def ktab_gate(locs, times, params):
    """The engine's bound for the exponent-in-table exponential (ktab_ok), restated: every vector's
    tauX^2 r2max / 2 + tauT^2 span^2 / 2 <= 700."""
    P = np.atleast_2d(params)
    r2max = float(np.sum((locs.max(axis=0) - locs.min(axis=0)) ** 2))
    span = float(np.max(times) - np.min(times))
    return bool(np.all(0.5 * P[:, 1] ** 2 * r2max + 0.5 * P[:, 2] ** 2 * span ** 2 <= 700.0))


def events_at_scale(n, sx_diag, seed):
    """n events, uniform on a square whose diagonal times sx is sx_diag (params P0), times uniform
    on [0, HORIZON) h and sorted."""
    rng = np.random.Generator(np.random.PCG64(seed))
    side = sx_diag / (synth.P0[1] * SX_PER_TAUX) / np.sqrt(2.0)
    locs = rng.uniform(0.0, side, size=(n, 2))
    times = np.sort(rng.uniform(0.0, HORIZON, size=n))
    return locs, times


def coincident_events(n, seed):
    """n events in groups of four sharing one location; two of each group also share a time."""
    rng = np.random.Generator(np.random.PCG64(seed))
    m = n // 4
    base = rng.uniform(0.0, 20.0, size=(m, 2))
    locs = np.repeat(base, 4, axis=0)
    t = rng.uniform(0.0, HORIZON, size=(m, 3))
    times = np.concatenate([t[:, :1], t], axis=1).reshape(-1)  # the first two of a group tie
    order = np.argsort(times, kind="stable")
    return locs[order], times[order]


def check(locs, times, params, rows_per_lane, flags, tol_ll, tol_g, tol_p, kind=_lib.PAIR_KERNEL_SYMMETRIC,
          base2=True):
    assert base2_gate(locs, times, params, split=kind == _lib.PAIR_KERNEL_SPLIT) == base2
    ref = oracle.total(locs, times, params, precision="double", probs=True)
    e = Engine(2, len(times), flags=flags)
    e.set_pair_config(kind, rows_per_lane, 4096)
    e.update_locations(locs)
    e.set_times(times)
    e.set_parameters(params)
    ll, g = e.log_likelihood_and_gradient()
    assert e.last_pass() == (kind, base2)
    ktab = kind == _lib.PAIR_KERNEL_SYMMETRIC and base2 and not (flags & _lib.FLOAT) and ktab_gate(locs, times, params)
    assert e.exponent_in_table() == ktab
    gerr = np.abs(g - ref["grad"]) / np.maximum(np.abs(ref["grad"]), 1e-3 * ref["scale"][1:])
    p = e.probs_self_excite()
    assert e.last_pass() == (kind, base2)
    perr = np.max(np.abs(p - ref["probs"]) / np.maximum(np.abs(ref["probs"]), 1e-300))
    print("N=%d A=%d flags=%d kind=%d base2=%d ktab=%d: ll %.2e grad %.2e probs %.2e" % (
        len(times), rows_per_lane, flags, kind, base2, ktab, rel_err(ll, ref["ll"]), gerr.max(), perr), flush=True)
    e.close()
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


def test_base2_form_in_the_split_passes():
    locs, times = events_at_scale(16384, 23.5, seed=670)
    check(locs, times, synth.P0, 2, _lib.OPENCL, 1e-12, 1e-12, 1e-12, kind=_lib.PAIR_KERNEL_SPLIT)


def test_split_pass_wide_window_takes_the_natural_form():
    """tauT = 5 / h on 8,000 events over 650 h: the time-sorted blocks (512 events, ~42 h) pass
    every gate of the base-2 form (largest background exponent 5.3e6), but a window of the culled
    order (1,024 events, ~83 h) has st x span ~ 350, so the split passes must take the natural
    form; the time-sorted pass keeps the base-2 form."""
    rng = np.random.Generator(np.random.PCG64(671))
    n = 8000
    locs = rng.uniform(0.0, 4.0, size=(n, 2))
    times = np.sort(rng.uniform(0.0, 650.0, size=n))
    p = np.array([1.0, 1 / 1.6, 5.0, 1.0, 1.0, 1.0])
    assert base2_gate(locs, times, p, split=False)
    check(locs, times, p, 2, _lib.OPENCL, 1e-12, 1e-12, 1e-12, kind=_lib.PAIR_KERNEL_SPLIT, base2=False)
    check(locs, times, p, 2, _lib.OPENCL, 1e-12, 1e-12, 1e-12, kind=_lib.PAIR_KERNEL_SYMMETRIC, base2=True)


def test_exponent_in_table_variant_is_bitwise_the_ldexp_variant():
    """One vector (P0 on 20 km x 6,000 h: largest background exponent ~315) in the two kernel
    variants: alone, its call runs the exponent-in-table exponential; in a batch with a vector
    whose exponents reach ~1,300 (tauT = 0.008 / h) the launch runs the table exponential with its
    ldexp. Same chunking (one plan at both K), so the log-likelihood and gradient must agree bit
    for bit; the single call is also checked against the FP64 oracle."""
    rng = np.random.Generator(np.random.PCG64(672))
    n = 16384
    locs = rng.uniform(0.0, 20.0, size=(n, 2))
    times = np.sort(rng.uniform(0.0, HORIZON, size=n))
    pa = np.array(synth.P0, dtype=float)
    pb = pa.copy()
    pb[2] = 0.008
    assert base2_gate(locs, times, pa) and base2_gate(locs, times, pb)
    assert ktab_gate(locs, times, pa) and not ktab_gate(locs, times, pb)
    e = Engine(2, n, flags=_lib.OPENCL)
    e.set_pair_config(_lib.PAIR_KERNEL_SYMMETRIC, 2, 4096)
    e.update_locations(locs)
    e.set_times(times)
    e.set_parameters(pa)
    ll_a, g_a = e.log_likelihood_and_gradient()
    assert e.last_pass() == (_lib.PAIR_KERNEL_SYMMETRIC, True) and e.exponent_in_table()
    e.set_parameters_batch(np.stack([pa, pb]))
    ll_b, g_b = e.log_likelihood_batch()
    assert e.last_pass() == (_lib.PAIR_KERNEL_SYMMETRIC, True) and not e.exponent_in_table()
    print("ll single %r batch %r" % (ll_a, ll_b[0]), flush=True)
    assert ll_b[0] == ll_a
    assert np.array_equal(g_b[0], g_a)
    ref = oracle.total(locs, times, pa, precision="double")
    assert rel_err(ll_a, ref["ll"]) <= 1e-12
    assert np.all(np.abs(g_a - ref["grad"]) <= 1e-12 * np.maximum(np.abs(ref["grad"]), 1e-3 * ref["scale"][1:]))
    e.close()
