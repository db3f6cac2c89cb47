"""The dcData fixture (tests/golden/dcdata.csv, from the reference's data/dcData.RData)."""
import os

import numpy as np

import oracle
from conftest import ROOT


def load_dcdata():
    d = np.loadtxt(os.path.join(ROOT, "tests", "golden", "dcdata.csv"), delimiter=",", skiprows=1)
    return d[:, :2].copy(), d[:, 2].copy()


def test_fixture_shape_and_order():
    locs, times = load_dcdata()
    assert locs.shape == (5022, 2)
    assert np.all(np.diff(times) >= 0)          # README.md: is.unsorted(dcData$Time) is FALSE
    assert (np.diff(times) == 0).sum() == 53    # tied times: the strict dt > 0 mask matters
    assert times.max() < 8760                   # hours of 2010 (R/data.R says minutes)


def test_oracle_on_prefix_is_consistent():
    # the first 400 events: long double and compensated double oracles agree
    locs, times = load_dcdata()
    p = [1.0, 1 / 1.6, 1 / (14 * 24), 1.0, 1.0, 1.0]
    a = oracle.total(locs[:400], times[:400], p)
    b = oracle.total(locs[:400], times[:400], p, precision="double")
    assert abs(a["ll"] - b["ll"]) <= 1e-14 * abs(a["ll"])
