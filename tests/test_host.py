"""Host-side functions exported by libmlc.so (mlc_host.h: locations, chi2
set-up tables, MPD bins, trapezium) against the oracle, the reference's own
compiled sources (oracle/_ref) and the committed golden vectors.  CPU only;
integer and FP64 results must be identical."""
import json
import os

import numpy as np
import pytest

from microsn_amd import api
from oracle import oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))


def test_locations_random_golden_vectors():
    with open(os.path.join(HERE, "golden", "locations_random.json")) as f:
        cases = json.load(f)["cases"]
    for c in cases:
        x, y = api.locations_random(c["Nloc"], c["Nmap"], c["Noff"], c["seed"])
        assert x.tolist() == c["x"] and y.tolist() == c["y"], c["seed"]
        ox, oy = orc.locations_random(c["Nloc"], c["Nmap"], c["Noff"], c["seed"])
        assert ox.tolist() == c["x"] and oy.tolist() == c["y"], c["seed"]


@pytest.mark.parametrize("args", [(99, 10000, 120), (7, 100, 3), (1, 10, 0), (13, 500, 249)])
def test_locations_on_grid_matches_oracle(args):
    x, y = api.locations_on_grid(*args)
    ox, oy = orc.locations_on_grid(*args)
    assert np.array_equal(x, ox) and np.array_equal(y, oy)


def test_find_indices_and_setup_chi2_match_oracle():
    rng = np.random.default_rng(0)
    x = np.concatenate([[0.0], np.sort(rng.random(6))])
    t = np.sort(rng.uniform(-0.1, 1.2, 40))
    assert np.array_equal(api.find_indices(t, x), orc.find_indices(t, x))
    d, s = rng.uniform(0.5, 1.5, 12), rng.uniform(0.05, 0.2, 12)
    tt = np.sort(rng.uniform(0, 30, 12))
    for quirk in (True, False):
        a = api.setup_chi2_calculation((0.4, 0.5), (1.0,), 1.7, d, tt, s, x * 0.4, quirk)
        b = orc.setup_chi2((0.4, 0.5), 1.0, 1.7, d, tt, s, x * 0.4, quirk)
        assert a["Njp"] == b["Njp"] and a["const_term"] == b["const_term"]
        for k in ("indA", "indB", "facA", "facB", "d", "s"):
            assert np.array_equal(a[k], b[k]), k


def test_mpd_functions_match_reference_and_oracle():
    for log_scale in (True, False):
        assert np.array_equal(api.mpd_bins(500, 1e-4, 400.0, log_scale), orc.mpd_bins(500, 1e-4, 400.0, log_scale))
    rng = np.random.default_rng(1)
    bA, bB, bR = api.mpd_bins(50, 0.01, 50.0), api.mpd_bins(60, 0.01, 60.0), api.mpd_bins(40, 0.01, 10.0)
    cA, cB = rng.random(50), rng.random(60)
    nA, nB = api.mpd_normalize(bA, cA), api.mpd_normalize(bB, cB)
    assert np.array_equal(nA, orc.mpd_normalize(bA, cA))
    assert np.array_equal(api.mpd_divide_by(bA, nA, bB, nB, bR), orc.mpd_divide_by(bA, nA, bB, nB, bR))
    x = np.sort(rng.random(21))
    f1, f2 = rng.random(21), rng.random(21)
    assert api.trapezium(f1, f2, x) == orc.trapezium(f1, f2, x)
    r = orc.ref()
    if r is not None:
        assert api.trapezium(f1, f2, x) == r.ref_trapezium(21, orc._p(f1), orc._p(f2), orc._p(x))
