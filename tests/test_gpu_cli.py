"""microlc end to end on synthetic GERLUMPH maps: every output file and the
last stdout line against the FP64 oracle run through the same steps as
driver.cpp (locations, expanding_source, MPDs and their ratio, pair chi2,
bins, trapezium)."""
import json
import os
import subprocess

import numpy as np
import pytest

from microsn_amd import mapio

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "microsn_amd", "microlc")


def _maps(tmp_path, N):
    rng = np.random.default_rng(8)
    dirs = []
    for name, avgN in (("A", 120.0), ("B", 95.0)):
        g = rng.standard_normal((N, N))
        counts = np.maximum(0, np.round(avgN * np.exp(0.6 * g))).astype(np.int32)
        d = str(tmp_path / name)
        mapio.write_gerlumph(d, counts, avgmu=1.0 / ((1 - 0.4) ** 2 - 0.35 ** 2), avgN=avgN, width=5.0,
                             k=0.4, g=0.35, s=0.0)
        dirs.append(d)
    return dirs


def _pairs(path):
    return np.loadtxt(path)


@pytest.mark.parametrize("shape,loc", [("gaussian", "grid"), ("uniform", "grid"), ("gaussian", "random")])
def test_microlc_matches_oracle(tmp_path, orc, shape, loc):
    assert os.path.exists(EXE), "microlc not built"
    N = 200
    dA, dB = _maps(tmp_path, N)
    locations = ({"type": "grid", "Nmap": N, "ss": 5.0, "Nside": 12} if loc == "grid" else
                 {"type": "random", "Nmap": N, "ss": 5.0, "seed": 321, "Nloc": 150})
    inp = {
        "sizes": [0, 0.05, 0.1, 0.2], "shape": shape,
        "locations": locations,
        "mpds": {"mpd": {"Nbins": 60, "min": 0.01, "max": 50},
                 "ratio": {"Nbins": 40, "min": 0.01, "max": 30}},
        "DL": 1570, "DS": 1750, "DLS": 850,
        "data": [{"d": [0.7, 1.2, 1.4, 1.1, 0.9], "s": [0.1, 0.15, 0.1, 0.1, 0.15], "t": [5, 7, 12, 14, 17]}],
        "maps": {"A": dA, "B": dB},
    }
    p = tmp_path / "in.json"
    p.write_text(json.dumps(inp))
    out = tmp_path / "out"
    out.mkdir()
    print("microlc", shape, loc, flush=True)
    r = subprocess.run([EXE, str(p)], cwd=out, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.strip().splitlines()
    assert lines[0] == f"Number of profiles set to: 4 ({shape})"
    assert "CONVOLUTIONS" in lines and "MPDS" in lines and "CHI-SQUARED" in lines
    assert lines[-1].startswith("GPU: ")

    # ---- the same steps on the oracle
    metaA, cA = mapio.read_gerlumph(dA)
    metaB, cB = mapio.read_gerlumph(dB)
    mA = orc.gerlumph_normalise(cA, metaA["avgmu"], metaA["avgN"], metaA["k"], metaA["g"])
    mB = orc.gerlumph_normalise(cB, metaB["avgmu"], metaB["avgN"], metaB["k"], metaB["g"])
    Noff = int(np.ceil(N / 5.0 * 0.2))
    if loc == "grid":
        lx, ly = orc.locations_on_grid(12, N, Noff)
    else:  # rand_input.json's form: std::mt19937(seed), inclusive uniform ints
        lx, ly = orc.locations_random(150, N, Noff, 321)
    LCA, DLCA = orc.expanding_source(mA, 5.0, inp["sizes"], shape, lx, ly)
    LCB, DLCB = orc.expanding_source(mB, 5.0, inp["sizes"], shape, lx, ly)
    assert np.allclose(_pairs(out / "lcurvesA.dat").reshape(LCA.shape), LCA, rtol=0, atol=6e-7)
    assert np.allclose(_pairs(out / "lcurvesB.dat").reshape(LCB.shape), LCB, rtol=0, atol=6e-7)

    bA, hA = orc.binned_mpd(mA, 60, 0.01, 50.0)
    bB, hB = orc.binned_mpd(mB, 60, 0.01, 50.0)
    hA, hB = orc.mpd_normalize(bA, hA), orc.mpd_normalize(bB, hB)
    bR = orc.mpd_bins(40, 0.01, 30.0)
    cR = orc.mpd_divide_by(bA, hA, bB, hB, bR)
    for name, b, c in (("mpd_A.dat", bA, hA), ("mpd_B.dat", bB, hB), ("mpd_R.dat", bR, cR)):
        got = _pairs(out / name)
        assert np.allclose(got[:, 0], b, rtol=1e-6, atol=0) and np.allclose(got[:, 1], c, rtol=1e-6, atol=1e-300)

    si, lo, npb = orc.setup_integral(LCA, LCB, bR)
    Dfac = np.sqrt(1750 * 850 / 1570)
    v = orc.setup_chi2((0.4, 0.5), 1.0, Dfac, inp["data"][0]["d"], inp["data"][0]["t"], inp["data"][0]["s"],
                       inp["sizes"])
    chi2 = orc.chi2_pairs(v, LCA, DLCA, LCB, DLCB, si)
    bc, be = orc.bin_chi2(chi2, lo, npb, gpu_convention=True)
    n_got = _pairs(out / "binned_N.dat")
    assert np.array_equal(n_got[:, 1].astype(np.int64), npb.astype(np.int64))
    e_got = _pairs(out / "binned_exp.dat")
    assert np.allclose(e_got[:, 1], be, rtol=1e-6, atol=1e-300)
    integral = orc.trapezium(be, cR, bR)
    got = float(lines[-1].split()[1])
    assert abs(got - integral) <= 5e-6 * abs(integral) + 1e-300


def test_load_map_on_device_is_bitexact(tmp_path, ctx):
    rng = np.random.default_rng(3)
    counts = rng.integers(0, 5000, (64, 64)).astype(np.int32)
    d = str(tmp_path / "m")
    mapio.write_gerlumph(d, counts, avgmu=2.1, avgN=137.0, width=10.0, k=0.45, g=0.3)
    meta, dmap = mapio.load_map(ctx, d)
    assert np.array_equal(dmap.cpu().numpy(), mapio.magnification(meta, counts))


def test_chi2_driver_output():
    """chi2_driver.cu:8-79 with N = 1000: stdout in the reference's order, the
    likelihood N^2 exp(-10.0070457) of the constant fake data from both the
    device run (the kernel's indexing) and the host run (the CPU twin's)."""
    exe = os.path.join(ROOT, "microsn_amd", "chi2_driver")
    assert os.path.exists(exe), "chi2_driver not built"
    r = subprocess.run([exe, "1000", "7", "10"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.strip().splitlines()
    assert lines[0] == "GPU run:"
    assert lines[1].startswith("Elapsed time: ") and lines[2].startswith("Elapsed time: ")
    assert lines[3].startswith("Likelihood: ")
    assert lines[4].startswith("number of free bytes in gpu memory: ")
    assert lines[5] == "CPU run:" and lines[6].startswith("Elapsed time: ")
    assert lines[7].startswith("Likelihood: ")
    chi2 = float(np.float32(35 * ((0.97 - 0.42 / 1.1) / 1.1) ** 2))
    want = 1000 * 1000 * np.exp(-chi2)
    for line in (lines[3], lines[7]):
        assert abs(float(line.split()[1]) - want) < 1e-5 * want
