"""Pair mode at the reference's own scale (simulating_lcurves/grid_input.json:
two 10000^2 GERLUMPH maps over 25 Einstein radii, 5 Gaussian source sizes,
99 x 99 = 9801 grid locations, so 9801^2 = 9.6e7 image pairs, 300 ratio
bins), checked through size-independent properties:
  * microlc end to end on synthetic 10000^2 GERLUMPH maps: the per-bin pair
    counts of binned_N.dat equal an exact host count of the ratios
    z = LCA[a,0] / LCB[b,0] (the point samples at the grid locations) under
    the bin edges, over all 9.6e7 pairs;
  * setup_integral_GPU on the same point samples: sorted_ind is a
    permutation of the 9.6e7 pairs, the keys it orders are non-decreasing
    with ties in pair-index order (std::stable_sort semantics), checked on
    the device over all pairs and on the host on a random sample, and the
    bin table equals the host count."""
import json
import os
import subprocess

import numpy as np
import pytest
import torch

from microsn_amd import api, mapio

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "microsn_amd", "microlc")
N, SS, NSIDE = 10000, 25.0, 99
SIZES = [0, 0.05, 0.1, 0.2, 0.3]


def _grid_input(dA, dB):
    return {"sizes": SIZES, "shape": "gaussian",
            "locations": {"type": "grid", "Nmap": N, "ss": SS, "Nside": NSIDE},
            "mpds": {"mpd": {"Nbins": 500, "min": 0.0001, "max": 400},
                     "ratio": {"Nbins": 300, "min": 0.0001, "max": 300}},
            "DL": 1570, "DS": 1750, "DLS": 850,
            "data": [{"d": [0.7, 1.2, 1.4, 1.1, 0.9], "s": [0.1, 0.15, 0.1, 0.1, 0.15], "t": [5, 7, 12, 14, 17]}],
            "maps": {"A": dA, "B": dB}}


def _counts(seed, avgN):
    rng = np.random.default_rng(seed)
    out = np.empty((N, N), np.int32)
    for r0 in range(0, N, 1000):  # in slabs, to bound host memory
        g = rng.standard_normal((1000, N))
        out[r0:r0 + 1000] = np.maximum(0, np.round(avgN * np.exp(0.6 * g))).astype(np.int32)
    return out


def _host_bin_counts(a, b, edges):
    """#{pairs with edge[i-1] < a_x / b_y <= edge[i]} exactly (IEEE FP64
    division as on the device), over all pairs, in chunks of rows."""
    n = np.zeros(edges.size + 1, np.int64)
    for i0 in range(0, a.size, 512):
        z = a[i0:i0 + 512, None] / b[None, :]
        z = z[~np.isnan(z)]
        n += np.bincount(np.searchsorted(edges, z.ravel(), side="left"), minlength=edges.size + 1)
    return n[:edges.size]


@pytest.fixture(scope="module")
def grid_maps(tmp_path_factory):
    d = tmp_path_factory.mktemp("gerlumph10000")
    dirs = []
    for name, seed, avgN in (("A", 81, 120.0), ("B", 82, 95.0)):
        counts = _counts(seed, avgN)
        p = str(d / name)
        mapio.write_gerlumph(p, counts, avgmu=1.0 / ((1 - 0.4) ** 2 - 0.35 ** 2), avgN=avgN, width=SS,
                             k=0.4, g=0.35, s=0.0)
        dirs.append(p)
        del counts
    return d, dirs


def _point_samples(orc, path):
    """LC[:, 0] of expanding_source: the normalised map at the grid locations."""
    meta, counts = mapio.read_gerlumph(path)
    Noff = int(np.ceil(N / SS * SIZES[-1]))
    lx, ly = orc.locations_on_grid(NSIDE, N, Noff)
    c = np.ascontiguousarray(counts[ly, lx])
    return orc.gerlumph_normalise(c, meta["avgmu"], meta["avgN"], meta["k"], meta["g"])


def test_microlc_grid_input_scale(grid_maps, orc):
    assert os.path.exists(EXE), "microlc not built"
    d, (dA, dB) = grid_maps
    p = d / "grid_input.json"
    p.write_text(json.dumps(_grid_input(dA, dB)))
    out = d / "out"
    out.mkdir(exist_ok=True)
    print("microlc at grid_input scale", flush=True)
    r = subprocess.run([EXE, str(p)], cwd=out, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    lines = r.stdout.strip().splitlines()
    assert lines[0] == "Number of profiles set to: 5 (gaussian)"
    assert any(line.startswith("Sampling from 9801 locations") for line in lines)
    assert lines[-1].startswith("GPU: ") and np.isfinite(float(lines[-1].split()[1]))
    a, b = _point_samples(orc, dA), _point_samples(orc, dB)
    edges = orc.mpd_bins(300, 0.0001, 300.0)
    want = _host_bin_counts(a, b, edges)
    got = np.loadtxt(out / "binned_N.dat")
    assert got.shape == (300, 2)
    assert np.array_equal(got[:, 1].astype(np.int64), want)
    assert np.allclose(got[:, 0], edges, rtol=1e-6, atol=0)


def test_setup_integral_grid_input_scale(grid_maps, orc, ctx):
    _, (dA, dB) = grid_maps
    a, b = _point_samples(orc, dA), _point_samples(orc, dB)
    nloc, nprof = a.size, len(SIZES)
    assert nloc == 9801
    LCA, LCB = api.SimLC(ctx, nloc, nprof), api.SimLC(ctx, nloc, nprof)
    LCA.LC[:, 0] = torch.tensor(a, device="cuda")
    LCB.LC[:, 0] = torch.tensor(b, device="cuda")
    edges = orc.mpd_bins(300, 0.0001, 300.0)
    sb = api.Chi2SortBins(ctx, nloc, edges.size)
    print("setup_integral 9.6e7 pairs", flush=True)
    api.setup_integral_GPU(ctx, LCA, LCB, edges, sb)
    torch.cuda.synchronize()
    npairs = nloc * nloc
    rank = sb.d_sorted_ind.long()
    # a permutation: scattering the pair indices by rank fills every slot once
    perm = torch.full((npairs,), -1, dtype=torch.long, device="cuda")
    perm[rank] = torch.arange(npairs, device="cuda")
    assert bool((perm >= 0).all())
    # keys in that order are non-decreasing, ties in pair-index order
    z = (torch.tensor(a, device="cuda")[:, None] / torch.tensor(b, device="cuda")[None, :]).reshape(-1)
    zs = z[perm]
    assert not bool(torch.isnan(zs).any())  # no 0/0 among these samples
    dz = zs[1:] - zs[:-1]
    assert bool((dz >= 0).all())
    tie = dz == 0
    assert bool((perm[1:][tie] > perm[:-1][tie]).all())
    # the same on the host for a random sample of pairs
    rng = np.random.default_rng(9)
    smp = np.sort(rng.choice(npairs, 200_000, replace=False))
    r_s = rank[torch.tensor(smp, device="cuda")].cpu().numpy()
    z_s = (a[smp // nloc] / b[smp % nloc])
    order = np.argsort(r_s)
    assert np.array_equal(order, np.argsort(z_s, kind="stable"))
    # the bin table against the exact host count
    want = _host_bin_counts(a, b, edges)
    assert np.array_equal(sb.n_per_bin.astype(np.int64), want)
    assert np.array_equal(sb.lower_ind.astype(np.int64), np.concatenate([[0], np.cumsum(want)[:-1]]))
