"""Empty inputs through the C ABI: a sort of no keys, bin tables over no
keys (host and device paths: every bin empty, the reference kernel's
sentinel for its means), and light curves whose observation sets hold no
epochs at all (chi2 and gradient against the FP64 oracle: zero for curves
on the map, +inf for the curves the band kernels flag without reading an
epoch)."""
import numpy as np
import pytest
import torch

from microsn_amd import api
from microsn_amd import lightcurves as lcm

pytestmark = pytest.mark.gpu


def test_sort_of_no_keys(ctx):
    keys = torch.empty(0, dtype=torch.float64, device="cuda")
    perm, rank, sk = api.sort_f64(ctx, keys, rank=True, sorted_keys=True)
    torch.cuda.synchronize()
    assert perm.numel() == 0 and rank.numel() == 0 and sk.numel() == 0


def test_bin_tables_of_no_keys(ctx, orc):
    keys = torch.empty(0, dtype=torch.float64, device="cuda")
    bins = orc.mpd_bins(100, 1.0, 1.0e7, True)
    lo, npb = api.bin_edges(ctx, keys, bins)
    assert not lo.any() and not npb.any()
    dlo, dnpb, bc, be = api.bin_table_device(ctx, keys, bins)
    torch.cuda.synchronize()
    assert not dlo.any() and not dnpb.any()
    # empty bins: kernel_bin_chi2's sentinel (gpu_functions.cu:789-792)
    assert torch.all(bc == 2000.0) and torch.all(be == float(np.exp(-1000.0)))


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_light_curves_without_epochs(ctx, orc, prec):
    rng = np.random.default_rng(70)
    N = 128
    m = 5.0 * np.exp(0.5 * rng.standard_normal((N, N)))
    spec = api.MapSpectrum(ctx, torch.tensor(m, device="cuda"), prec)
    nodes = lcm.NodeSet(ctx, spec, ["gaussian"], [[[2.0], [3.0]]], [1.0, 1.5])
    empty = (np.zeros(0), np.zeros(0), np.zeros(0))
    sets = [[empty, empty], [empty, empty]]
    obs = lcm.Observations(ctx, 2, sets)
    assert obs.nepoch == 0
    n = 40
    prm = np.stack([rng.uniform(20, N - 21, n), rng.uniform(20, N - 21, n), rng.uniform(0, 2 * np.pi, n),
                    np.full(n, 0.5), np.full(n, 2.0), np.zeros(n), np.zeros(n)], 1)
    prm[3, 0] = N + 10.0  # off the map, but no epoch ever reads it
    prof = np.zeros(n, np.int32)
    prof[5] = 1  # no such profile
    oi = (np.arange(n) % 2).astype(np.int32)
    oi[7] = 2  # no such observation set
    c, g = lcm.chi2_grad(ctx, nodes, obs, torch.tensor(prm, device="cuda"), torch.tensor(prof, device="cuda"),
                         torch.tensor(oi, device="cuda"))
    torch.cuda.synchronize()
    c, g = c.cpu().numpy(), g.cpu().numpy()
    # invalid profile or observation set: flagged (the oracle takes valid
    # indices only)
    assert np.isinf(c[5]) and np.isinf(c[7]) and not g[[5, 7]].any()
    ok = np.ones(n, bool)
    ok[[5, 7]] = False
    maps = nodes.rowmajor().double().cpu().numpy()
    rc, rg = orc.lc_chi2_grad(maps, nodes.B, nodes.node_off, nodes.node_cnt, nodes.node_rad, None,
                              nodes.band_scale, prm[ok], prof[ok], oi[ok], np.zeros(5, np.int32), np.zeros(0),
                              np.zeros(0), np.zeros(0))
    assert np.array_equal(c[ok], rc) and not rc.any()  # the off-map curve reads no epoch either
    assert np.array_equal(g[ok], rg) and not rg.any()
    spec.close()
