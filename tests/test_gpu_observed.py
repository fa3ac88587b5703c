"""The reference's own observed data (data.csv: 4 lensed images, 6 filters,
1 to 10 epochs per filter, kept as tests/golden/data.csv) through the device
light-curve chi2: one observation set per image, one band per filter, a
per-image magnitude offset on every curve, an expanding thin-disc source
whose size grows with wavelength.  The FP64 device path is checked against
the FP64 host oracle on every curve, and the FP32 path against the FP64
path."""
import os

import numpy as np
import pytest
import torch

from microsn_amd import api, observed
from microsn_amd import lightcurves as lcm

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_data_csv_through_device_chi2(ctx, orc):
    data = observed.load_csv(os.path.join(HERE, "golden", "data.csv"))
    assert data.nimages == 4 and len(data.filters) == 6
    # band scale: thin-disc size ~ wavelength^(4/3), the filters' pivot
    # wavelengths in nm relative to F390W
    lam = np.array([390.0, 475.0, 625.0, 814.0, 1100.0, 1600.0])
    bs = (lam / lam[0]) ** (4.0 / 3.0)
    rng = np.random.default_rng(31)
    N = 1024
    gen = torch.Generator(device="cuda").manual_seed(31)
    m = torch.exp(0.5 * torch.randn((N, N), dtype=torch.float64, device="cuda", generator=gen))
    s0 = api.MapSpectrum(ctx, m, "f64")
    m = s0.convolve("gaussian", 3.0)
    s0.close()
    m = (m * (5.0 / m.mean())).contiguous()
    # r0 in [1, 2] px, w in [0, 0.1] px/day over the 36.6-day span: the
    # band-b radius stays inside bs[b] * [0.8, 6.5]
    node_r = [0.8, 2.0, 4.0, 6.5]
    radii = [[[bs[b] * r for r in node_r] for b in range(6)]]
    ntraj = 5000
    x0 = rng.uniform(150, N - 151, ntraj)
    y0 = rng.uniform(150, N - 151, ntraj)
    th = rng.uniform(0, 2 * np.pi, ntraj)
    v = rng.uniform(0.5, 3.0, ntraj)
    r0 = rng.uniform(1.0, 2.0, ntraj)
    w = rng.uniform(0.0, 0.1, ntraj)
    # curve = (trajectory, image): image i's observation set and its offset
    base = np.array([-4.0, -2.9, -1.8, -1.5])
    params = np.concatenate([np.stack([x0, y0, th, v, r0, w, base[i] + rng.normal(0, 0.1, ntraj)], 1)
                             for i in range(4)])
    obs_idx = np.repeat(np.arange(4, dtype=np.int32), ntraj)
    prof = np.zeros(params.shape[0], np.int32)
    res = {}
    for prec in ("f64", "f32"):
        spec = api.MapSpectrum(ctx, m, prec)
        nodes = lcm.NodeSet(ctx, spec, ["exponential"], radii, bs)
        obs = observed.observations(ctx, data)
        assert obs.nobs == 4 and obs.B == 6 and obs.nepoch == 4 * 41
        c, g = lcm.chi2_grad(ctx, nodes, obs, torch.tensor(params, device="cuda"),
                             torch.tensor(prof, device="cuda"), torch.tensor(obs_idx, device="cuda"))
        torch.cuda.synchronize()
        res[prec] = (c.cpu().numpy(), g.cpu().numpy(), nodes, obs, spec)
    c64, g64, n64, obs, _ = res["f64"]
    assert np.all(np.isfinite(c64))
    maps = n64.rowmajor().cpu().numpy()
    rc, rg = orc.lc_chi2_grad(maps, 6, n64.node_off, n64.node_cnt, n64.node_rad, None, n64.band_scale, params,
                              prof, obs_idx, obs.off.cpu().numpy(), obs.t.cpu().numpy(),
                              obs.mobs.cpu().numpy(), obs.sig.cpu().numpy())
    assert np.max(np.abs(c64 - rc) / np.abs(rc)) < 1e-11
    assert np.max(np.abs(g64 - rg) / np.abs(rg).max(0)) < 1e-10
    c32, g32 = res["f32"][0], res["f32"][1]
    assert np.max(np.abs(c32 - c64) / np.abs(c64)) < 1e-5
    assert np.max(np.abs(g32 - g64) / np.abs(g64).max(0)) < 1e-4
    # the offset gradient of image i is -2 sum (mobs - m)/sig^2 over its 41
    # epochs: nonzero and different per image
    assert np.all(g64[:, 6] != 0.0)
    for prec in res:
        res[prec][4].close()
