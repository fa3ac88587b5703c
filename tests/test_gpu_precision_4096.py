"""The precision workload at its full shape: a 4096^2 map, Gaussian, uniform
disc and thin disc profiles at three bands with half-light radii 2 / 4 / 8 px
(nine node maps), 10^5 trajectories x 3 profiles, 100 irregular epochs per
band.  Checks, on the same inputs:
  * FP32 fused path against the FP64 fused path: chi2 relative error at most
    1e-5 on every curve, and the sorted 100-bin tables identical;
  * a sample of 2 400 curves of the FP64 path against the FP64 host oracle on
    the same FP64 node maps (the oracle's own bilinear + Hermite + chi2 +
    adjoint), to 1e-11;
  * the FP32 gradient against the FP64 gradient, relative to each
    component's scale."""
import numpy as np
import pytest
import torch

from microsn_amd import api
from microsn_amd import lightcurves as lcm

pytestmark = pytest.mark.gpu

N, NTRAJ, EPOCHS, TSPAN = 4096, 100_000, 100, 100.0
SHAPES = ("gaussian", "uniform", "exponential")
BAND_SCALE = (1.0, 2.0, 4.0)
R0 = 2.0


def _map(ctx, seed=1001):
    """Log-normal exp(0.5 g), smoothed with a 3 px Gaussian, mean 5."""
    gen = torch.Generator(device="cuda").manual_seed(seed)
    g = torch.randn((N, N), dtype=torch.float64, device="cuda", generator=gen)
    s = api.MapSpectrum(ctx, torch.exp(0.5 * g), "f64")
    m = s.convolve("gaussian", 3.0 * np.sqrt(2 * np.log(2)))
    s.close()
    return (m * (5.0 / m.mean())).contiguous()


def _inputs(seed=2001):
    rng = np.random.default_rng(seed)
    v = 200.0 / TSPAN
    margin = 216.0
    traj = np.stack([rng.uniform(margin, N - 1 - margin, NTRAJ), rng.uniform(margin, N - 1 - margin, NTRAJ),
                     rng.uniform(0, 2 * np.pi, NTRAJ), np.full(NTRAJ, v), np.full(NTRAJ, R0),
                     np.zeros(NTRAJ), rng.normal(0, 0.05, NTRAJ)], 1)
    params = np.repeat(traj, len(SHAPES), axis=0)
    prof = np.tile(np.arange(len(SHAPES), dtype=np.int32), NTRAJ)
    times = [np.sort(rng.uniform(0, TSPAN, EPOCHS)) for _ in BAND_SCALE]
    mobs = [rng.normal(-1.75, 0.1, EPOCHS) for _ in BAND_SCALE]
    sig = [np.full(EPOCHS, 0.05) for _ in BAND_SCALE]
    return params, prof, times, mobs, sig


def test_precision_4096_full_shape(ctx, orc):
    m = _map(ctx)
    params, prof, times, mobs, sig = _inputs()
    n = params.shape[0]
    obs = lcm.Observations(ctx, 3, [[(times[b], mobs[b], sig[b]) for b in range(3)]])
    P = torch.tensor(params, device="cuda")
    pr = torch.tensor(prof, device="cuda")
    oi = torch.zeros(n, dtype=torch.int32, device="cuda")
    rr = [[[R0 * s] for s in BAND_SCALE] for _ in SHAPES]
    out = {}
    for prec in ("f32", "f64"):
        print(f"precision 4096^2: {prec} fused path", flush=True)
        spec = api.MapSpectrum(ctx, m, prec)
        nodes = lcm.NodeSet(ctx, spec, SHAPES, rr, BAND_SCALE, compute=False)
        c, g = lcm.chi2_grad(ctx, nodes, obs, P, pr, oi, convolve=True)
        _, _, sk = api.sort_f64(ctx, c, sorted_keys=True)
        lo, npb, _, _ = api.bin_table_device(ctx, sk, api.mpd_bins(100, 1.0, 1.0e7, True))
        torch.cuda.synchronize()
        out[prec] = dict(c=c.cpu().numpy(), g=g.cpu().numpy(), lo=lo.cpu().numpy(), npb=npb.cpu().numpy(),
                         nodes=nodes, spec=spec)
    c32, c64 = out["f32"]["c"], out["f64"]["c"]
    fin = np.isfinite(c64)
    assert fin.all(), "every trajectory stays on the map"
    rel = np.abs(c32 - c64) / np.abs(c64)
    print(f"precision 4096^2: FP32 vs FP64 chi2 max rel err {rel.max():.3e}", flush=True)
    assert rel.max() <= 1e-5
    assert np.array_equal(out["f32"]["lo"], out["f64"]["lo"])
    assert np.array_equal(out["f32"]["npb"], out["f64"]["npb"])
    gs = np.abs(out["f64"]["g"]).max(0)
    gerr = (np.abs(out["f32"]["g"] - out["f64"]["g"]) / gs).max()
    assert gerr < 1e-4, gerr

    # FP64 device path against the host oracle on the same FP64 node maps
    n64 = out["f64"]["nodes"]
    maps = n64.rowmajor().cpu().numpy()
    off = np.cumsum([0] + [len(t) for t in times]).astype(np.int32)
    sample = np.random.default_rng(5).choice(n, 2400, replace=False)
    sample.sort()
    rc, rg = orc.lc_chi2_grad(maps, 3, n64.node_off, n64.node_cnt, n64.node_rad, None, n64.band_scale,
                              params[sample], prof[sample], np.zeros(sample.size, np.int32), off,
                              np.concatenate(times), np.concatenate(mobs), np.concatenate(sig))
    assert np.max(np.abs(out["f64"]["c"][sample] - rc) / np.abs(rc)) < 1e-11
    scale = np.maximum(np.abs(rg).max(0), 1e-300)
    assert np.max(np.abs(out["f64"]["g"][sample] - rg) / scale) < 1e-10
    for prec in ("f32", "f64"):
        out[prec]["spec"].close()
