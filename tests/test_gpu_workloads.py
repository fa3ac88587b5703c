"""The workloads other than the benchmarked one, run at their own shapes as
parity cases:

  plain 1000^2       1000^2 FP64 map, Gaussian source (sigma 5 px), 1000
                     trajectories of 200 px, 50 irregular epochs, chi2 sorted
                     into 100 bins: FP64 and FP32 paths against the FP64
                     oracle, order and bins bit-exact
  photosphere 2048^2 2048^2 FP32 map, expanding photosphere r(t) = r0 + w t
                     (r0 1 px, w in 0.05..0.5 px/day), 16384 trajectories:
                     the FP32 adjoint against the FP64 adjoint, the FP64
                     adjoint against central differences (h = 1e-3)
  batching 64 maps   64 maps of 1024^2..4096^2 (16 systems x 4 images),
                     20..400 epochs per curve, all curves in one batched
                     call; the whole step captured as one HIP graph
  scale 10000^2      10000^2 FP32 maps of 4 images x 5 bands with disc
                     profiles of 1..20 px, 1e6 trajectories per image, 200
                     epochs, 1000-bin table over all 4e6 chi2, the whole step
                     captured as one HIP graph: replays bit-equal to eager,
                     then size-independent properties (FP32 against FP64 on a
                     sample, sortedness and stable order against numpy, bin
                     table against a host search of the sorted keys)
The workload builders are bench.py's (batch_64_maps_workload,
scale_10000_workload).
"""
import numpy as np
import pytest
import torch

from microsn_amd import api
from microsn_amd import lightcurves as lcm

pytestmark = pytest.mark.gpu


def _lognormal_map(ctx, N, seed, mean=5.0, smooth_sigma=3.0):
    """exp(0.5 g) smoothed with a Gaussian of smooth_sigma px (on the
    device, FP64), scaled to the given mean."""
    gen = torch.Generator(device="cuda").manual_seed(seed)
    g = torch.randn((N, N), dtype=torch.float64, device="cuda", generator=gen)
    f = torch.exp(0.5 * g)
    s = api.MapSpectrum(ctx, f, "f64")
    m = s.convolve("gaussian", smooth_sigma * np.sqrt(2 * np.log(2)))
    s.close()
    return m * (mean / m.mean())


def _trajectories(rng, n, N, L, T, margin):
    x0 = rng.uniform(margin, N - 1 - margin, n)
    y0 = rng.uniform(margin, N - 1 - margin, n)
    th = rng.uniform(0, 2 * np.pi, n)
    return np.stack([x0, y0, th, np.full(n, L / T), np.zeros(n), np.zeros(n), rng.normal(0, 0.05, n)], 1)


def _bilinear(m, x, y):
    ix, iy = np.floor(x).astype(int), np.floor(y).astype(int)
    fx, fy = x - ix, y - iy
    return ((1 - fy) * ((1 - fx) * m[iy, ix] + fx * m[iy, ix + 1]) +
            fy * ((1 - fx) * m[iy + 1, ix] + fx * m[iy + 1, ix + 1]))


def test_plain_1000_end_to_end(ctx, orc):
    N, n, E = 1000, 1000, 50
    rng = np.random.default_rng(1000)
    m = _lognormal_map(ctx, N, 1000).cpu().numpy()
    rhalf = 5.0 * np.sqrt(2 * np.log(2))          # Gaussian sigma 5 px
    prof, dprof = orc.profile("gaussian", rhalf, 1.0, deriv=True)
    val = orc.fft_convolve(m, orc.set_kernel(prof, N, N))
    der = orc.fft_convolve(m, orc.set_kernel(dprof, N, N))
    maps = np.stack([val, der], -1)[None]
    T = 100.0
    t = np.sort(rng.uniform(0, T, E))
    prm = _trajectories(rng, n, N, 200.0, T, 230.0)
    prm[:, 4] = rhalf
    truth = np.array([N / 2, N / 2 + 3.3, 0.4, 2.0, rhalf, 0.0, 0.0])
    xt = truth[0] + truth[3] * np.cos(truth[2]) * t
    yt = truth[1] + truth[3] * np.sin(truth[2]) * t
    mobs = -2.5 * np.log10(_bilinear(val, xt, yt)) + rng.normal(0, 0.05, E)
    sig = np.full(E, 0.05)
    zeros = np.zeros(n, np.int32)
    rc, rg = orc.lc_chi2_grad(maps, 1, [0], [1], [rhalf], None, [1.0], prm, zeros, zeros,
                              np.array([0, E], np.int32), t, mobs, sig)
    assert np.all(np.isfinite(rc))
    bins = orc.mpd_bins(100, 1.0, 1e6)
    r_lo, r_npb = orc.bin_edges(rc[orc.stable_argsort(rc)], bins)

    obs = lcm.Observations(ctx, 1, [[(t, mobs, sig)]])
    P = torch.tensor(prm, device="cuda")
    pr = torch.zeros(n, dtype=torch.int32, device="cuda")
    for prec, tol in (("f64", 1e-11), ("f32", 1e-5)):
        spec = api.MapSpectrum(ctx, torch.tensor(m, device="cuda"), prec)
        nodes = lcm.NodeSet(ctx, spec, ["gaussian"], [[[rhalf]]], [1.0])
        for plain in ((False, True) if prec == "f64" else (False,)):
            print("plain 1000^2", prec, "plain" if plain else "fused", flush=True)
            c, g = lcm.chi2_grad(ctx, nodes, obs, P, pr, pr, plain=plain)
            cc = c.cpu().numpy()
            assert np.max(np.abs(cc - rc) / rc) < tol
            if prec == "f64":
                scale = np.abs(rg).max(axis=1, keepdims=True)
                assert np.max(np.abs(g.cpu().numpy() - rg) / scale) < 1e-9
            # device sort and bin table: bit-exact with std::stable_sort and
            # std::upper_bound on the same keys, and the same bins as the oracle's chi2
            perm, _, sk = api.sort_f64(ctx, c, sorted_keys=True)
            want = orc.stable_argsort(cc)
            assert np.array_equal(perm.cpu().numpy().astype(np.uint32), want)
            lo, npb = api.bin_edges(ctx, sk, bins)
            o_lo, o_npb = orc.bin_edges(cc[want], bins)
            assert np.array_equal(lo, o_lo) and np.array_equal(npb, o_npb)
            assert np.array_equal(lo, r_lo) and np.array_equal(npb, r_npb)
            sb = api.Chi2SortBins(ctx, 1, bins.size)
            sb.lower_ind[:], sb.n_per_bin[:] = lo, npb
            bc, be = api.bin_chi2_GPU(ctx, sb, sk)
            obc, obe = orc.bin_chi2(cc[want], lo, npb)
            assert np.allclose(bc, obc, rtol=1e-12, atol=0) and np.allclose(be, obe, rtol=1e-12, atol=1e-300)
        spec.close()


def test_photosphere_2048_differentiation(ctx):
    N, n, E, T = 2048, 16384, 40, 20.0
    rng = np.random.default_rng(2002)
    m = _lognormal_map(ctx, N, 2002)
    radii = [1.0, 1.5, 2.0, 2.75, 3.5, 4.5, 5.5, 6.75, 8.0, 9.5, 11.0, 12.5]
    t = np.sort(rng.uniform(0, T, E))
    sets = [[(t, rng.normal(-1.7, 0.05, E), np.full(E, 0.05))]]
    obs = lcm.Observations(ctx, 1, sets)
    prm = _trajectories(rng, n, N, 60.0, T, 80.0)
    prm[:, 4] = 1.0                                   # r0 = 1 px
    prm[:, 5] = rng.uniform(0.05, 0.5, n)             # w: photosphere growth, px/day
    P = torch.tensor(prm, device="cuda")
    pr = torch.zeros(n, dtype=torch.int32, device="cuda")
    out = {}
    for prec in ("f32", "f64"):
        spec = api.MapSpectrum(ctx, m, prec)
        nodes = lcm.NodeSet(ctx, spec, ["uniform"], [[radii]], [1.0])
        print("photosphere 2048^2", prec, flush=True)
        out[prec] = lcm.chi2_grad(ctx, nodes, obs, P, pr, pr)
        if prec == "f64":
            nodes64 = nodes
    c32, g32 = (x.cpu().numpy() for x in out["f32"])
    c64, g64 = (x.cpu().numpy() for x in out["f64"])
    assert np.all(np.isfinite(c64))
    assert np.max(np.abs(c32 - c64) / c64) <= 1e-5
    scale = np.abs(g64).max(axis=1, keepdims=True)
    assert np.max(np.abs(g32 - g64) / scale) < 1e-3
    assert np.median(np.abs(g32 - g64) / scale) < 1e-5
    # central differences of the FP64 chi2, h = 1e-3 in every parameter, on
    # 256 curves; bilinear kinks at pixel lines make a few differences
    # inexact, so the check is on the bulk
    k = 256
    Pk, prk = P[:k].clone(), pr[:k]
    h = 1e-3
    err = []
    for q in range(7):
        Pp, Pm = Pk.clone(), Pk.clone()
        Pp[:, q] += h
        Pm[:, q] -= h
        cp, _ = lcm.chi2_grad(ctx, nodes64, obs, Pp, prk, prk)
        cm, _ = lcm.chi2_grad(ctx, nodes64, obs, Pm, prk, prk)
        fd = ((cp - cm) / (2 * h)).cpu().numpy()
        err.append(np.abs(fd - g64[:k, q]) / scale[:k, 0])
    err = np.array(err)
    assert np.median(err) < 1e-3, np.median(err, axis=1)
    assert np.quantile(err, 0.9) < 2e-2, np.quantile(err, 0.9, axis=1)


def test_scale_10000_captured(ctx):
    """Four 10000^2 FP32 maps (one per image) x 5 bands of discs 1..20 px,
    1e6 trajectories per image, 200 epochs: map spectra, 20 node maps, the
    batched chi2 + adjoint of all 4e6 curves, their stable sort and a
    1000-bin table captured as one HIP graph; replays equal the eager step
    bit for bit.  Then FP32 against FP64 on a sample, the device sort against
    numpy's stable argsort and the bin table against a host search."""
    import bench

    w = bench.scale_10000_workload(ctx)
    print("scale 10000^2 eager step", flush=True)
    eager = [x.clone() for x in bench.batch_step(ctx, w)]
    torch.cuda.synchronize()
    print("scale 10000^2 capture", flush=True)
    g, out, gen = bench.capture(ctx, lambda: bench.batch_step(ctx, w))
    for r in range(2):
        assert ctx.generation() == gen
        for x in out:
            x.zero_()
        g.replay()
        torch.cuda.synchronize()
        for a, b in zip(out, eager):
            assert torch.equal(a, b), r
    c, gr, perm, sk = eager[:4]
    assert bool(torch.isfinite(c).all()) and bool(torch.isfinite(gr).all())
    # FP32 against FP64 on 20000 curves of image 0 (FP64 node maps)
    ntraj, radii = w["ntraj"], w["radii"]
    s64 = api.MapSpectrum(ctx, w["maps"][0], "f64")
    n64 = lcm.NodeSet(ctx, s64, ["uniform"], [[[r] for r in radii]], radii)
    rng = np.random.default_rng(4005)
    idx = torch.tensor(rng.choice(ntraj, 20000, replace=False), device="cuda")
    z = torch.zeros(20000, dtype=torch.int32, device="cuda")
    c64, _ = lcm.chi2_grad(ctx, n64, w["obs"], w["params"][idx].contiguous(), z, z)
    rel = ((c[idx] - c64).abs() / c64.abs()).max()
    assert float(rel) <= 1e-5, float(rel)
    del n64
    s64.close()
    host = c.cpu().numpy()
    print("scale 10000^2 sort", host.size, flush=True)
    want = np.argsort(host, kind="stable")
    assert np.array_equal(perm.cpu().numpy().astype(np.int64), want)
    sk_h = sk.cpu().numpy()
    assert np.array_equal(sk_h, host[want])
    lo, npb, bc = (eager[4].cpu().numpy(), eager[5].cpu().numpy(), eager[6].cpu().numpy())
    up = np.searchsorted(sk_h, w["bins"], side="right")
    assert np.array_equal(np.concatenate([[0], up[:-1]]), lo.astype(np.int64))
    assert np.array_equal(np.diff(np.concatenate([[0], up])), npb.astype(np.int64))
    for i in np.nonzero(npb)[0][::50]:
        seg = sk_h[lo[i]: lo[i] + npb[i]]
        assert abs(bc[i] - seg.mean()) <= 1e-12 * abs(seg.mean())
    for sp in w["specs"]:
        sp.close()


def test_batching_64_maps_one_call(ctx):
    """64 maps of 1024^2..4096^2 with kappa, gamma in 0.2..0.6 ((1-k)^2-g^2 >
    0.05), 16 systems of 4 images with flux ratios log-uniform in 0.1..1 (the
    per-image magnitude offset), light curves of 20..400 epochs, all curves
    of all maps in one batched call: FP32 within 1e-5 of FP64 on every curve,
    the pooled bin tables identical, every map's curves equal to a per-map
    call bit for bit at the same launch shape, and the whole step (64 map
    spectra, 64 node maps, the batched call, sort and bins) captured as one
    HIP graph whose replays equal the eager step."""
    import bench

    ctx.tune_force(16, 1)
    try:
        res = {}
        for prec in ("f32", "f64"):
            w = bench.batch_64_maps_workload(ctx, prec)
            print("batching 64 maps", prec, "maps", len(w["maps"]), flush=True)
            res[prec] = (w, [x.clone() for x in bench.batch_step(ctx, w)])
        w, out32 = res["f32"]
        _, out64 = res["f64"]
        c32, c64 = out32[0], out64[0]
        rel = ((c32 - c64).abs() / c64.abs()).max()
        assert float(rel) <= 1e-5, float(rel)
        assert torch.equal(out32[4], out64[4]) and torch.equal(out32[5], out64[5])
        midx = w["map_idx"].cpu().numpy()
        for m, ns in enumerate(w["nodesets"]):
            sel = torch.tensor(np.nonzero(midx == m)[0], device="cuda")
            rc, rg = lcm.chi2_grad(ctx, ns, w["obs"], w["params"][sel].contiguous(), w["prof"][sel].contiguous(),
                                   w["obs_idx"][sel].contiguous())
            assert torch.equal(rc, c32[sel]) and torch.equal(rg, out32[1][sel]), m
        g, out, gen = bench.capture(ctx, lambda: bench.batch_step(ctx, w))
        for r in range(2):
            assert ctx.generation() == gen
            for x in out:
                x.zero_()
            g.replay()
            torch.cuda.synchronize()
            for a, b in zip(out, out32):
                assert torch.equal(a, b), r
        for prec in res:
            for sp in res[prec][0]["specs"]:
                sp.close()
    finally:
        ctx.tune_mode(ctx.TUNE_AUTO)
