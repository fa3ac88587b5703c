"""Light curves/s with chi2 + gradient at a 4096x4096 map (MI355X, one GPU).

Workload (one "step" = one pass of the hot path over the batch, inputs
resident in HBM):
  * map 4096 x 4096: log-normal field exp(0.5 g), g ~ N(0,1) (seed 1001),
    smoothed with a circular Gaussian of sigma 3 px, scaled to mean
    1/((1-k)^2-g^2) = 5 (k = g = 0.4)
  * three source profiles (Gaussian, uniform disc, thin disc) at three bands
    with half-light radii 2 / 4 / 8 px -> 9 convolved (mu, dmu/dr) maps
  * 100 000 straight trajectories (origin and angle uniform, 200 px over the
    100-day baseline), each with the three profiles -> 300 000 light curves
  * 100 irregular epochs per band (300 per curve) against one synthetic
    observed curve (0.05 mag Gaussian noise)
  * per step: map spectrum, 9 fused convolutions, chi2 + 7-parameter adjoint
    of every curve, stable sort of the chi2 values and their 100-bin table
The metric counts light curves (1 trajectory x 1 profile, all bands and
epochs) per second.  The same step is also timed on the FP64 plain path
(reference-structured: full-size kernels, separate multiply, synchronised
stages, two-launch chi2), and the FP64 CPU oracle is timed on a bounded
sample of the same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N = 4096
NTRAJ = 100_000
EPOCHS = 100
TSPAN = 100.0
SHAPES = ("gaussian", "uniform", "exponential")
BAND_SCALE = (1.0, 2.0, 4.0)
R0 = 2.0
NBINS = 100
HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def make_map(seed=1001, n=N, sigma=3.0, mean=5.0):
    rng = np.random.default_rng(seed)
    f = np.exp(0.5 * rng.standard_normal((n, n)))
    k = np.fft.fftfreq(n)
    kr = np.fft.rfftfreq(n)
    H = np.exp(-2.0 * np.pi ** 2 * sigma ** 2 * (k[:, None] ** 2 + kr[None, :] ** 2))
    f = np.fft.irfft2(np.fft.rfft2(f) * H, s=(n, n))
    return f * (mean / f.mean())


def make_inputs(seed=2001):
    rng = np.random.default_rng(seed)
    L = 200.0
    v = L / TSPAN
    margin = L + 16.0
    x0 = rng.uniform(margin, N - 1 - margin, NTRAJ)
    y0 = rng.uniform(margin, N - 1 - margin, NTRAJ)
    th = rng.uniform(0.0, 2.0 * np.pi, NTRAJ)
    dm = rng.normal(0.0, 0.05, NTRAJ)
    traj = np.stack([x0, y0, th, np.full(NTRAJ, v), np.full(NTRAJ, R0), np.zeros(NTRAJ), dm], 1)
    params = np.repeat(traj, len(SHAPES), axis=0)  # curve = (trajectory, profile)
    prof = np.tile(np.arange(len(SHAPES), dtype=np.int32), NTRAJ)
    times = [np.sort(rng.uniform(0.0, TSPAN, EPOCHS)) for _ in BAND_SCALE]
    truth = np.array([N / 2 + 3.3, N / 2 - 7.1, 0.7, v, R0, 0.0, 0.0])
    return params, prof, times, truth, rng


def bilinear(img, x, y):
    ix = np.minimum(np.floor(x).astype(int), img.shape[1] - 2)
    iy = np.minimum(np.floor(y).astype(int), img.shape[0] - 2)
    fx, fy = x - ix, y - iy
    return ((1 - fy) * ((1 - fx) * img[iy, ix] + fx * img[iy, ix + 1])
            + fy * ((1 - fx) * img[iy + 1, ix] + fx * img[iy + 1, ix + 1]))


def algorithmic_bytes(name, esz, ncurves):
    """Bytes one launch must move at least (value bytes of its operands).

    A chi2 launch covers one (profile, band): the curves of that profile,
    one band's epochs each.  It must read its node map of (mu, dmu/dr) once
    and every curve's 64-byte record and 8 FP64 accumulators (read and
    written); the bilinear taps its samples gather (tap_bytes) come through
    the caches, most of them many times over."""
    W = N // 2 + 1
    spec = N * W * 2 * esz          # one half spectrum
    real = N * N * esz              # one real map
    curves = ncurves // len(SHAPES)
    return {
        "conv_node_col": spec + 2 * spec,              # read map spectrum, write 2 spectra
        "conv_node_yfilter": spec + 2 * spec,          # read row spectrum, write 2 spectra
        "conv_node_fused": spec + 2 * real,            # read row spectrum, write (mu, dmu/dr)
        "fft_row_c2r_node": 2 * spec + 2 * real,       # read 2 spectra, write (mu, dmu/dr)
        "fft_row_r2c": N * N * 8 + spec,               # read FP64 map, write spectrum
        "lc_chi2_grad": 2 * real + curves * (64 + 2 * 64),
        # the FP64 plain path's stages (esz = 8)
        "fft_col_fwd": 2 * spec,                       # read + write the spectrum, in place
        "fft_col_inv": 2 * spec,
        "fft_row_c2r": spec + real,                    # read spectrum, write real map
        "plain_make_kernel": real,                     # write the full-size kernel array
        "plain_cmul": 3 * spec,                        # read two spectra, write one
        "plain_interleave": 4 * real,                  # read value + derivative, write the pairs
        # per (curve, epoch) of all bands: 4 taps of (mu, dmu/dr), write 4 FP64 terms
        "lc_plain_samples": ncurves * EPOCHS * len(BAND_SCALE) * (4 * 2 * esz + 32),
        "lc_plain_reduce": ncurves * EPOCHS * len(BAND_SCALE) * (32 + 24) + ncurves * 64,
    }.get(name)


# What bounds the two dominant kernels below the HBM roofline, from their
# counters (profiles/counters_4096_f32.md).
LIMITER = {
    "lc_chi2_grad": "L1-to-L2 read requests: nearly every lane's tap load misses L1 "
                    "(about 0.65 L2 requests per sample, 0.14 per clock per CU at 370 clocks of latency); "
                    "its VALU alone is about half the launch",
    "conv_node_fused": "filter phase: L1-to-L2 read requests (0.13 per clock per CU); FFT phase: VALU "
                       "(no packed FP32)",
}


def tap_bytes(esz, ncurves):
    """Bytes of bilinear taps one chi2 launch gathers: 4 (mu, dmu/dr) taps per
    sample (L2/L1 traffic, not HBM)."""
    return ncurves // len(SHAPES) * EPOCHS * 4 * 2 * esz


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE x 2 + WRITE_SIZE, calibrated on a device copy), or
    None when it holds no figure for the kernel."""
    path = os.path.join(ROOT, "profiles", "traffic_4096_f32.json")
    try:
        with open(path) as f:
            return json.load(f)["bytes_per_launch"].get(kernel)
    except (OSError, ValueError, KeyError):
        return None


def run_step(ctx, ws, plain=False):
    import microsn_amd.api as api
    import microsn_amd.lightcurves as lcm

    ws["spec"].update(ws["map"])
    if plain:
        ws["nodes"].compute(plain)
    # fused path: each node map is convolved right before the band launch that reads it
    chi2, grad = lcm.chi2_grad(ctx, ws["nodes"], ws["obs"], ws["params"], ws["prof"], ws["obs_idx"],
                               plain=plain, convolve=not plain)
    perm, _, sk = api.sort_f64(ctx, chi2, sorted_keys=True)
    # bin table and binned means on the device: no host round trip in a step
    lo, npb, _, be = api.bin_table_device(ctx, sk, ws["bins"])
    return chi2, grad, lo, npb, be


def setup(ctx, prec, m_dev, params, prof, times, mobs, sig):
    import torch

    import microsn_amd.api as api
    import microsn_amd.lightcurves as lcm

    spec = api.MapSpectrum(ctx, m_dev, prec)
    rr = [[[R0 * sc] for sc in BAND_SCALE] for _ in SHAPES]
    nodes = lcm.NodeSet(ctx, spec, SHAPES, rr, BAND_SCALE, compute=False)
    obs = lcm.Observations(ctx, len(BAND_SCALE), [[(times[b], mobs[b], sig[b]) for b in range(len(BAND_SCALE))]])
    bins = api.mpd_bins(NBINS, 1.0, 1.0e7, True)
    return dict(spec=spec, nodes=nodes, obs=obs, map=m_dev, bins=bins,
                params=torch.tensor(params, device="cuda"), prof=torch.tensor(prof, device="cuda"),
                obs_idx=torch.zeros(len(prof), dtype=torch.int32, device="cuda"))


def time_steps(ctx, ws, steps, warmup, plain, timing):
    import torch

    for _ in range(warmup):
        run_step(ctx, ws, plain)
    torch.cuda.synchronize()
    if timing:
        ctx.timing_reset()
        ctx.timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = run_step(ctx, ws, plain)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    kern = {}
    if timing:
        ctx.timing(False)
        kern = ctx.timing_all()
    return dt, out, kern


def capture_step(ctx, ws):
    """One FP32 step captured into a HIP graph (torch.cuda.graph on a side
    stream the library is pointed at); the step runs once on that stream
    first so the workspace and device tables are in place.  Returns the
    graph and the output tensors its replays write."""
    import torch

    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    ctx.set_stream(s)
    try:
        with torch.cuda.stream(s):
            run_step(ctx, ws)
        s.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            out = run_step(ctx, ws)
    finally:
        ctx.set_stream(torch.cuda.current_stream())
    return g, out, ctx.generation()


def time_graph(ctx, g, gen, steps, warmup):
    """Replays of a graph captured at context generation gen (the library
    has reallocated or rewritten nothing it refers to since)."""
    import torch

    if ctx.generation() != gen:
        raise RuntimeError("the captured step refers to memory the library has since replaced")
    for _ in range(warmup):
        g.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.replay()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


# ---- the batching (64 maps) and scale (10000^2) workloads: bench extras,
# also run by the configuration tests
SIZES3 = (1024, 1280, 1536, 2048, 2560, 3072, 4096)


def lognormal_map(ctx, n, seed, mean=5.0, smooth_sigma=3.0):
    """exp(0.5 g), g ~ N(0, 1) on the device (seeded), smoothed with a
    Gaussian of smooth_sigma px (FP64 convolution of the library), scaled to
    the given mean."""
    import torch

    import microsn_amd.api as api

    gen = torch.Generator(device="cuda").manual_seed(seed)
    g = torch.randn((n, n), dtype=torch.float64, device="cuda", generator=gen)
    s = api.MapSpectrum(ctx, torch.exp(0.5 * g), "f64")
    m = s.convolve("gaussian", smooth_sigma * np.sqrt(2 * np.log(2)))
    s.close()
    return (m * (mean / m.mean())).contiguous()


def trajectories(rng, n, size, length, tspan, margin):
    """Straight trajectories of the given length over tspan, origins uniform
    at least `margin` px inside a size x size map, angles uniform."""
    x0 = rng.uniform(margin, size - 1 - margin, n)
    y0 = rng.uniform(margin, size - 1 - margin, n)
    th = rng.uniform(0, 2 * np.pi, n)
    return np.stack([x0, y0, th, np.full(n, length / tspan), np.zeros(n), np.zeros(n), rng.normal(0, 0.05, n)], 1)


def batch_64_maps_workload(ctx, prec="f32", ncur=2048, seed=3003):
    """64 maps (16 systems x 4 images) of 1024^2..4096^2, kappa and gamma
    uniform in 0.2..0.6 with (1-k)^2 - g^2 > 0.05 (map mean 1/|(1-k)^2-g^2|),
    flux ratios log-uniform in 0.1..1 as each image's magnitude offset, a
    Gaussian source of 2 px, ncur trajectories per map, each curve with one
    of 64 observation sets of 20..400 irregular epochs."""
    import torch

    import microsn_amd.api as api
    import microsn_amd.lightcurves as lcm

    rng = np.random.default_rng(seed)
    nsys, nimg, nsets = 16, 4, 64
    sets = []
    for _ in range(nsets):
        ne = int(rng.integers(20, 401))
        t = np.sort(rng.uniform(0, 100.0, ne))
        sets.append([(t, rng.normal(-1.0, 0.3, ne), np.full(ne, 0.05))])
    obs = lcm.Observations(ctx, 1, sets)
    maps, specs, nodesets, prm, midx, oidx = [], [], [], [], [], []
    for s in range(nsys):
        ratios = np.exp(rng.uniform(np.log(0.1), np.log(1.0), nimg))
        for i in range(nimg):
            n = int(rng.choice(SIZES3))
            while True:
                k, g = rng.uniform(0.2, 0.6, 2)
                if (1 - k) ** 2 - g ** 2 > 0.05:
                    break
            m = lognormal_map(ctx, n, 3100 + s * nimg + i, mean=1.0 / abs((1 - k) ** 2 - g ** 2))
            spec = api.MapSpectrum(ctx, m, prec)
            maps.append(m)
            specs.append(spec)
            nodesets.append(lcm.NodeSet(ctx, spec, ["gaussian"], [[[2.0]]], [1.0], compute=False))
            p = trajectories(rng, ncur, n, 200.0, 100.0, 230.0)
            p[:, 4] = 2.0
            p[:, 6] = -2.5 * np.log10(ratios[i])
            prm.append(p)
            midx.append(np.full(ncur, len(maps) - 1, np.int32))
            oidx.append(rng.integers(0, nsets, ncur).astype(np.int32))
    t = lambda a: torch.tensor(np.concatenate(a), device="cuda")  # noqa: E731
    return dict(maps=maps, specs=specs, nodesets=nodesets, batch=lcm.NodeBatch(nodesets), obs=obs,
                params=t(prm), map_idx=t(midx), prof=torch.zeros(len(maps) * ncur, dtype=torch.int32, device="cuda"),
                obs_idx=t(oidx), bins=api.mpd_bins(100, 1.0, 1e7))


def scale_10000_workload(ctx, prec="f32", ntraj=1_000_000, seed=4004):
    """Four 10000^2 maps (one per lensed image), uniform-disc sources of
    1, 2.5, 5, 10 and 20 px in five bands of 40 epochs each (200 per curve,
    0.02 mag noise), ntraj trajectories on every image (4 ntraj curves), a
    1000-bin table of all chi2."""
    import torch

    import microsn_amd.api as api
    import microsn_amd.lightcurves as lcm

    n, nimg, nb, E = 10000, 4, 5, 40
    radii = [1.0, 2.5, 5.0, 10.0, 20.0]
    rng = np.random.default_rng(seed)
    times = [np.sort(rng.uniform(0, 100.0, E)) for _ in range(nb)]
    obs = lcm.Observations(ctx, nb, [[(tb, rng.normal(-1.7, 0.02, E), np.full(E, 0.02)) for tb in times]])
    prm = trajectories(rng, ntraj, n, 200.0, 100.0, 230.0)
    prm[:, 4] = 1.0
    maps, specs, nodesets = [], [], []
    for img in range(nimg):
        m = lognormal_map(ctx, n, 4100 + img)
        spec = api.MapSpectrum(ctx, m, prec)
        maps.append(m)
        specs.append(spec)
        nodesets.append(lcm.NodeSet(ctx, spec, ["uniform"], [[[r] for r in radii]], radii, compute=False))
    P = torch.tensor(np.tile(prm, (nimg, 1)), device="cuda")
    M = torch.tensor(np.repeat(np.arange(nimg, dtype=np.int32), ntraj), device="cuda")
    z = torch.zeros(nimg * ntraj, dtype=torch.int32, device="cuda")
    return dict(maps=maps, specs=specs, nodesets=nodesets, batch=lcm.NodeBatch(nodesets), obs=obs, params=P,
                map_idx=M, prof=z, obs_idx=z, bins=api.mpd_bins(1000, 1.0, 1e7), radii=radii, ntraj=ntraj)


def batch_step(ctx, w):
    """One step of a batched workload: every map's spectrum and node maps
    (one batched call each), the chi2 + adjoint of all curves in one batched
    call, the stable sort of
    the chi2 values and their bin table, with no host synchronisation."""
    import microsn_amd.api as api
    import microsn_amd.lightcurves as lcm

    api.update_spectra(w["specs"], w["maps"])
    w["batch"].compute()
    c, g = lcm.chi2_grad_batch(ctx, w["batch"], w["obs"], w["params"], w["map_idx"], w["prof"], w["obs_idx"])
    perm, _, sk = api.sort_f64(ctx, c, sorted_keys=True)
    lo, npb, bc, be = api.bin_table_device(ctx, sk, w["bins"])
    return c, g, perm, sk, lo, npb, bc, be


def capture(ctx, fn):
    """fn() once on a side stream the library is pointed at (workspace,
    tables and kept launch shapes settle), then captured into a HIP graph.
    Returns (graph, the outputs its replays write, context generation)."""
    import torch

    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    ctx.set_stream(s)
    try:
        with torch.cuda.stream(s):
            fn()
        s.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            out = fn()
    finally:
        ctx.set_stream(torch.cuda.current_stream())
    return g, out, ctx.generation()


def time_batch_workload(ctx, which, steps):
    """The batching or scale workload: one eager step (per-kernel events),
    then the step captured as one HIP graph and its replays timed; the
    batched band kernel's roofline from its event-timed launch."""
    import torch

    w = batch_64_maps_workload(ctx) if which == "batch_64_maps" else scale_10000_workload(ctx)
    torch.cuda.synchronize()
    batch_step(ctx, w)  # first step: launch shapes tuned, tables and workspace in place
    torch.cuda.synchronize()
    ctx.timing_reset()
    ctx.timing(True)
    eager = batch_step(ctx, w)
    torch.cuda.synchronize()
    ctx.timing(False)
    kern = ctx.timing_all()
    g, out, gen = capture(ctx, lambda: batch_step(ctx, w))
    same = False
    g.replay()
    torch.cuda.synchronize()
    same = all(torch.equal(a, b) for a, b in zip(out, eager))
    if ctx.generation() != gen:
        raise RuntimeError("the captured step refers to memory the library has since replaced")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.replay()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    ncurves = int(w["params"].shape[0])
    off = w["obs"].off.cpu().numpy()
    oi = w["obs_idx"].cpu().numpy()
    B = w["obs"].B
    samples = int((off[oi * B + B] - off[oi * B]).sum())
    ms, launches = kern["lc_batch"]
    # per step: every node map read once, every record read and accumulator
    # written (per launch of the per-band form: its share); the gathered
    # FP32 (mu, dmu/dr) taps come through the caches
    node_bytes = sum(ns.nodes.numel() * ns.nodes.element_size() for ns in w["nodesets"])
    ab = node_bytes + ncurves * (64 + 64)
    taps = samples * 4 * 8
    dom = max(kern, key=lambda k: kern[k][0])
    max_pix = max(int(m.shape[0]) * int(m.shape[1]) for m in w["maps"])
    lanes, gmul, flags = ctx.tune_lookup(ctx.TUNE_KIND_BATCH, "f32", w["obs"].nepoch / (w["obs"].nobs * B),
                                         max_pix, with_flags=True)
    rec = {"ms_per_step": dt * 1e3, "light_curves_per_s": ncurves / dt, "curves": ncurves,
           "samples": samples, "maps": len(w["maps"]),
           "map_sizes": sorted({int(m.shape[0]) for m in w["maps"]}), "graph_replay_matches_eager": same,
           "steps": steps, "dominant_kernel": dom,
           "kernels_ms_per_step": {k: v[0] for k, v in sorted(kern.items())},
           "lc_batch_shape": {"lanes": lanes, "grid_multiple": gmul,
                              "launches_per_step": launches, "one_launch_per_band": bool(flags & ctx.TUNE_BATCH_PER_BAND)},
           "lc_batch_roofline": {"bound": "hbm", "achieved": ab / (ms / 1e3) / 1e9, "peak": HBM_PEAK / 1e9,
                                 "unit": "GB/s", "frac": ab / (ms / 1e3) / HBM_PEAK,
                                 "algorithmic_bytes_per_launch": ab / launches, "avg_launch_ms": ms / launches,
                                 "tap_bytes_per_step": taps, "tap_rate_GBps": taps / (ms / 1e3) / 1e9}}
    # the dominant kernel's roofline over the step (FP32 maps: 8-byte spectrum
    # bins and 8-byte (mu, dmu/dr) pixels; the map transform reads FP64)
    def step_bytes(name):
        tot = 0
        for m, ns in zip(w["maps"], w["nodesets"]):
            ny, nx = int(m.shape[0]), int(m.shape[1])
            spec, px, nn = ny * (nx // 2 + 1) * 8, ny * nx * 8, int(ns.nodes.shape[0])
            per = {"fft_row_c2r_node": nn * (2 * spec + px), "conv_node_fused": nn * (spec + px),
                   "conv_node_yfilter": nn * 3 * spec, "fft_row_r2c": ny * nx * 8 + spec}.get(name)
            if per is None:
                return None
            tot += per
        return tot
    if dom == "lc_batch":
        rec["dominant_roofline"] = dict(rec["lc_batch_roofline"], kernel=dom)
    elif step_bytes(dom):
        dms, dl = kern[dom]
        rec["dominant_roofline"] = {"kernel": dom, "bound": "hbm", "achieved": step_bytes(dom) / (dms / 1e3) / 1e9,
                                    "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                                    "frac": step_bytes(dom) / (dms / 1e3) / HBM_PEAK,
                                    "algorithmic_bytes_per_step": step_bytes(dom), "launches_per_step": dl,
                                    "ms_per_step": dms}
    for sp in w["specs"]:
        sp.close()
    del w, eager, out, g
    torch.cuda.empty_cache()
    return rec


def time_pair_mode(ctx, steps=3):
    """The reference's pair mode at grid_input.json's shape (9801 locations
    = 9.6e7 image pairs, 5 source sizes, 300 ratio bins, the 5-point observed
    curve): setup_integral (ratios z, their stable sort and the rank of every
    pair, bin table), the pair chi2 in sorted order and the binned means,
    timed per call on synthetic light curves (log-normal point samples, the
    convolved columns a smooth function of size)."""
    import torch

    import microsn_amd.api as api

    nloc, sizes = 9801, np.array([0.0, 0.05, 0.1, 0.2, 0.3])
    rng = np.random.default_rng(5005)
    lcs = []
    for _ in range(2):
        mu0 = np.exp(rng.normal(1.5, 0.6, nloc))
        grow = 1.0 + rng.normal(0.0, 0.05, (nloc, 1)) * sizes[None, 1:] / 0.3
        lc = api.SimLC(ctx, nloc, sizes.size)
        cols = np.concatenate([mu0[:, None], mu0[:, None] * grow], 1)
        lc.LC.copy_(torch.tensor(cols, device="cuda"))
        lc.DLC.copy_(torch.tensor(np.diff(cols, axis=1), device="cuda"))
        lcs.append(lc)
    LCA, LCB = lcs
    bins = api.mpd_bins(300, 0.0001, 300.0)
    sb = api.Chi2SortBins(ctx, nloc, bins.size)
    Dfac = np.sqrt(1750.0 * 850.0 / 1570.0)
    cv = api.setup_chi2_calculation((0.4, 0.5), (1.0,), Dfac, [0.7, 1.2, 1.4, 1.1, 0.9], [5, 7, 12, 14, 17],
                                    [0.1, 0.15, 0.1, 0.1, 0.15], sizes)
    chi2 = torch.empty(nloc * nloc, dtype=torch.float64, device="cuda")
    t = {"setup_integral": 0.0, "calculate_chi2": 0.0, "bin_chi2": 0.0}
    for it in range(steps + 1):  # the first round settles the workspace
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        api.setup_integral_GPU(ctx, LCA, LCB, bins, sb)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        api.calculate_chi2_GPU(ctx, cv, chi2, sb, LCA, LCB)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        api.bin_chi2_GPU(ctx, sb, chi2)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        if it > 0:
            t["setup_integral"] += (t1 - t0) / steps
            t["calculate_chi2"] += (t2 - t1) / steps
            t["bin_chi2"] += (t3 - t2) / steps
    ctx.timing_reset()
    ctx.timing(True)
    api.setup_integral_GPU(ctx, LCA, LCB, bins, sb)
    api.calculate_chi2_GPU(ctx, cv, chi2, sb, LCA, LCB)
    api.bin_chi2_GPU(ctx, sb, chi2)
    torch.cuda.synchronize()
    ctx.timing(False)
    kern = ctx.timing_all()
    total = sum(t.values())
    rec = {"pairs": nloc * nloc, "locations": nloc, "bins": int(bins.size),
           "ms": {k: v * 1e3 for k, v in t.items()}, "ms_total": total * 1e3, "pairs_per_s": nloc * nloc / total,
           "kernels_ms": {k: v[0] for k, v in sorted(kern.items())}}
    del chi2, sb, LCA, LCB, lcs
    torch.cuda.empty_cache()
    return rec


def replay_ms(ctx, fn, steps):
    """fn() captured as one HIP graph (after an eager run that settles the
    launch shapes and tables): ms per replay over `steps` replays, and
    whether a replay reproduces the eager outputs bit for bit."""
    import torch

    eager = [x.clone() for x in fn()]
    torch.cuda.synchronize()
    g, out, gen = capture(ctx, fn)
    g.replay()
    torch.cuda.synchronize()
    same = all(torch.equal(a, b) for a, b in zip(out, eager))
    if ctx.generation() != gen:
        raise RuntimeError("the captured step refers to memory the library has since replaced")
    t0 = time.perf_counter()
    for _ in range(steps):
        g.replay()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps * 1e3, same


def time_plain_1000(ctx, steps, cpu=True):
    """The CPU-runnable workload: 1000^2 map (log-normal, 3 px smoothing,
    mean 5), Gaussian source of sigma 5 px, 1000 trajectories of 200 px over
    100 days, 50 irregular epochs against one observed curve with 0.05 mag
    noise, chi2 + adjoint, sort, 100 bins: GPU FP32 and FP64 fused (graph
    replays), GPU FP64 plain (eager, reference-structured) and the FP64 host
    oracle (OpenMP) on the same step, in light curves/s."""
    import torch

    import microsn_amd.api as api
    import microsn_amd.lightcurves as lcm

    n, ncur, E, T = 1000, 1000, 50, 100.0
    rng = np.random.default_rng(1000)
    m = lognormal_map(ctx, n, 1000)
    rhalf = 5.0 * np.sqrt(2 * np.log(2))
    t = np.sort(rng.uniform(0, T, E))
    prm = trajectories(rng, ncur, n, 200.0, T, 230.0)
    prm[:, 4] = rhalf
    mobs = rng.normal(-1.75, 0.05, E)
    sig = np.full(E, 0.05)
    obs = lcm.Observations(ctx, 1, [[(t, mobs, sig)]])
    P = torch.tensor(prm, device="cuda")
    z = torch.zeros(ncur, dtype=torch.int32, device="cuda")
    bins = api.mpd_bins(100, 1.0, 1e6)
    rec = {"curves": ncur, "epochs": E, "map": n}
    for prec in ("f32", "f64"):
        spec = api.MapSpectrum(ctx, m, prec)
        nodes = lcm.NodeSet(ctx, spec, ["gaussian"], [[[rhalf]]], [1.0], compute=False)

        def step(spec=spec, nodes=nodes):
            spec.update(m)
            c, g = lcm.chi2_grad(ctx, nodes, obs, P, z, z, convolve=True)
            _, _, sk = api.sort_f64(ctx, c, sorted_keys=True)
            lo, npb, _, be = api.bin_table_device(ctx, sk, bins)
            return c, g, lo, npb, be
        ms, same = replay_ms(ctx, step, steps)
        rec[f"gpu_{prec}_fused"] = {"ms_per_step": ms, "light_curves_per_s": ncur / ms * 1e3,
                                    "graph_replay_matches_eager": same}
        if prec == "f64":
            def plain(spec=spec, nodes=nodes):
                spec.update(m)
                nodes.compute(plain=True)
                c, g = lcm.chi2_grad(ctx, nodes, obs, P, z, z, plain=True)
                _, _, sk = api.sort_f64(ctx, c, sorted_keys=True)
                return api.bin_table_device(ctx, sk, bins)
            plain()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                plain()
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) / steps * 1e3
            rec["gpu_f64_plain"] = {"ms_per_step": ms, "light_curves_per_s": ncur / ms * 1e3}
        spec.close()
    if cpu:
        from oracle import oracle as orc
        mh = m.cpu().numpy()
        t0 = time.perf_counter()
        v, dv = orc.profile("gaussian", rhalf, 1.0, deriv=True)
        maps = np.stack([orc.fft_convolve(mh, orc.set_kernel(v, n, n)),
                         orc.fft_convolve(mh, orc.set_kernel(dv, n, n))], -1)[None]
        zz = np.zeros(ncur, np.int32)
        c, _ = orc.lc_chi2_grad(maps, 1, [0], [1], [rhalf], None, [1.0], prm, zz, zz, np.array([0, E], np.int32),
                                t, mobs, sig)
        sk = c[orc.stable_argsort(c)]
        lo, npb = orc.bin_edges(sk, bins)
        orc.bin_chi2(sk, lo, npb)
        dt = time.perf_counter() - t0
        rec["cpu_oracle_f64"] = {"ms_per_step": dt * 1e3, "light_curves_per_s": ncur / dt,
                                 "threads": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))}
    return rec


def time_photosphere_2048(ctx, steps):
    """The differentiation workload: 2048^2 map, expanding uniform-disc
    photosphere r(t) = r0 + w t (r0 = 1 px, w uniform in 0.05..0.5 px/day,
    12 radius nodes from 1 to 12.5 px), 16384 trajectories of 60 px over 20
    days, 40 epochs: one step (map, 12 node maps, chi2 + 7-parameter
    adjoint) in FP32 and FP64, and, for scale, the 13 chi2 evaluations that
    central differences in the seven parameters would need (one evaluation
    per side of six of them plus the centre)."""
    import torch

    import microsn_amd.api as api
    import microsn_amd.lightcurves as lcm

    n, ncur, E, T = 2048, 16384, 40, 20.0
    rng = np.random.default_rng(2002)
    m = lognormal_map(ctx, n, 2002)
    radii = [1.0, 1.5, 2.0, 2.75, 3.5, 4.5, 5.5, 6.75, 8.0, 9.5, 11.0, 12.5]
    t = np.sort(rng.uniform(0, T, E))
    obs = lcm.Observations(ctx, 1, [[(t, rng.normal(-1.7, 0.05, E), np.full(E, 0.05))]])
    prm = trajectories(rng, ncur, n, 60.0, T, 80.0)
    prm[:, 4] = 1.0
    prm[:, 5] = rng.uniform(0.05, 0.5, ncur)
    P = torch.tensor(prm, device="cuda")
    z = torch.zeros(ncur, dtype=torch.int32, device="cuda")
    rec = {"curves": ncur, "epochs": E, "map": n, "nodes": len(radii)}
    for prec in ("f32", "f64"):
        spec = api.MapSpectrum(ctx, m, prec)
        nodes = lcm.NodeSet(ctx, spec, ["uniform"], [[radii]], [1.0], compute=False)

        def step(spec=spec, nodes=nodes):
            spec.update(m)
            return lcm.chi2_grad(ctx, nodes, obs, P, z, z, convolve=True)
        ms, same = replay_ms(ctx, step, steps)

        # the adjoint against central differences on the same node maps: one
        # chi2 + gradient call, and the 13 chi2 calls differences would take
        ms_one, _ = replay_ms(ctx, lambda nodes=nodes: lcm.chi2_grad(ctx, nodes, obs, P, z, z), steps)

        def fd13(nodes=nodes):
            return [lcm.chi2_grad(ctx, nodes, obs, P, z, z)[0] for _ in range(13)]
        ms_fd, _ = replay_ms(ctx, fd13, steps)
        rec[prec] = {"ms_per_step": ms, "light_curves_per_s": ncur / ms * 1e3, "graph_replay_matches_eager": same,
                     "ms_chi2_and_adjoint_call": ms_one, "ms_13_chi2_calls": ms_fd}
        spec.close()
    return rec


def cpu_model():
    """The host CPU's model name (lscpu's "Model name", from /proc/cpuinfo),
    with its family, model and stepping numbers (a virtual machine may report
    only a short name)."""
    info = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" not in line:
                    if info:
                        break  # the first processor's block is enough
                    continue
                k, v = (x.strip() for x in line.split(":", 1))
                info.setdefault(k, v)
    except OSError:
        return None
    name = info.get("model name")
    if name is None:
        return None
    ids = [f"{k} {info[k]}" for k in ("cpu family", "model", "stepping") if k in info]
    return f"{name} ({', '.join(ids)})" if ids else name


def cpu_baseline(m, params, prof, times, mobs, sig, bins):
    """One whole step of the workload on the host with the FP64 oracle (C++,
    OpenMP): the 9 node maps (value and d/dr, each an FFT convolution of the
    map), chi2 and adjoint of all curves, the stable sort of the chi2 values
    and the bin table.  Returns the baseline record and the host chi2."""
    from oracle import oracle as orc

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    B, P = len(BAND_SCALE), len(SHAPES)
    t0 = time.perf_counter()
    nodes = np.empty((P * B, N, N, 2))
    for p, shape in enumerate(SHAPES):
        for b, sc in enumerate(BAND_SCALE):
            v, dv = orc.profile(shape, R0 * sc, 1.0, deriv=True)
            nodes[p * B + b, ..., 0] = orc.fft_convolve(m, orc.set_kernel(v, N, N))
            nodes[p * B + b, ..., 1] = orc.fft_convolve(m, orc.set_kernel(dv, N, N))
    t_conv = time.perf_counter() - t0
    off = np.cumsum([0] + [len(t) for t in times]).astype(np.int32)
    tt, mo, sg = np.concatenate(times), np.concatenate(mobs), np.concatenate(sig)
    t1 = time.perf_counter()
    chi2, _ = orc.lc_chi2_grad(nodes, B, np.arange(P * B, dtype=np.int32), np.ones(P * B, np.int32),
                               np.array([R0 * sc for _ in SHAPES for sc in BAND_SCALE]),
                               np.arange(P * B, dtype=np.int32), np.array(BAND_SCALE), params, prof,
                               np.zeros(len(prof), np.int32), off, tt, mo, sg)
    t_curves = time.perf_counter() - t1
    t2 = time.perf_counter()
    sk = chi2[orc.stable_argsort(chi2)]
    lo, npb = orc.bin_edges(sk, bins)
    orc.bin_chi2(sk, lo, npb)
    t_sort = time.perf_counter() - t2
    t_step = time.perf_counter() - t0
    ncurves = len(prof)
    return {"value": ncurves / t_step, "unit": "light curves/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "sample": (f"one whole step on this machine's host CPUs with the FP64 oracle: 9 node maps as 18 "
                       f"4096^2 FFT convolutions ({t_conv:.2f} s; each transforms the map again, "
                       f"54 2-D FFTs where 37 would do), chi2 + adjoint of {ncurves} curves x 300 "
                       f"epochs ({t_curves:.2f} s), stable sort + bin table ({t_sort:.2f} s); "
                       f"{t_step:.1f} s in all")}, chi2


def pmc_step(lanes, grid_multiple):
    """Two FP32 steps of the workload, then a 512 MiB device-to-device copy
    whose FETCH_SIZE / WRITE_SIZE calibrate the byte counters.  The band
    launches use the given shape (the one the bench's own tuner keeps for this
    workload, reported as band_launch_shape): under counter collection the
    tuner's timings, and so its choice, would differ from run to run."""
    import torch

    import microsn_amd.api as api

    ctx = api.Context(0)
    ctx.tune_force(lanes, grid_multiple)
    m_dev = torch.tensor(make_map(), device="cuda")
    params, prof, times, _, _ = make_inputs()
    mobs = [np.full(len(t), -1.7) for t in times]
    sig = [np.full(len(t), 0.05) for t in times]
    ws = setup(ctx, "f32", m_dev, params, prof, times, mobs, sig)
    for _ in range(2):
        run_step(ctx, ws)
    torch.cuda.synchronize()
    a = torch.empty(128 * 1024 * 1024, dtype=torch.float32, device="cuda")
    b = torch.ones_like(a)
    a.copy_(b)
    torch.cuda.synchronize()
    print("pmc step done", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the FP64 plain leg")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batching (64 maps) and scale (10000^2) workloads")
    ap.add_argument("--extras-only", action="store_true",
                    help="run only the batching, scale and pair-mode workloads (for profiling them)")
    ap.add_argument("--eager", action="store_true",
                    help="time the FP32 step as eager launches instead of replays of its captured graph")
    ap.add_argument("--pmc-step", action="store_true",
                    help="run two FP32 steps and a 512 MiB device copy (the byte-counter "
                         "calibration) for rocprofv3 --pmc, print nothing else")
    ap.add_argument("--pair-only", action="store_true",
                    help="run only the pair-mode workload, one timed round (for profiling it)")
    ap.add_argument("--pmc-lanes", type=int, default=16,
                    help="band launch shape of --pmc-step: lanes per curve")
    ap.add_argument("--pmc-grid", type=int, default=1,
                    help="band launch shape of --pmc-step: grid as a multiple of one fill")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(curves shard by trajectory; see the design notes). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch

    import microsn_amd.api as api
    import microsn_amd.lightcurves as lcm

    if args.pmc_step:
        pmc_step(args.pmc_lanes, args.pmc_grid)
        return
    if args.pair_only:
        print(json.dumps({"extras": {"pair_mode_grid_input": time_pair_mode(api.Context(0), steps=1)}}))
        return
    if args.extras_only:
        ctx = api.Context(0)
        print(json.dumps({"extras": {"batching_64_maps": time_batch_workload(ctx, "batch_64_maps", args.steps),
                                     "scale_10000": time_batch_workload(ctx, "scale_10000",
                                                                                max(3, args.steps // 4)),
                                     "pair_mode_grid_input": time_pair_mode(ctx),
                                     "plain_1000": time_plain_1000(ctx, args.steps, cpu=False),
                                     "photosphere_2048": time_photosphere_2048(ctx, args.steps)}}))
        return
    t_setup = time.perf_counter()
    m = make_map()
    params, prof, times, truth, rng = make_inputs()
    ctx = api.Context(0)
    m_dev = torch.tensor(m, device="cuda")

    # synthetic observed curve: the true trajectory on the FP64 Gaussian maps + noise
    s64 = api.MapSpectrum(ctx, m_dev, "f64")
    mobs, sig = [], []
    for b, sc in enumerate(BAND_SCALE):
        v = s64.convolve("gaussian", R0 * sc).cpu().numpy()
        t = times[b]
        x = truth[0] + truth[3] * np.cos(truth[2]) * t
        y = truth[1] + truth[3] * np.sin(truth[2]) * t
        mobs.append(-2.5 * np.log10(bilinear(v, x, y)) + rng.normal(0.0, 0.05, t.size))
        sig.append(np.full(t.size, 0.05))
    s64.close()
    ncurves = len(prof)

    # ---- FP32 fused path (the metric)
    ws32 = setup(ctx, "f32", m_dev, params, prof, times, mobs, sig)
    t_setup = time.perf_counter() - t_setup
    dt_eager, out32, _ = time_steps(ctx, ws32, args.steps, args.warmup, plain=False, timing=False)
    graph_info = None
    dt32 = dt_eager
    if not args.eager:
        # the same step as one graph replay; its outputs must equal the eager step's
        g32, gout, gen = capture_step(ctx, ws32)
        dt32 = time_graph(ctx, g32, gen, args.steps, args.warmup)
        same = all(torch.equal(a, b) for a, b in zip(gout, out32))
        if not same:
            print("graph replay differs from the eager step: timing the eager step", file=sys.stderr)
            dt32 = dt_eager
        graph_info = {"graph_replay_matches_eager": same, "ms_per_step_eager": dt_eager * 1e3}
    # the same steps again with a HIP event pair around every launch: the
    # per-kernel breakdown and the roofline's launch durations
    ksteps = max(1, min(args.steps, 5))
    dt32_ev, _, kern = time_steps(ctx, ws32, ksteps, 0, plain=False, timing=True)
    chi32, grad32, lo32, npb32, be32 = out32
    band_shape = dict(zip(("lanes", "grid_multiple"),
                          ctx.tune_lookup(ctx.TUNE_KIND_BANDS, "f32", EPOCHS, float(N) * N)))

    # ---- FP64 plain path (same workload)
    plain = None
    err = {}
    if not args.no_plain:
        ws64 = setup(ctx, "f64", m_dev, params, prof, times, mobs, sig)
        k64 = max(2, args.steps // 5)
        dt64, out64, _ = time_steps(ctx, ws64, k64, 1, plain=True, timing=False)
        # one more step with an event pair around every launch: the plain
        # path's per-kernel breakdown (its own roofline beside the FP32 one)
        _, _, kern64 = time_steps(ctx, ws64, 1, 0, plain=True, timing=True)
        chi64, grad64, lo64, npb64, be64 = out64
        lo32, npb32, lo64, npb64 = (x.cpu().numpy() for x in (lo32, npb32, lo64, npb64))
        fin = torch.isfinite(chi64)
        rel = ((chi32 - chi64).abs() / chi64.abs())[fin]
        gs = grad64.abs().amax(dim=0)
        n32 = ws32["nodes"].nodes[..., 0].double()
        n64 = ws64["nodes"].nodes[..., 0]
        map_err = float(((n32 - n64).abs() / n64.abs()).max())
        err = {"chi2_max_rel_err": float(rel.max()), "chi2_median_rel_err": float(rel.median()),
               "grad_max_err_over_scale": float(((grad32 - grad64).abs() / gs).max()),
               "convolved_map_max_rel_err": map_err,
               "lightcurve_max_err_mag": 2.5 / np.log(10) * map_err,
               "bins_identical_to_fp64": bool(np.array_equal(lo32, lo64) and np.array_equal(npb32, npb64)),
               "curves_changing_bin": int(np.abs(npb32.astype(np.int64) - npb64.astype(np.int64)).sum() // 2)}
        plain = {"value": ncurves / dt64, "ms_per_step": dt64 * 1e3, "steps": k64, "dtype": "f64",
                 "kernels_ms_per_step": {k: v[0] for k, v in sorted(kern64.items())},
                 "roofline_by_kernel": {
                     k: {"achieved": algorithmic_bytes(k, 8, ncurves) / (v[0] / v[1] / 1e3) / 1e9,
                         "unit": "GB/s",
                         "frac": algorithmic_bytes(k, 8, ncurves) / (v[0] / v[1] / 1e3) / HBM_PEAK,
                         "avg_launch_ms": v[0] / v[1]}
                     for k, v in kern64.items() if algorithmic_bytes(k, 8, ncurves)}}
        del ws64
        torch.cuda.empty_cache()

    # ---- roofline of the dominant kernel
    def roofline(k):
        ms_tot, launches = kern[k]
        avg_s = ms_tot / launches / 1e3
        ab = algorithmic_bytes(k, 4, ncurves)
        tr = measured_traffic(k)
        r = {"kernel": k, "bound": "hbm", "achieved": ab / avg_s / 1e9, "peak": HBM_PEAK / 1e9,
             "unit": "GB/s", "frac": ab / avg_s / HBM_PEAK, "traffic": tr,
             # the same launch time against the HBM bytes the counters measured
             "frac_of_measured_traffic": tr / avg_s / HBM_PEAK if tr else None,
             "algorithmic_bytes_per_launch": ab, "avg_launch_ms": avg_s * 1e3,
             "ms_per_step": ms_tot / max(1, ksteps)}
        if k == "lc_chi2_grad":  # its gathers, served by the caches
            r["tap_bytes_per_launch"] = tap_bytes(4, ncurves)
            r["tap_rate_GBps"] = tap_bytes(4, ncurves) / avg_s / 1e9
        if k in LIMITER:
            r["limiter"] = LIMITER[k]
        return r

    # the dominant kernel by time per step, and every kernel with algorithmic bytes
    timed = [k for k in kern if algorithmic_bytes(k, 4, ncurves)]
    dom = max(timed, key=lambda k: kern[k][0])
    roof = roofline(dom)
    roof_all = {k: roofline(k) for k in sorted(timed, key=lambda k: -kern[k][0])}

    extras = None
    if not args.no_extras:
        del ws32
        torch.cuda.empty_cache()
        extras = {"batching_64_maps": time_batch_workload(ctx, "batch_64_maps", args.steps),
                  "scale_10000": time_batch_workload(ctx, "scale_10000", max(3, args.steps // 4)),
                  "pair_mode_grid_input": time_pair_mode(ctx),
                  "plain_1000": time_plain_1000(ctx, args.steps, cpu=not args.no_cpu),
                  "photosphere_2048": time_photosphere_2048(ctx, args.steps)}

    cpu = None
    if not args.no_cpu:
        cpu, chi_host = cpu_baseline(m, params, prof, times, mobs, sig, api.mpd_bins(NBINS, 1.0, 1.0e7, True))
        fin = np.isfinite(chi_host)
        err["chi2_max_rel_err_vs_host_oracle"] = float(
            np.max(np.abs(chi32.cpu().numpy()[fin] - chi_host[fin]) / np.abs(chi_host[fin])))

    line = {
        "metric": "light curves/s with chi2+grad @4096² map, FP32 vs FP64 plain, ≤1e-5 chi2 err",
        "value": ncurves / dt32, "unit": "light curves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt32 * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded log-normal map, random trajectories)",
        "config": {"workload": "4096^2 map, 3 profiles x 3 bands (2/4/8 px), 1e5 trajectories x 3 profiles, "
                               "100 epochs/band, chi2 + 7-param adjoint + sort + 100-bin table",
                   "map": N, "curves": ncurves, "epochs_per_curve": EPOCHS * len(BAND_SCALE),
                   "nodes": len(SHAPES) * len(BAND_SCALE)},
        "fp64_plain": plain, "speedup_vs_fp64_plain": (ncurves / dt32) / plain["value"] if plain else None,
        "precision": err, "roofline": roof, "roofline_by_kernel": roof_all, "cpu_baseline": cpu,
        "kernels_ms_per_step": {k: v[0] / ksteps for k, v in sorted(kern.items())},
        "ms_per_step_with_kernel_events": dt32_ev * 1e3,
        "hip_graph": graph_info,
        "band_launch_shape": band_shape,
        "setup_s": t_setup,
        "extras": extras,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
