"""The batched light-curve call (mlc_lc_chi2_grad_batch: several maps of
different sizes, all bands in one persistent launch or one launch per band)
and the launch-shape
tuner, through the C ABI:
  * one map: the batched call equals the per-band launches of
    mlc_lc_chi2_grad bit for bit at every launch shape (8, 16, 32 lanes per
    curve), FP32 and FP64;
  * five maps of different sizes, uneven epochs per observation set, single-
    and multi-node profiles, curves with invalid map or profile indices:
    every curve equals the per-map call on its own map bit for bit;
  * a captured batched call replays bit-equal to the eager call;
  * every batched case also as one launch per band (the tuner's other
    launch form), bit-equal to the persistent launch;
  * the tuner keeps one of its candidates per key, the result equals a
    forced run of that shape, and kept shapes (with the batched launch
    form) survive a save/load;
  * the multi-node band kernel forced onto single-node bands (radius below
    the node, and a NaN radius) agrees with the single-node kernel."""
import numpy as np
import pytest
import torch

from microsn_amd import api
from microsn_amd import lightcurves as lcm

pytestmark = pytest.mark.gpu


@pytest.fixture
def tuned(ctx):
    yield ctx
    ctx.tune_mode(ctx.TUNE_AUTO)


def _map(ctx, N, seed):
    g = torch.randn((N, N), dtype=torch.float64, device="cuda",
                    generator=torch.Generator(device="cuda").manual_seed(seed))
    return torch.exp(0.5 * g) * 5.0


def _obs(ctx, rng, nsets, B, emax=90):
    sets = []
    for _ in range(nsets):
        bands = []
        for _b in range(B):
            e = int(rng.integers(1, emax))
            t = np.sort(rng.uniform(0, 40, e))
            bands.append((t, rng.normal(-1.7, 0.2, e), rng.uniform(0.03, 0.08, e)))
        sets.append(bands)
    return lcm.Observations(ctx, B, sets)


def _curves(rng, n, sizes, map_idx, nsets, P):
    N = np.array([sizes[m] if 0 <= m < len(sizes) else 256 for m in map_idx], float)
    prm = np.stack([rng.uniform(50, N - 51), rng.uniform(50, N - 51), rng.uniform(0, 2 * np.pi, n),
                    rng.uniform(0.1, 1.0, n), rng.uniform(1.0, 3.0, n), rng.uniform(0.0, 0.05, n),
                    rng.normal(0, 0.1, n)], 1)
    prm[::89, 0] = N[::89] + 3.0  # off the map: chi2 = inf
    prof = rng.integers(0, P, n).astype(np.int32)
    prof[::211] = P  # invalid profile
    return prm, prof, rng.integers(0, nsets, n).astype(np.int32)


def _nodes(ctx, m, prec):
    spec = api.MapSpectrum(ctx, m, prec)
    radii = [[[2.0], [3.0]], [[1.5, 2.5, 4.0], [2.0, 3.5, 6.0]]]
    return spec, lcm.NodeSet(ctx, spec, ["gaussian", "exponential"], radii, [1.0, 1.5])


def _t(x):
    return torch.tensor(x, device="cuda")


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("lanes", [8, 16, 32])
@pytest.mark.parametrize("per_band", [False, True])
def test_batch_one_map_equals_band_launches(tuned, prec, lanes, per_band):
    ctx = tuned
    rng = np.random.default_rng(100 + lanes)
    spec, nodes = _nodes(ctx, _map(ctx, 512, 5), prec)
    obs = _obs(ctx, rng, 9, 2)
    n = 20000
    prm, prof, oi = _curves(rng, n, [512], np.zeros(n, int), 9, 2)
    ctx.tune_force(lanes, 1, ctx.TUNE_BATCH_PER_BAND if per_band else 0)
    P, pr, o = _t(prm), _t(prof), _t(oi)
    ref = lcm.chi2_grad(ctx, nodes, obs, P, pr, o)
    got = lcm.chi2_grad_batch(ctx, lcm.NodeBatch([nodes]), obs, P, torch.zeros(n, dtype=torch.int32,
                                                                               device="cuda"), pr, o)
    torch.cuda.synchronize()
    assert torch.equal(got[0], ref[0]) and torch.equal(got[1], ref[1])
    assert torch.isinf(ref[0]).any() and torch.isfinite(ref[0]).any()
    spec.close()


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("per_band", [False, True])
def test_batch_many_maps_equals_per_map_calls(tuned, prec, per_band):
    ctx = tuned
    rng = np.random.default_rng(7)
    sizes = [256, 384, 512, 320, 640]
    specs, sets = [], []
    for i, N in enumerate(sizes):
        s, ns = _nodes(ctx, _map(ctx, N, 50 + i), prec)
        specs.append(s)
        sets.append(ns)
    obs = _obs(ctx, rng, 13, 2, emax=200)
    n = 30000
    midx = rng.integers(0, len(sizes), n).astype(np.int32)
    midx[::97] = -1
    midx[5::97] = len(sizes)
    prm, prof, oi = _curves(rng, n, sizes, midx, 13, 2)
    ctx.tune_force(16, 1, ctx.TUNE_BATCH_PER_BAND if per_band else 0)
    c, g = lcm.chi2_grad_batch(ctx, lcm.NodeBatch(sets), obs, _t(prm), _t(midx), _t(prof), _t(oi))
    torch.cuda.synchronize()
    bad = (midx < 0) | (midx >= len(sizes))
    assert torch.isinf(c[_t(bad)]).all() and (g[_t(bad)] == 0).all()
    for m in range(len(sizes)):
        sel = np.nonzero(midx == m)[0]
        rc, rg = lcm.chi2_grad(ctx, sets[m], obs, _t(prm[sel]), _t(prof[sel]), _t(oi[sel]))
        torch.cuda.synchronize()
        assert torch.equal(c[_t(sel)], rc), m
        assert torch.equal(g[_t(sel)], rg), m
    for s in specs:
        s.close()


@pytest.mark.parametrize("mode", ["auto", "per_band"])
def test_batch_graph_replay_equals_eager(tuned, mode):
    ctx = tuned
    if mode == "per_band":
        ctx.tune_force(16, 1, ctx.TUNE_BATCH_PER_BAND)
    rng = np.random.default_rng(8)
    sizes = [256, 512]
    specs, sets = zip(*[_nodes(ctx, _map(ctx, N, 70 + i), "f32") for i, N in enumerate(sizes)])
    batch = lcm.NodeBatch(sets)
    obs = _obs(ctx, rng, 5, 2)
    n = 8000
    midx = rng.integers(0, 2, n).astype(np.int32)
    prm, prof, oi = _curves(rng, n, sizes, midx, 5, 2)
    P, M, pr, o = _t(prm), _t(midx), _t(prof), _t(oi)
    eager = [x.clone() for x in lcm.chi2_grad_batch(ctx, batch, obs, P, M, pr, o)]  # tunes its key
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    ctx.set_stream(s)
    try:
        with torch.cuda.stream(s):
            lcm.chi2_grad_batch(ctx, batch, obs, P, M, pr, o)
        s.synchronize()
        gen = ctx.generation()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=s):
            out = lcm.chi2_grad_batch(ctx, batch, obs, P, M, pr, o)
    finally:
        ctx.set_stream(torch.cuda.current_stream())
    for r in range(3):
        assert ctx.generation() == gen
        for x in out:
            x.zero_()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(out[0], eager[0]) and torch.equal(out[1], eager[1]), r
        again = lcm.chi2_grad_batch(ctx, batch, obs, P, M, pr, o)
        torch.cuda.synchronize()
        assert torch.equal(again[0], eager[0])
    for sp in specs:
        sp.close()


def test_tuner_keeps_a_candidate_and_reuses_it(tuned, tmp_path):
    ctx = tuned
    rng = np.random.default_rng(9)
    spec, nodes = _nodes(ctx, _map(ctx, 128, 9), "f32")
    # 3 epochs per band on a 128^2 map: a key no other test uses
    sets = [[(np.sort(rng.uniform(0, 10, 3)), rng.normal(-1.7, .2, 3), np.full(3, .05)) for _ in range(2)]]
    obs = lcm.Observations(ctx, 2, sets)
    n = 4000
    prm, prof, oi = _curves(rng, n, [128], np.zeros(n, int), 1, 2)
    prm[:, 0] = rng.uniform(40, 87, n)
    prm[:, 1] = rng.uniform(40, 87, n)
    prm[:, 3] = 0.5
    ctx.tune_mode(ctx.TUNE_AUTO)
    P, pr, o = _t(prm), _t(prof), _t(oi)
    assert ctx.tune_lookup(ctx.TUNE_KIND_BANDS, "f32", 3, 128 * 128) == (0, 0)
    auto = lcm.chi2_grad(ctx, nodes, obs, P, pr, o)
    lanes, gm = ctx.tune_lookup(ctx.TUNE_KIND_BANDS, "f32", 3, 128 * 128)
    assert lanes in (8, 16, 32) and gm in (1, 2)
    again = lcm.chi2_grad(ctx, nodes, obs, P, pr, o)
    ctx.tune_force(lanes, gm)
    forced = lcm.chi2_grad(ctx, nodes, obs, P, pr, o)
    torch.cuda.synchronize()
    assert torch.equal(auto[0], again[0]) and torch.equal(auto[0], forced[0])
    assert torch.equal(auto[1], forced[1])
    path = tmp_path / "shapes.txt"
    ctx.tune_save(path)
    other = api.Context(0)
    try:
        other.tune_load(path)
        assert other.tune_lookup(other.TUNE_KIND_BANDS, "f32", 3, 128 * 128) == (lanes, gm)
    finally:
        other.close()
    spec.close()


def test_tuner_batch_kind_keeps_the_launch_form(tuned, tmp_path):
    """A two-band batched call under the tuner keeps lanes, grid multiple and
    the launch form (persistent or one launch per band) for its key; the
    result equals a forced run of the kept shape and the form survives a
    save/load."""
    ctx = tuned
    rng = np.random.default_rng(12)
    sizes = [192, 160]
    specs, sets = zip(*[_nodes(ctx, _map(ctx, N, 90 + i), "f64") for i, N in enumerate(sizes)])
    batch = lcm.NodeBatch(sets)
    # 7 epochs per band, largest map 192^2: a key no other test uses
    sets_obs = [[(np.sort(rng.uniform(0, 10, 7)), rng.normal(-1.7, .2, 7), np.full(7, .05)) for _ in range(2)]]
    obs = lcm.Observations(ctx, 2, sets_obs)
    n = 5000
    midx = rng.integers(0, 2, n).astype(np.int32)
    prm, prof, oi = _curves(rng, n, sizes, midx, 1, 2)
    args = (_t(prm), _t(midx), _t(prof), _t(oi))
    ctx.tune_mode(ctx.TUNE_AUTO)
    key = (ctx.TUNE_KIND_BATCH, "f64", 7, 192 * 192)
    assert ctx.tune_lookup(*key, with_flags=True) == (0, 0, 0)
    auto = lcm.chi2_grad_batch(ctx, batch, obs, *args)
    lanes, gm, fl = ctx.tune_lookup(*key, with_flags=True)
    assert lanes in (8, 16, 32) and gm in (1, 2) and fl in (0, ctx.TUNE_BATCH_PER_BAND)
    ctx.tune_force(lanes, gm, fl)
    forced = lcm.chi2_grad_batch(ctx, batch, obs, *args)
    ctx.tune_force(lanes, gm, fl ^ ctx.TUNE_BATCH_PER_BAND)
    other_form = lcm.chi2_grad_batch(ctx, batch, obs, *args)
    torch.cuda.synchronize()
    for x, y in ((auto, forced), (auto, other_form)):
        assert torch.equal(x[0], y[0]) and torch.equal(x[1], y[1])
    path = tmp_path / "shapes.txt"
    ctx.tune_save(path)
    other = api.Context(0)
    try:
        other.tune_load(path)
        assert other.tune_lookup(*key, with_flags=True) == (lanes, gm, fl)
    finally:
        other.close()
    for sp in specs:
        sp.close()


def test_multi_node_kernel_on_single_node_bands(tuned):
    """Single-node bands through the multi-node kernel (forced), with radii
    below the node and one NaN radius: the same chi2 as the single-node
    kernel to FP32 rounding, the NaN curve flagged by both."""
    ctx = tuned
    rng = np.random.default_rng(10)
    spec = api.MapSpectrum(ctx, _map(ctx, 256, 10), "f32")
    nodes = lcm.NodeSet(ctx, spec, ["gaussian"], [[[3.0], [4.0]]], [1.0, 1.3])
    obs = _obs(ctx, rng, 3, 2)
    n = 3000
    prm, prof, oi = _curves(rng, n, [256], np.zeros(n, int), 3, 1)
    prof[:] = 0
    prm[:, 4] = rng.uniform(0.5, 2.0, n)  # r below the node radius of both bands
    prm[7, 4] = np.nan
    P, pr, o = _t(prm), _t(prof), _t(oi)
    ctx.tune_force(16, 1)
    one = lcm.chi2_grad(ctx, nodes, obs, P, pr, o)
    ctx.tune_force(16, 1, ctx.TUNE_GENERAL_NODES)
    gen = lcm.chi2_grad(ctx, nodes, obs, P, pr, o)
    torch.cuda.synchronize()
    c1, c2 = one[0].cpu().numpy(), gen[0].cpu().numpy()
    assert np.isinf(c1[7]) and np.isinf(c2[7])
    fin = np.isfinite(c1)
    assert np.array_equal(fin, np.isfinite(c2)) and fin.sum() > n // 2
    assert np.max(np.abs(c1[fin] - c2[fin]) / np.abs(c1[fin])) < 1e-6
    spec.close()
