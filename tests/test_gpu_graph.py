"""One whole step of the light-curve path (map update, node maps, chi2 +
adjoint, sort, bin table) captured into a HIP graph through torch.cuda.graph
and replayed: every replay, with eager work and synchronisations between the
replays, gives the eager step's outputs bit for bit."""
import numpy as np
import pytest
import torch

from microsn_amd import api
from microsn_amd import lightcurves as lcm

pytestmark = pytest.mark.gpu


def _workload(ctx, N=512, n=4096, seed=77, prec="f32"):
    rng = np.random.default_rng(seed)
    g = torch.randn((N, N), dtype=torch.float64, device="cuda",
                    generator=torch.Generator(device="cuda").manual_seed(seed))
    m = torch.exp(0.5 * g) * 5.0
    # FP64 maps keep the mean in the spectrum (subtract_mean off by default):
    # their update must issue kernels only, like the FP32 path
    spec = api.MapSpectrum(ctx, m, prec)
    shapes, bs = ("gaussian", "uniform"), (1.0, 2.0)
    nodes = lcm.NodeSet(ctx, spec, shapes, [[[2.0], [4.0]], [[2.0], [4.0]]], bs, compute=False)
    times = [np.sort(rng.uniform(0, 50, 30)) for _ in bs]
    obs = lcm.Observations(ctx, len(bs), [[(t, rng.normal(-1.7, 0.2, t.size), np.full(t.size, 0.05))
                                           for t in times]])
    x0 = rng.uniform(120, N - 121, n)
    y0 = rng.uniform(120, N - 121, n)
    th = rng.uniform(0, 2 * np.pi, n)
    prm = np.stack([x0, y0, th, np.full(n, 2.0), np.full(n, 2.0), np.zeros(n), rng.normal(0, .05, n)], 1)
    return dict(m=m, spec=spec, nodes=nodes, obs=obs, P=torch.tensor(prm, device="cuda"),
                prof=torch.tensor(np.arange(n) % 2, dtype=torch.int32, device="cuda"),
                oi=torch.zeros(n, dtype=torch.int32, device="cuda"),
                bins=api.mpd_bins(50, 1.0, 1e7))


def _step(ctx, w, convolve=False):
    w["spec"].update(w["m"])
    if not convolve:
        w["nodes"].compute()
    # convolve=True: node maps and chi2 in one call
    c, g = lcm.chi2_grad(ctx, w["nodes"], w["obs"], w["P"], w["prof"], w["oi"], convolve=convolve)
    perm, _, sk = api.sort_f64(ctx, c, sorted_keys=True)
    lo, npb, bc, be = api.bin_table_device(ctx, sk, w["bins"])
    return c, g, perm, sk, lo, npb, bc, be, w["nodes"].nodes


# Under capture the combined call's node kernels and the band launches of
# the two profiles become parallel branches on the context's spread streams;
# eager calls keep them in order.  1000^2: the node maps come from the fused
# mixed-radix kernel, so they are branches as well.
@pytest.mark.parametrize("convolve,prec,N", [(False, "f32", 512), (True, "f32", 512), (True, "f64", 512),
                                             (True, "f32", 1000)])
def test_captured_step_replays_match_eager(ctx, convolve, prec, N):
    w = _workload(ctx, N=N, prec=prec)
    _step_c = lambda c, w: _step(c, w, convolve)
    eager = [t.clone() for t in _step_c(ctx, w)]
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    ctx.set_stream(s)
    try:
        with torch.cuda.stream(s):
            _step_c(ctx, w)  # the workspace, tables and caches settle on this stream
        s.synchronize()
        graph = torch.cuda.CUDAGraph()
        gen = ctx.generation()
        with torch.cuda.graph(graph, stream=s):
            out = _step_c(ctx, w)
        # the capture itself reallocated nothing and rewrote no table
        assert ctx.generation() == gen
    finally:
        ctx.set_stream(torch.cuda.current_stream())
    for r in range(3):
        assert ctx.generation() == gen, "a call since the capture invalidated the graph"
        for t in out:
            t.zero_()
        graph.replay()
        torch.cuda.synchronize()
        for a, b in zip(out, eager):
            assert torch.equal(a, b), r
        # eager work between replays, on the library's usual stream
        again = _step_c(ctx, w)
        torch.cuda.synchronize()
        assert torch.equal(again[0], eager[0])
