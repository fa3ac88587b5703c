"""The batched map calls (mlc_maps_update, mlc_maps_convolve_nodes) through
the C ABI: maps of different sizes (powers of two, widths of the compile-time
mixed-radix kernels and of the runtime-plan row kernels), with and without
mean subtraction, one map without nodes, FP32 and FP64:
  * the node maps equal those of the per-map calls (mlc_map_update then
    mlc_map_convolve_nodes on each map) bit for bit;
  * a captured batched update + convolution replays bit-equal to the eager
    calls, with eager work between the replays;
  * a map listed twice in one batched call gives the per-map calls' bits;
  * maps of different precisions or contexts, and a map of the wrong shape,
    are refused."""
import numpy as np
import pytest
import torch

from microsn_amd import api
from microsn_amd import lightcurves as lcm
from microsn_amd._lib import MlcError

pytestmark = pytest.mark.gpu

SIZES = [(256, 256), (320, 192), (1000, 1000), (512, 384), (1200, 600), (128, 128)]
SHAPES = ["gaussian", "uniform", "exponential"]


def _maps(seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    return [torch.exp(0.5 * torch.randn((ny, nx), dtype=torch.float64, device="cuda", generator=g)) * 5.0
            for nx, ny in SIZES]


def _sets(ctx, maps, prec, with_mean):
    specs, sets = [], []
    for i, m in enumerate(maps):
        sp = api.MapSpectrum(ctx, m, prec, subtract_mean=with_mean[i])
        # the last map has no nodes
        radii = [[[2.0 + i, 5.5]], [[3.0]], [[1.5 + 0.5 * i]]] if i < len(maps) - 1 else [[[]], [[]], [[]]]
        ns = lcm.NodeSet(ctx, sp, SHAPES, radii, [1.0], compute=False)
        specs.append(sp)
        sets.append(ns)
    return specs, sets


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_batched_map_calls_equal_per_map_calls(ctx, prec):
    maps0, maps1 = _maps(11), _maps(12)
    with_mean = [True, False, True, True, False, True] if prec == "f32" else [False, True, False, False, True, False]
    spa, seta = _sets(ctx, maps0, prec, with_mean)
    spb, setb = _sets(ctx, maps0, prec, with_mean)
    print("batched map calls", prec, flush=True)
    # a new map into every spectrum, then the node maps
    for sp, m, ns in zip(spa, maps1, seta):
        sp.update(m)
        ns.compute()
    api.update_spectra(spb, maps1)
    lcm.NodeBatch(setb).compute()
    torch.cuda.synchronize()
    for i, (a, b) in enumerate(zip(seta, setb)):
        assert a.nodes.shape[0] == b.nodes.shape[0]
        assert torch.equal(a.nodes, b.nodes), i
    assert seta[0].nodes.shape[0] == 4 and seta[-1].nodes.shape[0] == 0
    # mean subtraction really changes the spectrum path: the value maps of
    # the two kinds of map agree with the oracle-free identity mean(node) ~
    # mean(map) (the kernels sum to one)
    for i, ns in enumerate(setb[:-1]):
        mu = ns.rowmajor()[0, :, :, 0].double().mean()
        assert abs(float(mu) / float(maps1[i].mean()) - 1) < 1e-5, i
    for sp in spa + spb:
        sp.close()


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_batched_map_calls_with_a_map_listed_twice(ctx, prec):
    """One spectrum behind two node sets (the y-filter path at 1200 x 600 and
    the fused path at 256^2), and a spectrum updated twice in one call: the
    same bits as the per-map calls (a map listed twice is not spread over
    streams)."""
    maps0, maps1 = _maps(31), _maps(32)
    for idx in (4, 0):
        sp = api.MapSpectrum(ctx, maps0[idx], prec)
        a = lcm.NodeSet(ctx, sp, ["gaussian", "uniform"], [[[2.0, 4.5]], [[3.0]]], [1.0], compute=False)
        b = lcm.NodeSet(ctx, sp, ["exponential", "gaussian"], [[[1.5, 6.0]], [[2.5]]], [1.0], compute=False)
        print("map listed twice", prec, SIZES[idx], flush=True)
        api.update_spectra([sp, sp], [maps1[idx], maps1[idx]])
        lcm.NodeBatch([a, b]).compute()
        torch.cuda.synchronize()
        ra, rb = a.nodes.clone(), b.nodes.clone()
        sp.update(maps1[idx])
        a.compute()
        b.compute()
        torch.cuda.synchronize()
        assert torch.equal(ra, a.nodes) and torch.equal(rb, b.nodes), idx
        sp.close()


def test_batched_map_calls_captured(ctx):
    maps0, maps1 = _maps(21), _maps(22)
    specs, sets = _sets(ctx, maps0, "f32", [True] * len(SIZES))
    batch = lcm.NodeBatch(sets)

    def step():
        api.update_spectra(specs, maps1)
        batch.compute()

    step()
    torch.cuda.synchronize()
    eager = [ns.nodes.clone() for ns in sets]
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    ctx.set_stream(s)
    try:
        with torch.cuda.stream(s):
            step()
        s.synchronize()
        gen = ctx.generation()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=s):
            step()
    finally:
        ctx.set_stream(torch.cuda.current_stream())
    for r in range(3):
        assert ctx.generation() == gen
        for ns in sets:
            ns.nodes.zero_()
        graph.replay()
        torch.cuda.synchronize()
        for ns, e in zip(sets, eager):
            assert torch.equal(ns.nodes, e), r
        # eager work between replays: a per-map update of another map
        specs[0].update(maps0[0])
        specs[0].update(maps1[0])
        torch.cuda.synchronize()
    for sp in specs:
        sp.close()


def test_batched_map_calls_refuse_mixed_inputs(ctx):
    m = _maps(31)
    a = api.MapSpectrum(ctx, m[0], "f32")
    b = api.MapSpectrum(ctx, m[1], "f64")
    with pytest.raises(MlcError, match="different precisions"):
        api.update_spectra([a, b], [m[0], m[1]])
    with pytest.raises(ValueError, match="shape"):
        api.update_spectra([a], [m[1]])
    other = api.Context(0)
    try:
        c = api.MapSpectrum(other, m[0], "f32")
        with pytest.raises(ValueError, match="different contexts"):
            api.update_spectra([a, c], [m[0], m[0]])
        c.close()
    finally:
        other.close()
    for sp in (a, b):
        sp.close()
