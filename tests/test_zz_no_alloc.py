"""mlc_lc_chi2_grad_noalloc (scratch in a static device array, no
allocation) against mlc_lc_chi2_grad (the same kernels on the context's
workspace), through the C ABI: bit-identical chi2 and gradient on two calls
in a row, the second smaller than the first (the static scratch keeps the
first call's contents), and a call too large for the static scratch refused
before any launch."""
import numpy as np
import pytest
import torch

from microsn_amd import api
from microsn_amd import lightcurves as lcm
from microsn_amd._lib import MlcError, lib, ptr

pytestmark = pytest.mark.gpu


def _call(ctx, fn, nodes, obs, P, prof, oi):
    n = P.shape[0]
    chi2 = torch.full((n,), -1.0, dtype=torch.float64, device="cuda")
    grad = torch.full((n, 7), -1.0, dtype=torch.float64, device="cuda")
    ctx.check(fn(ctx.h, nodes.prec, ptr(nodes.nodes), nodes.nodes.shape[0], nodes.Nx, nodes.Ny, nodes.P,
                 nodes.B, ptr(nodes.node_off), ptr(nodes.node_cnt), ptr(nodes.node_rad),
                 ptr(nodes.band_scale), n, ptr(P), ptr(prof), ptr(oi), obs.nobs, ptr(obs.off), obs.nepoch,
                 ptr(obs.t), ptr(obs.mobs), ptr(obs.sig), ptr(chi2), ptr(grad)))
    return chi2, grad


def _workload(ctx, prec, n, nsets, seed):
    rng = np.random.default_rng(seed)
    N = 384
    g = torch.randn((N, N), dtype=torch.float64, device="cuda",
                    generator=torch.Generator(device="cuda").manual_seed(seed))
    m = torch.exp(0.5 * g) * 5.0
    spec = api.MapSpectrum(ctx, m, prec)
    # profile 0: one node per band (single-node kernel); profile 1: three
    # radius nodes per band (Hermite kernel)
    radii = [[[2.0], [3.0]], [[1.5, 2.5, 4.0], [2.0, 3.5, 6.0]]]
    nodes = lcm.NodeSet(ctx, spec, ["gaussian", "exponential"], radii, [1.0, 1.5])
    sets = []
    for _ in range(nsets):
        bands = []
        for _b in range(2):
            e = int(rng.integers(5, 60))
            t = np.sort(rng.uniform(0, 40, e))
            bands.append((t, rng.normal(-1.7, 0.2, e), rng.uniform(0.03, 0.08, e)))
        sets.append(bands)
    obs = lcm.Observations(ctx, 2, sets)
    prm = np.stack([rng.uniform(60, N - 61, n), rng.uniform(60, N - 61, n), rng.uniform(0, 2 * np.pi, n),
                    rng.uniform(0.2, 1.2, n), rng.uniform(1.0, 3.0, n), rng.uniform(0.0, 0.05, n),
                    rng.normal(0, 0.1, n)], 1)
    prm[::97, 0] = N + 5.0  # some curves start off the map: chi2 = inf, zero gradient
    prof = rng.integers(0, 2, n).astype(np.int32)
    oi = rng.integers(0, nsets, n).astype(np.int32)
    return (spec, nodes, obs, torch.tensor(prm, device="cuda"), torch.tensor(prof, device="cuda"),
            torch.tensor(oi, device="cuda"))


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_noalloc_equals_workspace_entry_point(ctx, prec):
    L = lib()
    for call, (n, nsets, seed) in enumerate([(40000, 7, 11), (3000, 3, 12)]):
        print(f"noalloc case prec={prec} call={call} n={n} sets={nsets} seed={seed}", flush=True)
        spec, nodes, obs, P, prof, oi = _workload(ctx, prec, n, nsets, seed)
        ref = _call(ctx, L.mlc_lc_chi2_grad, nodes, obs, P, prof, oi)
        got = _call(ctx, L.mlc_lc_chi2_grad_noalloc, nodes, obs, P, prof, oi)
        torch.cuda.synchronize()
        assert torch.equal(got[0], ref[0]), call
        assert torch.equal(got[1], ref[1]), call
        assert torch.isinf(ref[0]).any() and torch.isfinite(ref[0]).any()
        spec.close()


def test_noalloc_refuses_calls_beyond_its_scratch(ctx):
    """600 000 curves need about 82 MB of scratch, more than the 64 MiB static
    array: refused with an argument error before any launch."""
    spec, nodes, obs, _, _, _ = _workload(ctx, "f32", 16, 1, 13)
    n = 600_000
    P = torch.zeros((n, 7), dtype=torch.float64, device="cuda")
    prof = torch.zeros(n, dtype=torch.int32, device="cuda")
    oi = torch.zeros(n, dtype=torch.int32, device="cuda")
    with pytest.raises(MlcError) as e:
        _call(ctx, lib().mlc_lc_chi2_grad_noalloc, nodes, obs, P, prof, oi)
    assert e.value.code == 1 and "static scratch" in str(e.value)
    torch.cuda.synchronize()
    spec.close()
