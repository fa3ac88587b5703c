"""GPU parity of the trajectory light curves, per-curve magnitude chi2 and
adjoint (mlc_map_convolve node maps + mlc_lc_chi2_grad) against the FP64
oracle, central finite differences and the FP64 path."""
import numpy as np
import pytest
import torch

from microsn_amd import api
from microsn_amd import lightcurves as lcm

pytestmark = pytest.mark.gpu


def _map(rng, Ny, Nx):
    g = rng.standard_normal((Ny, Nx))
    return 5.0 * np.exp(0.5 * g) / np.exp(0.125)


def _setup(ctx, prec, rng, Ny=128, Nx=160, radii=((2.0, 3.0, 4.5), (2.5,)), shapes=("gaussian", "exponential"),
           band_scale=(1.0, 1.5), smooth=True):
    m = _map(rng, Ny, Nx)
    if smooth:  # smoothed field, as the synthetic configurations use
        s0 = api.MapSpectrum(ctx, torch.tensor(m, device="cuda"), "f64")
        m = s0.convolve("gaussian", 3.0 * 1.18).cpu().numpy()
    spec = api.MapSpectrum(ctx, torch.tensor(m, device="cuda"), prec)
    P, B = len(shapes), len(band_scale)
    rr = [[list(radii[(p + b) % len(radii)]) for b in range(B)] for p in range(P)]
    nodes = lcm.NodeSet(ctx, spec, shapes, rr, band_scale)
    sets = []
    for o in range(3):
        bands = []
        for b in range(B):
            ne = int(rng.integers(1 if o else 0, 40))
            t = np.sort(rng.uniform(0, 30, ne))
            bands.append((t, rng.normal(-1.7, 0.3, ne), rng.uniform(0.03, 0.08, ne)))
        sets.append(bands)
    obs = lcm.Observations(ctx, B, sets)
    return m, spec, nodes, obs, sets


def _params(rng, n, Ny, Nx, P, nobs, tmax=30.0):
    L = 0.9 * tmax * 1.5
    x0 = rng.uniform(L, Nx - 1 - L, n)
    y0 = rng.uniform(L, Ny - 1 - L, n) if Ny - 1 - 2 * L > 0 else np.full(n, (Ny - 1) / 2)
    th = rng.uniform(0, 2 * np.pi, n)
    v = rng.uniform(0.2, 1.0, n)
    r0 = rng.uniform(2.0, 3.5, n)
    w = rng.uniform(0.0, 0.04, n)
    dm = rng.normal(0, 0.1, n)
    prm = np.stack([x0, y0, th, v, r0, w, dm], 1)
    return prm, rng.integers(0, P, n).astype(np.int32), rng.integers(0, nobs, n).astype(np.int32)


def _oracle(orc, nodes, sets, prm, prof, obsi):
    maps = nodes.rowmajor().double().cpu().numpy()
    off = [0]
    t, mo, sg = [], [], []
    for bands in sets:
        for (tb, mb, sb) in bands:
            t.append(tb); mo.append(mb); sg.append(sb)
            off.append(off[-1] + len(tb))
    cat = lambda xs: np.concatenate(xs) if xs else np.zeros(0)
    return orc.lc_chi2_grad(maps, nodes.B, nodes.node_off, nodes.node_cnt, nodes.node_rad, None,
                            nodes.band_scale, prm, prof, obsi, np.array(off, np.int32), cat(t), cat(mo),
                            cat(sg))


@pytest.mark.parametrize("plain", [False, True])
def test_f64_matches_oracle(ctx, orc, plain):
    rng = np.random.default_rng(10)
    m, spec, nodes, obs, sets = _setup(ctx, "f64", rng)
    prm, prof, obsi = _params(rng, 300, 128, 160, nodes.P, obs.nobs)
    prm[0, 0] = -5.0  # leaves the map: +inf, zero gradient
    c, g = lcm.chi2_grad(ctx, nodes, obs, torch.tensor(prm, device="cuda"), torch.tensor(prof, device="cuda"),
                         torch.tensor(obsi, device="cuda"), plain=plain)
    rc, rg = _oracle(orc, nodes, sets, prm, prof, obsi)
    c, g = c.cpu().numpy(), g.cpu().numpy()
    assert np.isinf(c[0]) and np.isinf(rc[0]) and not g[0].any()
    fin = np.isfinite(rc)
    assert np.array_equal(np.isfinite(c), fin)
    assert np.max(np.abs(c[fin] - rc[fin]) / np.maximum(np.abs(rc[fin]), 1e-30)) < 1e-11
    scale = np.maximum(np.abs(rg[fin]).max(axis=1, keepdims=True), 1e-30)
    assert np.max(np.abs(g[fin] - rg[fin]) / scale) < 1e-10


def test_f32_matches_oracle_on_f32_maps(ctx, orc):
    """The FP32 fused path (single- and multi-node bands) against the FP64
    oracle run on the same FP32 node maps, with curves that start off the
    map, leave it part-way, end exactly on the last pixel centre, or meet a
    NaN epoch time."""
    rng = np.random.default_rng(16)
    m, spec, nodes, obs, sets = _setup(ctx, "f32", rng)
    B, Ny, Nx = nodes.B, 128, 160
    t_nan = np.array([1.0, np.nan, 7.0])
    sets.append([(np.array([3.0, 8.0]), np.array([-1.6, -1.8]), np.array([0.05, 0.05])),
                 (t_nan, np.full(3, -1.7), np.full(3, 0.05))][:B])
    sets.append([(np.array([2.0, 5.0, 9.0]), np.full(3, -1.7), np.full(3, 0.06)),
                 (np.array([1.0, 4.0, 12.0]), np.full(3, -1.5), np.full(3, 0.04))][:B])
    obs = lcm.Observations(ctx, B, sets)
    prm, prof, obsi = _params(rng, 400, Ny, Nx, nodes.P, 3)
    prm[0, 0] = -5.0                                   # off the map from the start
    prm[1, :4] = (Nx - 1 - 10.0, 60.0, 0.0, 1.0)       # leaves through x = Nx - 1
    obsi[1] = 1
    prm[2, :4] = (Nx - 1 - 12.0, 50.0, 0.0, 1.0)       # x = Nx - 1 exactly at t = 12
    obsi[2] = 4
    obsi[3] = 3                                        # NaN epoch time
    bad_prof = prof.copy()
    bad_prof[4], bad_prof[5] = -1, nodes.P              # no such profile: (+inf, 0)
    c, g = lcm.chi2_grad(ctx, nodes, obs, torch.tensor(prm, device="cuda"),
                         torch.tensor(bad_prof, device="cuda"), torch.tensor(obsi, device="cuda"))
    rc, rg = _oracle(orc, nodes, sets, prm, prof, obsi)
    c, g = c.cpu().numpy(), g.cpu().numpy()
    assert np.all(np.isinf(c[4:6])) and not g[4:6].any()
    rc[4:6], rg[4:6] = np.inf, 0.0
    fin = np.isfinite(rc)
    assert not fin[0] and not fin[3] and fin[2]
    assert np.array_equal(np.isfinite(c), fin)
    assert not g[~fin].any()
    assert np.max(np.abs(c[fin] - rc[fin]) / np.abs(rc[fin])) < 1e-5
    scale = np.maximum(np.abs(rg[fin]).max(axis=1, keepdims=True), 1e-30)
    assert np.max(np.abs(g[fin] - rg[fin]) / scale) < 1e-4


@pytest.mark.parametrize("ne", [100, 128, 129, 300])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_shared_observation_set_matches_oracle(ctx, orc, prec, ne):
    """Every curve on one observation set: a band of at most 128 epochs is
    staged in LDS once per wavefront and read from there, a longer band from
    global memory; chi2 and gradient against the oracle on the same maps,
    including a curve that leaves the map (+inf) between staged curves."""
    rng = np.random.default_rng(ne)
    m, spec, nodes, _, _ = _setup(ctx, prec, rng)
    B = nodes.B
    sets = [[(np.sort(rng.uniform(0, 30, ne)), rng.normal(-1.7, 0.3, ne), rng.uniform(0.03, 0.08, ne))
             for _ in range(B)]]
    obs = lcm.Observations(ctx, B, sets)
    prm, prof, obsi = _params(rng, 600, 128, 160, nodes.P, 1)
    prm[37, 0] = -5.0
    c, g = lcm.chi2_grad(ctx, nodes, obs, torch.tensor(prm, device="cuda"), torch.tensor(prof, device="cuda"),
                         torch.tensor(obsi, device="cuda"))
    rc, rg = _oracle(orc, nodes, sets, prm, prof, obsi)
    c, g = c.cpu().numpy(), g.cpu().numpy()
    fin = np.isfinite(rc)
    assert not fin[37] and fin.sum() > 500
    assert np.array_equal(np.isfinite(c), fin)
    tc, tg = (1e-11, 1e-10) if prec == "f64" else (1e-5, 1e-4)
    assert np.max(np.abs(c[fin] - rc[fin]) / np.abs(rc[fin])) < tc
    scale = np.maximum(np.abs(rg[fin]).max(axis=1, keepdims=True), 1e-30)
    assert np.max(np.abs(g[fin] - rg[fin]) / scale) < tg


def test_adjoint_matches_central_differences(ctx):
    rng = np.random.default_rng(12)
    m, spec, nodes, obs, sets = _setup(ctx, "f64", rng)
    prm, prof, obsi = _params(rng, 8, 128, 160, nodes.P, obs.nobs)
    obsi[:] = 1
    P = torch.tensor(prm, device="cuda")
    pr, ob = torch.tensor(prof, device="cuda"), torch.tensor(obsi, device="cuda")
    c0, g0 = lcm.chi2_grad(ctx, nodes, obs, P, pr, ob)
    h = np.array([1e-6, 1e-6, 1e-7, 1e-7, 1e-6, 1e-7, 1e-6])
    for q in range(7):
        Pp, Pm = P.clone(), P.clone()
        Pp[:, q] += h[q]
        Pm[:, q] -= h[q]
        cp, _ = lcm.chi2_grad(ctx, nodes, obs, Pp, pr, ob)
        cm, _ = lcm.chi2_grad(ctx, nodes, obs, Pm, pr, ob)
        fd = ((cp - cm) / (2 * h[q])).cpu().numpy()
        ad = g0[:, q].cpu().numpy()
        err = np.abs(fd - ad) / np.maximum(np.abs(ad), 1e-3 * np.abs(g0.cpu().numpy()).max())
        assert np.median(err) < 1e-4 and np.max(err) < 2e-2, (q, fd, ad)


def test_node_derivative_matches_convolution_fd(ctx):
    """The d/drhalf node map equals central differences of the convolved map
    in rhalf at a fixed grid size n (Gaussian: n = ceil(4 rhalf) stays 18)."""
    rng = np.random.default_rng(13)
    m = _map(rng, 96, 96)
    spec = api.MapSpectrum(ctx, torch.tensor(m, device="cuda"), "f64")
    r, h = 4.3, 1e-5
    node = spec.convolve("gaussian", r, value=False, node=True)
    vp = spec.convolve("gaussian", r + h)
    vm = spec.convolve("gaussian", r - h)
    fd = ((vp - vm) / (2 * h)).cpu().numpy()
    d = node[..., 1].cpu().numpy()
    assert np.max(np.abs(fd - d)) < 1e-7 * np.max(np.abs(d))


@pytest.mark.parametrize("rhalf", [3.0, 8.1])
def test_disc_node_derivative_against_convolution_fd(ctx, rhalf):
    """The uniform disc's d/drhalf node map (a one-pixel rim ramp) against
    central differences over rhalf +- 1 px of the device's own value
    convolutions: within 15% (L2), correlation above 0.999 (the discrete
    disc changes by whole pixel rings, so a pixel is the finest meaningful
    step)."""
    n, d = 128, 1.0
    rng = np.random.default_rng(3)
    f = np.exp(0.5 * rng.standard_normal((n, n)))
    k = np.fft.fftfreq(n)
    H = np.exp(-2.0 * np.pi ** 2 * 9.0 * (k[:, None] ** 2 + k[None, :] ** 2))
    m = np.real(np.fft.ifft2(np.fft.fft2(f) * H))
    m *= 5.0 / m.mean()
    spec = api.MapSpectrum(ctx, torch.tensor(m, device="cuda"), "f64")
    D = spec.convolve("uniform", rhalf, value=False, node=True)[..., 1].cpu().numpy()
    fd = ((spec.convolve("uniform", rhalf + d) - spec.convolve("uniform", rhalf - d)) / (2 * d)).cpu().numpy()
    assert np.linalg.norm(D - fd) / np.linalg.norm(fd) < 0.15
    assert np.corrcoef(D.ravel(), fd.ravel())[0, 1] > 0.999
    spec.close()


def test_f32_chi2_within_1e5_of_f64(ctx):
    """Three profiles at three bands with radii 2/4/8 px (one node each, the
    radius sits on the node): FP32 maps + FP32 taps against FP64."""
    rng = np.random.default_rng(14)
    shapes, bs = ("gaussian", "uniform", "exponential"), (1.0, 2.0, 4.0)
    sm, s64, _, obs, sets = _setup(ctx, "f64", np.random.default_rng(14), Ny=512, Nx=512,
                                   radii=((2.0,),), shapes=shapes, band_scale=bs)
    s32 = api.MapSpectrum(ctx, torch.tensor(sm, device="cuda"), "f32")
    rr = [[[2.0], [4.0], [8.0]] for _ in range(3)]
    n64 = lcm.NodeSet(ctx, s64, shapes, rr, bs)
    n32 = lcm.NodeSet(ctx, s32, shapes, rr, bs)
    prm, prof, obsi = _params(rng, 2000, 512, 512, 3, obs.nobs)
    prm[:, 4] = 2.0
    prm[:, 5] = 0.0
    args = [torch.tensor(a, device="cuda") for a in (prm, prof, obsi)]
    c64, g64 = lcm.chi2_grad(ctx, n64, obs, *args)
    c32, g32 = lcm.chi2_grad(ctx, n32, obs, *args)
    fin = torch.isfinite(c64) & (c64 > 0)
    rel = ((c32 - c64).abs() / c64.abs())[fin]
    assert float(rel.max()) <= 1e-5, float(rel.max())
    # light-curve error in magnitudes: node maps themselves
    mag = (2.5 / np.log(10)) * ((n32.nodes[..., 0].double() - n64.nodes[..., 0]).abs() / n64.nodes[..., 0])
    assert float(mag.max()) < 1e-5


def test_autograd_function(ctx):
    rng = np.random.default_rng(15)
    m, spec, nodes, obs, sets = _setup(ctx, "f32", rng)
    prm, prof, obsi = _params(rng, 64, 128, 160, nodes.P, obs.nobs)
    P = torch.tensor(prm, device="cuda", requires_grad=True)
    pr, ob = torch.tensor(prof, device="cuda"), torch.tensor(obsi, device="cuda")
    c = lcm.lightcurve_chi2(P, ctx, nodes, obs, pr, ob)
    w = torch.linspace(0.5, 1.5, 64, device="cuda", dtype=torch.float64)
    (c * w).sum().backward()
    c2, g2 = lcm.chi2_grad(ctx, nodes, obs, P.detach(), pr, ob)
    assert torch.equal(c.detach(), c2)
    assert torch.allclose(P.grad, w[:, None] * g2, rtol=0, atol=0)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_convolve_entry_matches_separate_calls(ctx, prec):
    """mlc_lc_chi2_grad_convolve (node maps and chi2 in one call) against
    mlc_map_convolve_nodes followed by
    mlc_lc_chi2_grad: node maps, chi2 and gradient bit for bit; with no
    curves every node map is still written."""
    rng = np.random.default_rng(77)
    _, _, nodes, obs, _ = _setup(ctx, prec, rng)
    prm, prof, obsi = _params(rng, 500, 128, 160, nodes.P, 3)
    prm[11, 0] = -5.0
    args = (torch.tensor(prm, device="cuda"), torch.tensor(prof, device="cuda"), torch.tensor(obsi, device="cuda"))
    nodes.compute()
    ref = nodes.nodes.clone()
    c0, g0 = lcm.chi2_grad(ctx, nodes, obs, *args)
    nodes.nodes.fill_(float("nan"))
    c1, g1 = lcm.chi2_grad(ctx, nodes, obs, *args, convolve=True)
    assert torch.equal(nodes.nodes, ref)
    assert not torch.isfinite(c0[11]) and torch.isfinite(c0).sum() > 400
    assert torch.equal(c1, c0) and torch.equal(g1, g0)
    nodes.nodes.fill_(float("nan"))
    e = (args[0][:0].contiguous(), args[1][:0].contiguous(), args[2][:0].contiguous())
    c2, g2 = lcm.chi2_grad(ctx, nodes, obs, *e, convolve=True)
    assert c2.shape == (0,) and torch.equal(nodes.nodes, ref)


def test_convolve_entry_rejects_bad_tables(ctx):
    """Bad node shapes or node tables are refused before any launch: the
    status is MLC_ERR_ARG and the node maps keep their contents."""
    import ctypes as C
    from microsn_amd._lib import lib, ptr

    rng = np.random.default_rng(78)
    _, _, nodes, obs, _ = _setup(ctx, "f32", rng)
    prm, prof, obsi = _params(rng, 50, 128, 160, nodes.P, 3)
    P_, pr, oi = (torch.tensor(prm, device="cuda"), torch.tensor(prof, device="cuda"),
                  torch.tensor(obsi, device="cuda"))
    chi2 = torch.empty(50, dtype=torch.float64, device="cuda")
    grad = torch.empty((50, 7), dtype=torch.float64, device="cuda")
    nodes.nodes.fill_(7.0)
    torch.cuda.synchronize()

    def call(shp, off):
        return lib().mlc_lc_chi2_grad_convolve(
            nodes.spectrum.h, shp.ctypes.data, nodes.pix, ptr(nodes.nodes), nodes.nodes.shape[0], nodes.P,
            nodes.B, off.ctypes.data, ptr(nodes.node_cnt), ptr(nodes.node_rad), ptr(nodes.band_scale), 50,
            ptr(P_), ptr(pr), ptr(oi), obs.nobs, ptr(obs.off), obs.nepoch, ptr(obs.t), ptr(obs.mobs),
            ptr(obs.sig), ptr(chi2), ptr(grad))

    good_shp = np.ascontiguousarray(nodes.node_shape, np.int32)
    bad_shp = good_shp.copy()
    bad_shp[-1] = 99
    bad_off = nodes.node_off.copy()
    bad_off[-1] = nodes.nodes.shape[0]  # past the last node
    assert call(bad_shp, nodes.node_off) == 1
    assert call(good_shp, bad_off) == 1
    torch.cuda.synchronize()
    assert bool((nodes.nodes == 7.0).all())
    assert call(good_shp, nodes.node_off) == 0
    torch.cuda.synchronize()
    assert not bool((nodes.nodes == 7.0).all())
