"""GPU parity of the library's 2-D FFT and of the map x profile convolution
(cuFFT D2Z/Z2D + kernel_multiplyFFTs + setKernel, gpu_functions.cu:151-200).

cuFFT itself pins nothing (no reference test holds its outputs), so the
transforms are checked against numpy's FFT in FP64 and the convolution
against the oracle's FP64 FFT convolution and direct circular sums."""
import numpy as np
import pytest
import torch

from microsn_amd import api

pytestmark = pytest.mark.gpu

SIZES = [(8, 2), (16, 16), (48, 64), (60, 100), (250, 256), (1000, 1000), (96, 4096), (4096, 40),
         # the widths with compile-time row kernels (2 M for M in 500 ... 5000)
         (8, 1280), (6, 1536), (16, 2000), (12, 2560), (8, 3072), (10, 5000), (8, 10000)]


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("Ny,Nx", SIZES)
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_r2c_c2r_match_numpy(ctx, Ny, Nx, prec):
    rng = np.random.default_rng(Nx * 7 + Ny)
    x = rng.standard_normal((Ny, Nx))
    dt = torch.float64 if prec == "f64" else torch.float32
    X = api.fft_r2c(ctx, torch.tensor(x, dtype=dt, device="cuda")).cpu().numpy()
    ref = np.fft.rfft2(x)
    tol = 1e-13 if prec == "f64" else 3e-6
    assert _rel(X, ref) < tol
    # inverse: unnormalised, input not modified
    Xin = torch.tensor(ref, dtype=torch.complex128 if prec == "f64" else torch.complex64, device="cuda")
    keep = Xin.clone()
    y = api.fft_c2r(ctx, Xin, Nx).cpu().numpy()
    assert torch.equal(Xin, keep)
    assert _rel(y, x * Nx * Ny) < tol


def _smooth_map(rng, Ny, Nx):
    g = rng.standard_normal((Ny, Nx))
    return 5.0 * np.exp(0.5 * g) / np.exp(0.125)


@pytest.mark.parametrize("shape", ["gaussian", "uniform", "exponential", "gaussian_hole",
                                   "thermal_hole", "wavy"])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_convolve_matches_oracle(ctx, orc, shape, prec):
    rng = np.random.default_rng(3)
    Ny, Nx = 64, 80
    m = _smooth_map(rng, Ny, Nx)
    rhalf, pix, rin, a = 4.3, 1.0, 1.5, 1.0
    spec = api.MapSpectrum(ctx, torch.tensor(m, device="cuda"), prec)
    val, node = spec.convolve(shape, rhalf, pix, rin=rin, a=a, value=True, node=True)
    prof, dprof = orc.profile(shape, rhalf, pix, rin=rin, a=a, deriv=True)
    ref = orc.fft_convolve(m, orc.set_kernel(prof, Nx, Ny))
    dref = orc.fft_convolve(m, orc.set_kernel(dprof, Nx, Ny))
    tol = 1e-12 if prec == "f64" else 2e-6
    v = val.cpu().numpy().astype(np.float64)
    nd = node.cpu().numpy().astype(np.float64)
    assert _rel(v, ref) < tol
    assert _rel(nd[..., 0], ref) < tol
    # the derivative map has zero mean; compare against its own scale
    assert np.max(np.abs(nd[..., 1] - dref)) < tol * 20 * max(np.max(np.abs(ref)), 1.0)
    # plain (reference-structured) path gives the same maps
    pv, pn = spec.convolve(shape, rhalf, pix, rin=rin, a=a, value=True, node=True, plain=True)
    assert _rel(pv.cpu().numpy().astype(np.float64), ref) < tol
    assert np.max(np.abs(pn.cpu().numpy()[..., 1].astype(np.float64) - dref)) < tol * 20 * np.max(np.abs(ref))


def test_convolve_direct_sum_and_identity(ctx, orc):
    """Direct circular convolution on a tiny map, and the mean preserved."""
    rng = np.random.default_rng(11)
    m = _smooth_map(rng, 24, 30)
    spec = api.MapSpectrum(ctx, torch.tensor(m, device="cuda"), "f64")
    v = spec.convolve("gaussian", 2.2).cpu().numpy()
    prof = orc.profile("gaussian", 2.2)
    direct = orc.direct_convolve(m, orc.set_kernel(prof, 30, 24))
    assert _rel(v, direct) < 1e-12
    assert abs(v.mean() - m.mean()) < 1e-12 * m.mean()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_convolve_4096_properties(ctx, prec):
    """Full-size map: mean preserved and a translated impulse lands where the
    wrapped kernel puts it (the +1/2 px corner convention of setKernel)."""
    N = 4096
    m = torch.full((N, N), 2.0, dtype=torch.float64, device="cuda")
    m[1000, 2000] += 1000.0
    spec = api.MapSpectrum(ctx, m, prec)
    v, nd = spec.convolve("gaussian", 5.0, node=True)
    v = v.double()
    # node maps (compile-time row FFT of the packed value + i derivative) agree with the value map
    assert float((nd[..., 0].double() - v).abs().max()) < (1e-12 if prec == "f64" else 2e-6) * 1002.0
    mean = float(v.mean())
    assert abs(mean - float(m.mean())) < (1e-12 if prec == "f64" else 1e-6) * float(m.mean())
    # centroid of the excess around the impulse: offset (-1/2, -1/2) px from
    # kernel[dy][dx] = P((dx+1/2) pix, (dy+1/2) pix)
    w = (v - 2.0)[980:1021, 1980:2021]
    yy, xx = torch.meshgrid(torch.arange(980, 1021, device="cuda", dtype=torch.float64),
                            torch.arange(1980, 2021, device="cuda", dtype=torch.float64), indexing="ij")
    cy = float((w * yy).sum() / w.sum())
    cx = float((w * xx).sum() / w.sum())
    assert abs(cy - 999.5) < 1e-3 and abs(cx - 1999.5) < 1e-3


def test_profile_on_device_matches_oracle(ctx, orc):
    for shape in ["gaussian", "uniform", "exponential", "gaussian_hole", "thermal_hole", "wavy"]:
        p, dp = api.profile(ctx, shape, 6.7, 1.0, rin=2.0, a=1.0)
        rp, rdp = orc.profile(shape, 6.7, 1.0, rin=2.0, a=1.0, deriv=True)
        assert p.shape == rp.shape
        assert np.max(np.abs(p.cpu().numpy() - rp)) < 1e-14
        assert np.max(np.abs(dp.cpu().numpy() - rdp)) < 1e-13


def test_unsupported_sizes_rejected(ctx):
    m = torch.ones((14, 22), dtype=torch.float64, device="cuda")  # 11 = prime
    with pytest.raises(Exception, match="prime factor"):
        api.MapSpectrum(ctx, m, "f32")
    m = torch.ones((16, 15), dtype=torch.float64, device="cuda")  # odd width
    with pytest.raises(Exception, match="even"):
        api.MapSpectrum(ctx, m, "f32")
    spec = api.MapSpectrum(ctx, torch.ones((32, 32), dtype=torch.float64, device="cuda"), "f32")
    with pytest.raises(Exception, match="larger than the map"):
        spec.convolve("gaussian", 40.0)


@pytest.mark.parametrize("Ny,Nx,shape,rhalf", [(36, 40, "gaussian", 8.0),     # rows wrap, partial row block
                                               (256, 256, "uniform", 50.0),   # FP64 falls back to column FFTs
                                               (100, 120, "exponential", 3.1),
                                               (64, 1024, "gaussian", 6.0),   # compile-time row FFTs
                                               (40, 2048, "uniform", 3.0),
                                               (24, 10240, "gaussian", 3.0),   # FP64 node rows in two passes
                                               (16, 16384, "uniform", 2.0)])   # longest compile-time rows
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_convolve_wide_kernels_match_oracle(ctx, orc, Ny, Nx, shape, rhalf, prec):
    rng = np.random.default_rng(Ny + Nx)
    m = _smooth_map(rng, Ny, Nx)
    spec = api.MapSpectrum(ctx, torch.tensor(m, device="cuda"), prec)
    val, node = spec.convolve(shape, rhalf, 1.0, value=True, node=True)
    prof, dprof = orc.profile(shape, rhalf, 1.0, deriv=True)
    ref = orc.fft_convolve(m, orc.set_kernel(prof, Nx, Ny))
    dref = orc.fft_convolve(m, orc.set_kernel(dprof, Nx, Ny))
    tol = 1e-12 if prec == "f64" else 2e-6
    assert _rel(val.cpu().numpy().astype(np.float64), ref) < tol
    nd = node.cpu().numpy().astype(np.float64)
    assert _rel(nd[..., 0], ref) < tol
    assert np.max(np.abs(nd[..., 1] - dref)) < tol * 20 * max(np.max(np.abs(ref)), 1.0)


@pytest.mark.parametrize("Ny,Nx,shape,rhalf", [(64, 1024, "gaussian", 6.0),
                                               (40, 2048, "uniform", 3.0),
                                               (100, 256, "wavy", 3.0),
                                               (64, 8192, "exponential", 2.0),    # 2 rows (FP32) / 1 row per block
                                               (48, 16384, "uniform", 2.0),       # FP32 one row per block
                                               (512, 4096, "gaussian", 8.0),
                                               (20, 1024, "gaussian", 4.0)])      # too few rows: separate kernels
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_node_only_request_fused_kernel(ctx, orc, Ny, Nx, shape, rhalf, prec):
    """A node map requested alone runs the y-filter and the row FFT in one
    kernel (k_node_fused); it matches the oracle's FFT convolutions within
    the usual tolerance, and the node map of a request that also wants the
    value map (separate y-filter and row kernels) within 4 ulp of the map's
    scale: the same sums in the same order, compiled into different code."""
    rng = np.random.default_rng(Ny * 3 + Nx)
    m = _smooth_map(rng, Ny, Nx)
    spec = api.MapSpectrum(ctx, torch.tensor(m, device="cuda"), prec)
    alone = spec.convolve(shape, rhalf, 1.0, value=False, node=True)
    _, both = spec.convolve(shape, rhalf, 1.0, value=True, node=True)
    eps = 2.0 ** -23 if prec == "f32" else 2.0 ** -52
    assert float((alone - both).abs().max()) <= 4 * eps * float(both.abs().max())
    prof, dprof = orc.profile(shape, rhalf, 1.0, deriv=True)
    ref = orc.fft_convolve(m, orc.set_kernel(prof, Nx, Ny))
    dref = orc.fft_convolve(m, orc.set_kernel(dprof, Nx, Ny))
    tol = 1e-12 if prec == "f64" else 2e-6
    nd = alone.cpu().numpy().astype(np.float64)
    assert _rel(nd[..., 0], ref) < tol
    assert np.max(np.abs(nd[..., 1] - dref)) < tol * 20 * max(np.max(np.abs(ref)), 1.0)


@pytest.mark.parametrize("Nx,Ny", [(2048, 48), (2560, 40), (4096, 32)])
@pytest.mark.parametrize("shape,rhalf", [("uniform", 1.0),      # h = 5: whole rounds of taps only
                                         ("uniform", 1.5),      # h = 6: one tap after a round
                                         ("gaussian", 1.0),     # h = 2: a partial round only
                                         ("exponential", 1.5),  # h = 3
                                         ("uniform", 4.0),      # h = 9: a round and four taps
                                         ("gaussian", 6.0)])    # h = 13: two rounds and three taps
def test_node_filter_rounds_and_tails(ctx, orc, Nx, Ny, shape, rhalf):
    """The fused node kernels filter with register rings of five tap slots
    (four rows per workgroup at 2048 and 4096 FP32 columns, two at 2560):
    tap counts that leave every remainder modulo five, against the separate
    y-filter and row kernels (within 4 ulp of the map's scale) and the
    oracle."""
    rng = np.random.default_rng(Nx + int(10 * rhalf))
    m = _smooth_map(rng, Ny, Nx)
    spec = api.MapSpectrum(ctx, torch.tensor(m, device="cuda"), "f32")
    alone = spec.convolve(shape, rhalf, 1.0, value=False, node=True)
    _, both = spec.convolve(shape, rhalf, 1.0, value=True, node=True)
    assert float((alone - both).abs().max()) <= 4 * 2.0 ** -23 * float(both.abs().max())
    prof, dprof = orc.profile(shape, rhalf, 1.0, deriv=True)
    ref = orc.fft_convolve(m, orc.set_kernel(prof, Nx, Ny))
    dref = orc.fft_convolve(m, orc.set_kernel(dprof, Nx, Ny))
    nd = alone.cpu().numpy().astype(np.float64)
    assert _rel(nd[..., 0], ref) < 2e-6
    assert np.max(np.abs(nd[..., 1] - dref)) < 2e-6 * 20 * max(np.max(np.abs(ref)), 1.0)
    spec.close()


@pytest.mark.parametrize("Ny,Nx", [(128, 160), (128, 256)])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_convolve_nodes_equals_one_by_one(ctx, Ny, Nx, prec):
    """mlc_map_convolve_nodes writes, bit for bit, what one mlc_map_convolve
    call per node writes: mixed shapes, radii and tap counts in one batch
    (at 128 x 256 in FP64 the 80-pixel Gaussian takes the column-FFT
    fallback)."""
    import ctypes as C

    from microsn_amd._lib import SHAPES, lib
    from microsn_amd.lightcurves import NodeSet

    rng = np.random.default_rng(Ny + Nx)
    m = _smooth_map(rng, Ny, Nx)
    spec = api.MapSpectrum(ctx, torch.tensor(m, device="cuda"), prec)
    shapes = ["gaussian", "uniform", "exponential", "gaussian", "wavy"]
    rads = [2.0, 3.5, 1.5, 20.0 if Nx == 256 else 6.0, 2.5]
    ns = NodeSet(ctx, spec, shapes, [[[r]] for r in rads], [1.0])  # batched
    for k, (s, r) in enumerate(zip(shapes, rads)):
        one = torch.empty_like(ns.nodes[k])
        ctx.check(lib().mlc_map_convolve(spec.h, SHAPES[s], r, 1.0, 0.0, 1.0, None, one.data_ptr()))
        assert torch.equal(one, ns.nodes[k]), (s, r)
    # an empty batch is a no-op; a null node pointer is rejected before any work
    ctx_h = spec.h
    assert lib().mlc_map_convolve_nodes(ctx_h, 0, None, None, 1.0, None, None, None) == 0
    shp = np.array([0], np.int32)
    rad = np.array([2.0])
    dst = (C.c_void_p * 1)(None)
    assert lib().mlc_map_convolve_nodes(ctx_h, 1, shp.ctypes.data, rad.ctypes.data, 1.0, None, None,
                                        C.addressof(dst)) != 0


def test_context_close_closes_its_maps():
    """Closing a context first closes the map spectra made on it (a map keeps
    a pointer to its context), so a map that outlives its context is inert."""
    c2 = api.Context(0)
    spec = api.MapSpectrum(c2, torch.ones((16, 16), dtype=torch.float64, device="cuda"), "f32")
    c2.close()
    assert spec.h is None
    spec.close()
    del spec


@pytest.mark.parametrize("Ny,Nx", [(40, 1000), (32, 1280), (24, 2560), (24, 3072), (20, 10000)])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_node_map_rows_at_fixed_widths(ctx, orc, Ny, Nx, prec):
    """Node maps (mu, dmu/drhalf) at widths whose rows run the compile-time
    row kernels: against the oracle's FP64 FFT convolutions of the profile
    and of its derivative."""
    rng = np.random.default_rng(Nx + Ny)
    m = _smooth_map(rng, Ny, Nx)
    spec = api.MapSpectrum(ctx, torch.tensor(m, device="cuda"), prec)
    got = spec.convolve("gaussian", 3.0, value=False, node=True).double().cpu().numpy()
    p, dp = orc.profile("gaussian", 3.0, 1.0, deriv=True)
    want_v = orc.fft_convolve(m, orc.set_kernel(p, Nx, Ny))
    want_d = orc.fft_convolve(m, orc.set_kernel(dp, Nx, Ny))
    tol = 1e-12 if prec == "f64" else 2e-6
    assert _rel(got[..., 0], want_v) < tol
    assert _rel(got[..., 1], want_d) < tol * 10
    spec.close()
