"""Independent pins of the oracle functions that the GPU tests lean on most:
profile values of every shape, the multi-band chi2 and its layout, and the
trajectory chi2 with its gradient.  Each check is a direct numpy
restatement of the reference's text (profile.cpp, chi2_functions.cuh) or a
closed form of the light-curve model, evaluated without the oracle's code,
chosen so that a dropped term, a wrong constant, a swapped index or a
transposed layout fails.  CPU only."""
import numpy as np
import pytest
from scipy.integrate import quad
from scipy.optimize import brentq

from oracle import oracle as orc


def _grid(n, pix):
    """createGrid (profile.cpp:143-157) with incl = orient = 0: pixel centres
    at (j - n/2) pix + pix/2 on both axes, r = hypot(x, y)."""
    c = (np.arange(n) - n // 2) * pix + pix / 2.0
    return np.hypot(c[None, :], c[:, None])


def _even(n):
    return n + (n % 2)


def _thin_disc(rhalf, pix):
    """Exponential (profile.cpp:402-446): sigma = 0.4106 rhalf (setByHalfRadius),
    Rhalf = sigma / 0.4106, n = ceil(4 Rhalf / pix) made even; value
    1/(exp((r/sigma)^0.75) - 1) for r < 2 Rhalf, 0 beyond.  The r < r_inner
    assignment is overwritten by the following if (no else), so the inner
    truncation never applies."""
    sigma = 0.4106 * rhalf
    Rh = sigma / 0.4106
    n = _even(int(np.ceil(4.0 * Rh / pix)))
    r = _grid(n, pix)
    v = np.where(r < 2 * Rh, 1.0 / np.expm1(np.minimum(r / sigma, 1e3) ** 0.75), 0.0)
    return v / v.sum()


def _gaussian_hole(rhalf, pix, Rin):
    """GaussianHole (profile.cpp:269-311): sdev^2 = (rhalf^2 - Rin^2)/(2 ln 2),
    Rhalf = sqrt(2 ln 2 sdev^2 + Rin^2), value 0 inside Rin, exp(-r^2/(2 sdev^2))
    outside."""
    sdev = np.sqrt((rhalf ** 2 - Rin ** 2) / (2.0 * np.log(2.0)))
    Rh = np.sqrt(2.0 * np.log(2.0) * sdev ** 2 + Rin ** 2)
    n = _even(int(np.ceil(4.0 * Rh / pix)))
    r = _grid(n, pix)
    v = np.where(r < Rin, 0.0, np.exp(-r * r / (2 * sdev ** 2)))
    return v / v.sum()


def _thermal_hole(rhalf, pix):
    """ThermalHole (profile.cpp:315-354): Rin = rhalf/4, Rhalf = 4 Rin, value
    (1 - sqrt(Rin/r)) (3 Rin / 2 pi) / r^3 for r > Rin, else 0."""
    Rin = rhalf / 4.0
    n = _even(int(np.ceil(4.0 * (4.0 * Rin) / pix)))
    r = _grid(n, pix)
    with np.errstate(divide="ignore", invalid="ignore"):
        v = np.where(r > Rin, (1.0 - np.sqrt(Rin / r)) * (3.0 * Rin / (2.0 * np.pi)) / r ** 3, 0.0)
    return v / v.sum()


def _wavy(rhalf, pix, a):
    """Wavy with 3 nodes (profile.cpp:358-398, profile.hpp:156): b = 3 pi/(2 rhalf),
    Rhalf = 3 pi/(2 b), value sin(b r)^2/(a r) for r < 2 Rhalf, else 0."""
    b = 3 * np.pi / (2.0 * rhalf)
    Rh = 3 * np.pi / (2.0 * b)
    n = _even(int(np.ceil(4.0 * Rh / pix)))
    r = _grid(n, pix)
    v = np.where(r < 2 * Rh, np.sin(b * r) ** 2 / (a * r), 0.0)
    return v / v.sum()


@pytest.mark.parametrize("rhalf,pix", [(2.0, 1.0), (4.0, 1.0), (8.0, 1.0), (3.3, 0.7), (0.05, 0.0025)])
def test_thin_disc_values_from_profile_cpp(rhalf, pix):
    p = orc.profile("exponential", rhalf, pix)
    want = _thin_disc(rhalf, pix)
    assert p.shape == want.shape
    assert np.max(np.abs(p - want)) <= 1e-13 * np.max(want)


@pytest.mark.parametrize("rhalf,pix,Rin", [(3.0, 1.0, 1.0), (6.5, 0.8, 2.5), (10.0, 1.0, 0.0)])
def test_gaussian_hole_values_from_profile_cpp(rhalf, pix, Rin):
    p = orc.profile("gaussian_hole", rhalf, pix, rin=Rin)
    want = _gaussian_hole(rhalf, pix, Rin)
    assert p.shape == want.shape
    assert np.max(np.abs(p - want)) <= 1e-13 * np.max(want)


@pytest.mark.parametrize("rhalf,pix", [(4.0, 1.0), (8.0, 1.0), (13.5, 0.9)])
def test_thermal_hole_values_from_profile_cpp(rhalf, pix):
    p = orc.profile("thermal_hole", rhalf, pix)
    want = _thermal_hole(rhalf, pix)
    assert p.shape == want.shape
    assert np.max(np.abs(p - want)) <= 1e-13 * np.max(want)


@pytest.mark.parametrize("rhalf,pix,a", [(4.0, 1.0, 1.0), (8.0, 1.0, 0.3), (13.5, 0.9, 2.0)])
def test_wavy_values_from_profile_cpp(rhalf, pix, a):
    p = orc.profile("wavy", rhalf, pix, a=a)
    want = _wavy(rhalf, pix, a)
    assert p.shape == want.shape
    assert np.max(np.abs(p - want)) <= 1e-13 * np.max(want)


def test_thin_disc_half_light_radius():
    """The constant 0.4106 of Exponential::setByHalfRadius is the half-light
    radius of the untruncated T ~ r^-3/4 disc, I(r) = 1/(exp((r/sigma)^(3/4)) - 1):
    the radius enclosing half the flux is sigma/0.4106 to 1e-4.  The oracle's
    profile on a fine grid, truncated at 2 Rhalf, must then hold the fraction
    F(Rhalf)/F(2 Rhalf) of its flux inside Rhalf (a wrong exponent or a
    wrong sigma factor moves this by more than 1e-2)."""
    I = lambda r: r / np.expm1(min(r, 1e3) ** 0.75)  # noqa: E731  (sigma = 1)
    tot = quad(I, 0, np.inf, limit=500)[0]
    enc = lambda R: quad(I, 0, R, limit=500)[0] / tot  # noqa: E731
    R50 = brentq(lambda R: enc(R) - 0.5, 0.5, 10.0)
    assert abs(R50 * 0.4106 - 1.0) < 1e-4
    rhalf, pix = 8.0, 0.02  # 1600 x 1600 pixels
    p = orc.profile("exponential", rhalf, pix)
    r = _grid(p.shape[0], pix)
    frac = p[r < rhalf].sum() / p.sum()
    sigma = 0.4106 * rhalf
    want = enc(rhalf / sigma) / enc(2 * rhalf / sigma)
    assert abs(frac - want) < 2e-3, (frac, want)


# ---------------------------------------------------------------- multi-band chi2
def _multiband_numpy(f_obs, df_obs, mu1, mu2, gpu_quirk):
    """chi2_functions.cuh restated in numpy float32.
    CPU twin (getChi2CPU, 160-190): for pair (x, y), one running float sum
    over filters then epochs of tmp*tmp/2, tmp = (f_obs[f,t] - mu1[x,f,t] /
    mu2[y,f,t]) / df_obs[f,t].
    GPU kernel (kernelChi2, 35-68): mu2 is read at x (filter_mu2 =
    &mu2[x*F*T + f*T]), one float partial per filter over its epochs, the
    partials added to chi2_all[N*y + x] (here in filter order)."""
    N, F, T = mu1.shape
    half = np.float32(2)
    chi = np.zeros((N, N), np.float32)  # [y, x]
    m1 = mu1[None, :, :, :]  # x on axis 1
    m2 = (mu2[None, :, :, :] if gpu_quirk else mu2[:, None, :, :])  # [y, x] broadcast
    m2 = np.broadcast_to(m2, (N, N, F, T))
    for f in range(F):
        part = np.zeros((N, N), np.float32)
        for t in range(T):
            tmp = (f_obs[f, t] - m1[..., f, t] / m2[..., f, t]) / df_obs[f, t]
            if gpu_quirk:
                part = part + tmp * tmp / half
            else:
                chi = chi + tmp * tmp / half
        if gpu_quirk:
            chi = chi + part
    return chi.reshape(-1)  # index N*y + x


@pytest.mark.parametrize("N,F,T", [(1, 1, 1), (3, 2, 5), (8, 7, 10), (5, 3, 1)])
@pytest.mark.parametrize("gpu_quirk", [False, True])
def test_multiband_chi2_layout_and_bits(N, F, T, gpu_quirk):
    rng = np.random.default_rng(1000 * N + 10 * F + T)
    f_obs = rng.uniform(0.5, 2.0, (F, T)).astype(np.float32)
    df_obs = rng.uniform(0.05, 0.3, (F, T)).astype(np.float32)
    mu1 = rng.uniform(0.5, 3.0, (N, F, T)).astype(np.float32)
    mu2 = rng.uniform(0.5, 3.0, (N, F, T)).astype(np.float32)
    chi_all, _, l64 = orc.chi2_multiband(f_obs, df_obs, mu1, mu2, gpu_quirk=gpu_quirk)
    want = _multiband_numpy(f_obs, df_obs, mu1, mu2, gpu_quirk)
    assert chi_all.dtype == np.float32
    assert np.array_equal(chi_all.view(np.uint32), want.view(np.uint32))
    assert abs(l64 - np.exp(-want.astype(np.float64)).sum()) < 1e-12 * max(1.0, l64)
    if N > 1:
        # the layout check has teeth: the transpose differs (and, on the
        # CPU-twin convention, so does swapping the roles of mu1 and mu2)
        tr = want.reshape(N, N).T.reshape(-1)
        assert not np.array_equal(chi_all, tr)


# ---------------------------------------------------------------- trajectory chi2 on affine maps
def _affine_maps(Ny, Nx, rad, a, bx, cy, d, e):
    """Node maps of mu(x, y, r) = a + bx x + cy y + d r + e r^2 at radii rad:
    (value, d mu/dr) per pixel; bilinear taps of an affine field and the
    cubic Hermite interpolant between nodes are exact for it."""
    y, x = np.mgrid[0:Ny, 0:Nx].astype(np.float64)
    maps = np.zeros((len(rad), Ny, Nx, 2))
    for k, r in enumerate(rad):
        maps[k, ..., 0] = a + bx * x + cy * y + d * r + e * r * r
        maps[k, ..., 1] = d + 2 * e * r
    return maps


def _affine_closed_form(prm, band_of, t, mobs, sig, bs, a, bx, cy, d, e):
    """chi2 = sum_j ((mobs_j - m_j)/sig_j)^2 with m = dm - 2.5 log10 mu along
    x = x0 + v cos(theta) t, y = y0 + v sin(theta) t, r = bs[band] (r0 + w t),
    and its gradient by the chain rule, all in closed form."""
    x0, y0, th, v, r0, w, dm = prm
    s = np.asarray(bs)[band_of]
    x = x0 + v * np.cos(th) * t
    y = y0 + v * np.sin(th) * t
    r = s * (r0 + w * t)
    mu = a + bx * x + cy * y + d * r + e * r * r
    m = dm - 2.5 * np.log10(mu)
    res = (mobs - m) / sig
    chi2 = np.sum(res * res)
    dchi_dm = -2.0 * res / sig               # d chi2 / d m_j
    dm_dmu = -2.5 / (np.log(10.0) * mu)      # d m / d mu
    dmu_dr = d + 2 * e * r
    dmu = np.stack([
        np.full_like(t, bx),                                        # x0
        np.full_like(t, cy),                                        # y0
        v * t * (-bx * np.sin(th) + cy * np.cos(th)),               # theta
        t * (bx * np.cos(th) + cy * np.sin(th)),                    # v
        dmu_dr * s,                                                 # r0
        dmu_dr * s * t,                                             # w
    ])
    g = (dmu * (dchi_dm * dm_dmu)[None, :]).sum(1)
    return chi2, np.concatenate([g, [np.sum(dchi_dm)]])


def test_lc_chi2_grad_on_affine_maps_closed_form():
    """orc_lc_chi2_grad against the closed form along straight trajectories
    on mu = a + b x + c y + d r + e r^2 (b != c, so x and y cannot swap; two
    bands with different scales and node sets; radii inside the node ranges,
    where the Hermite interpolant is exact)."""
    rng = np.random.default_rng(77)
    Ny, Nx, B = 60, 80, 2
    coef = dict(a=3.0, bx=0.021, cy=-0.013, d=0.17, e=-0.004)
    rad = [0.5, 2.0, 5.0, 1.0, 3.0, 9.0]  # band 0: nodes 0-2, band 1: nodes 3-5
    maps = _affine_maps(Ny, Nx, rad, **coef)
    bs = [1.0, 1.7]
    t = [np.sort(rng.uniform(0.0, 30.0, 23)), np.sort(rng.uniform(2.0, 35.0, 17))]
    mobs = [rng.normal(-1.2, 0.1, 23), rng.normal(-1.1, 0.1, 17)]
    sig = [rng.uniform(0.02, 0.08, 23), rng.uniform(0.02, 0.08, 17)]
    obs_off = np.array([0, 23, 40], np.int32)
    n = 24
    prm = np.stack([rng.uniform(25, 55, n), rng.uniform(20, 40, n), rng.uniform(0, 2 * np.pi, n),
                    rng.uniform(0.1, 0.6, n), rng.uniform(0.6, 1.0, n), rng.uniform(0.0, 0.08, n),
                    rng.normal(0.0, 0.2, n)], 1)
    chi2, grad = orc.lc_chi2_grad(maps, B, [0, 3], [3, 3], rad, None, bs, prm, np.zeros(n, np.int32),
                                  np.zeros(n, np.int32), obs_off, np.concatenate(t),
                                  np.concatenate(mobs), np.concatenate(sig))
    tt = np.concatenate(t)
    band_of = np.repeat([0, 1], [23, 17])
    for c in range(n):
        r = np.asarray(bs)[band_of] * (prm[c, 4] + prm[c, 5] * tt)
        lo = np.where(band_of == 0, rad[0], rad[3])
        hi = np.where(band_of == 0, rad[2], rad[5])
        assert np.all((r > lo) & (r < hi)), "radii must stay inside the node ranges"
        wc, wg = _affine_closed_form(prm[c], band_of, tt, np.concatenate(mobs), np.concatenate(sig),
                                     bs, **coef)
        assert abs(chi2[c] - wc) < 1e-11 * wc, (c, chi2[c], wc)
        assert np.all(np.abs(grad[c] - wg) < 1e-9 * np.maximum(1.0, np.abs(wg))), (c, grad[c], wg)


def test_lc_chi2_grad_affine_detects_swapped_axes():
    """The affine pin has teeth: the closed form with x and y slopes swapped,
    or theta measured from the y axis, disagrees with the oracle."""
    rng = np.random.default_rng(78)
    coef = dict(a=2.5, bx=0.03, cy=-0.01, d=0.0, e=0.0)
    maps = _affine_maps(40, 40, [1.0], **coef)
    t = np.linspace(0, 20, 21)
    mobs = rng.normal(-1.0, 0.1, 21)
    sig = np.full(21, 0.05)
    prm = np.array([[12.0, 15.0, 0.6, 0.5, 1.0, 0.0, 0.0]])
    chi2, grad = orc.lc_chi2_grad(maps, 1, [0], [1], [1.0], None, [1.0], prm, [0], [0],
                                  np.array([0, 21], np.int32), t, mobs, sig)
    band_of = np.zeros(21, int)
    good = _affine_closed_form(prm[0], band_of, t, mobs, sig, [1.0], **coef)
    assert abs(chi2[0] - good[0]) < 1e-11 * good[0]
    swapped = dict(coef, bx=coef["cy"], cy=coef["bx"])
    bad = _affine_closed_form(prm[0], band_of, t, mobs, sig, [1.0], **swapped)
    assert abs(chi2[0] - bad[0]) > 1e-3 * good[0]
    p2 = prm[0].copy()
    p2[2] = np.pi / 2 - p2[2]
    bad2 = _affine_closed_form(p2, band_of, t, mobs, sig, [1.0], **coef)
    assert abs(chi2[0] - bad2[0]) > 1e-3 * good[0]


def _smooth_field(seed, n):
    """A log-normal field smoothed with a sigma-3 px Gaussian, mean 5."""
    rng = np.random.default_rng(seed)
    f = np.exp(0.5 * rng.standard_normal((n, n)))
    k = np.fft.fftfreq(n)
    H = np.exp(-2.0 * np.pi ** 2 * 9.0 * (k[:, None] ** 2 + k[None, :] ** 2))
    m = np.real(np.fft.ifft2(np.fft.fft2(f) * H))
    return m * (5.0 / m.mean())


@pytest.mark.parametrize("rhalf", [3.0, 6.2, 8.1, 12.0])
def test_disc_radius_derivative_against_finite_differences(rhalf):
    """The uniform disc's d/drhalf (the rim's indicator step taken as a
    one-pixel ramp of height sqrt2/pix, R = sqrt2 rhalf, plus the
    normalisation term) against central differences of the convolved value
    maps over rhalf +- 1 px on a smooth map: the discrete disc gains or
    loses whole pixel rings, so only a step of about a pixel is meaningful,
    and there the two agree within 15% (L2) with correlation above 0.999.
    dR/drhalf = 1 instead of sqrt2 misses by 22-33%, twice the ramp by
    90%, a flipped sign or a dropped normalisation term by far more."""
    n, d = 128, 1.0
    m = _smooth_field(3, n)
    p, dp = orc.profile("uniform", rhalf, 1.0, deriv=True)
    D = orc.fft_convolve(m, orc.set_kernel(dp, n, n))
    vp = orc.fft_convolve(m, orc.set_kernel(orc.profile("uniform", rhalf + d, 1.0), n, n))
    vm = orc.fft_convolve(m, orc.set_kernel(orc.profile("uniform", rhalf - d, 1.0), n, n))
    fd = (vp - vm) / (2 * d)
    assert abs(dp.sum()) < 1e-12  # the normalised profile keeps its sum
    assert np.linalg.norm(D - fd) / np.linalg.norm(fd) < 0.15
    assert np.corrcoef(D.ravel(), fd.ravel())[0, 1] > 0.999
