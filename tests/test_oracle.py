"""Pins of the FP64 CPU oracle (oracle/), against which the GPU tests check
their results.  Each test checks an oracle function against something the oracle
does not share code with: the reference's own C++ compiled from its sources
(oracle/_ref: locations.cpp, mpd.cpp, integration.cpp), values derived by
hand from the reference's inputs (grid_input.json), closed forms, numpy, or
brute force on tiny inputs.  CPU only."""
import numpy as np
import pytest

from oracle import oracle as orc


def _ref():
    r = orc.ref()
    if r is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    return r


# grid_input.json of the reference: sizes (R_E), map 10000 px over 25 R_E,
# Nside 99, DL/DS/DLS (Mpc), one light curve (d, s, t)
GRID_SIZES = [0.0, 0.05, 0.1, 0.2, 0.3]
GRID_NMAP, GRID_SS, GRID_NSIDE = 10000, 25.0, 99
GRID_D = [0.7, 1.2, 1.4, 1.1, 0.9]
GRID_S = [0.1, 0.15, 0.1, 0.1, 0.15]
GRID_T = [5.0, 7.0, 12.0, 14.0, 17.0]
DL, DS, DLS = 1570.0, 1750.0, 850.0


# ---------------------------------------------------------------- profiles
def test_profile_sizes_of_grid_input():
    """profile.hpp:142-166 via std::to_string/std::stof parameters: the
    kernel sizes derived by hand for grid_input.json."""
    pix = orc.parity_round(GRID_SS / GRID_NMAP)
    for shape, want in [("gaussian", [82, 162, 322, 482]), ("uniform", [64, 120, 234, 346]),
                        ("exponential", [82, 162, 322, 482])]:
        got = [orc.profile_params(shape, orc.parity_round(s), pix)[0] for s in GRID_SIZES[1:]]
        assert got == want, (shape, got)


@pytest.mark.parametrize("shape", ["uniform", "gaussian", "gaussian_hole", "thermal_hole", "wavy",
                                   "exponential"])
def test_profile_normalised(shape):
    p, dp = orc.profile(shape, 3.0, 1.0, deriv=True)
    assert p.shape[0] % 2 == 0
    assert abs(p.sum() - 1.0) < 1e-13
    # d/drhalf of a normalised profile sums to zero
    assert abs(dp.sum()) < 1e-12 * max(1.0, np.abs(dp).sum())
    # radially symmetric on the grid: invariant under both flips and transposition
    assert np.allclose(p, p[::-1, :], rtol=0, atol=1e-16)
    assert np.allclose(p, p[:, ::-1], rtol=0, atol=1e-16)
    assert np.allclose(p, p.T, rtol=0, atol=1e-16)


def _centres(n, pix):
    c = (np.arange(n) - n / 2 + 0.5) * pix
    return np.hypot(c[None, :], c[:, None])


def test_gaussian_is_sampled_at_pixel_centres():
    """profile.cpp:228-263: sdev = rhalf/1.18, pixel centres at
    (j - n/2 + 1/2) pix (profile.cpp:126-150)."""
    rhalf, pix = 2.7, 0.8
    p = orc.profile("gaussian", rhalf, pix)
    r = _centres(p.shape[0], pix)
    g = np.exp(-r * r / (2 * (rhalf / 1.18) ** 2))
    assert np.max(np.abs(p - g / g.sum())) < 1e-15


def test_uniform_disc_pixel_count():
    """profile.cpp:185-222: value 1 inside R = sqrt(2) rhalf, brute-force count."""
    for rhalf in (1.5, 4.0, 9.3):
        p = orc.profile("uniform", rhalf, 1.0)
        r = _centres(p.shape[0], 1.0)
        inside = r < np.sqrt(2.0) * rhalf
        assert np.count_nonzero(p) == np.count_nonzero(inside)
        assert np.allclose(p[inside], 1.0 / inside.sum(), rtol=1e-15)


@pytest.mark.parametrize("shape,rin,a", [("gaussian", 0.0, 1.0), ("exponential", 0.0, 1.0),
                                         ("gaussian_hole", 1.0, 1.0), ("thermal_hole", 1.0, 1.0),
                                         ("wavy", 0.0, 1.0), ("wavy", 0.0, -0.6)])
def test_profile_derivative_matches_central_difference(shape, rin, a):
    rhalf, h = 3.3, 1e-5
    p, dp = orc.profile(shape, rhalf, 1.0, rin=rin, a=a, deriv=True)
    n = p.shape[0]
    pp = orc.profile(shape, rhalf + h, 1.0, rin=rin, a=a)
    pm = orc.profile(shape, rhalf - h, 1.0, rin=rin, a=a)
    assert pp.shape[0] == n == pm.shape[0]
    fd = (pp - pm) / (2 * h)
    assert np.max(np.abs(fd - dp)) < 1e-6 * np.max(np.abs(dp))


def test_thermal_hole_and_wavy_invariants():
    """Properties the reference's definitions fix (profile.cpp:315-398)
    without restating their formulas: the thermal hole is dark inside
    Rin = rhalf / 4 and bright outside; the wavy profile is dark from
    2 rhalf outward, and its amplitude a drops out of the normalised profile
    (sign included)."""
    for rhalf in (4.0, 8.0, 13.5):
        p = orc.profile("thermal_hole", rhalf, 1.0)
        r = _centres(p.shape[0], 1.0)
        assert np.all(p[r <= rhalf / 4] == 0.0) and np.all(p[r > rhalf / 4] > 0.0)
        w1 = orc.profile("wavy", rhalf, 1.0, a=1.0)
        r = _centres(w1.shape[0], 1.0)
        assert np.all(w1[r >= 2 * rhalf] == 0.0) and np.all(w1[r < 2 * rhalf] >= 0.0)
        for a in (0.3, -0.6, 7.0):
            assert np.allclose(orc.profile("wavy", rhalf, 1.0, a=a), w1, rtol=1e-13, atol=0)


# ---------------------------------------------------------------- FFT and convolution
@pytest.mark.parametrize("L", [1, 2, 3, 5, 8, 12, 30, 64, 100, 127, 243, 1000])
def test_fft1d_matches_numpy(L):
    rng = np.random.default_rng(L)
    x = rng.standard_normal(L) + 1j * rng.standard_normal(L)
    assert np.max(np.abs(orc.fft1d(x) - np.fft.fft(x))) < 1e-12 * max(1.0, np.sqrt(L))
    assert np.max(np.abs(orc.fft1d(x, +1) - L * np.fft.ifft(x))) < 1e-12 * max(1.0, np.sqrt(L))


def test_fft_convolution_equals_direct_and_numpy():
    rng = np.random.default_rng(3)
    m = rng.random((24, 20))
    k = orc.set_kernel(orc.profile("gaussian", 1.7, 1.0), 20, 24)
    a = orc.fft_convolve(m, k)
    b = orc.direct_convolve(m, k)
    c = np.fft.irfft2(np.fft.rfft2(m) * np.fft.rfft2(k), s=m.shape)
    assert np.max(np.abs(a - b)) < 1e-13
    assert np.max(np.abs(a - c)) < 1e-13
    assert abs(a.sum() - m.sum()) < 1e-11  # kernel sums to one


def test_set_kernel_wraps_quadrants_into_corners():
    """magnification_map.cpp:172-194: K[dy, dx] = P((dx+1/2) pix, (dy+1/2) pix)."""
    p = np.arange(36.0).reshape(6, 6)
    k = orc.set_kernel(p, 16, 12)
    h = 3
    assert k.sum() == p.sum()
    assert k[0, 0] == p[h, h] and k[0, 1] == p[h, h + 1] and k[1, 0] == p[h + 1, h]
    assert k[-1, -1] == p[h - 1, h - 1] and k[-1, 0] == p[h - 1, h] and k[0, -1] == p[h, h - 1]
    assert np.count_nonzero(k) == np.count_nonzero(p)


def test_two_pixel_disc_averages_the_corner_neighbourhood():
    """A 2 x 2 uniform kernel: the convolved map is the source centred at the
    pixel corner (+1/2, +1/2), i.e. the mean of pixels (y..y+1, x..x+1)
    (circular)."""
    rng = np.random.default_rng(4)
    m = rng.random((10, 14))
    k = orc.set_kernel(np.full((2, 2), 0.25), 14, 10)
    want = 0.25 * (m + np.roll(m, -1, 0) + np.roll(m, -1, 1) + np.roll(np.roll(m, -1, 0), -1, 1))
    assert np.max(np.abs(orc.fft_convolve(m, k) - want)) < 1e-14


# ---------------------------------------------------------------- locations
def test_locations_on_grid_closed_form_and_reference():
    Noff = int(np.ceil(GRID_NMAP / GRID_SS * GRID_SIZES[-1]))  # driver.cpp:60
    assert Noff == 120
    x, y = orc.locations_on_grid(GRID_NSIDE, GRID_NMAP, Noff)
    col = 120 + 49 + 98 * np.arange(GRID_NSIDE)  # dd = (10000 - 240) // 99 = 98
    assert np.array_equal(x.reshape(GRID_NSIDE, GRID_NSIDE), np.tile(col, (GRID_NSIDE, 1)))
    assert np.array_equal(y.reshape(GRID_NSIDE, GRID_NSIDE), np.tile(col[:, None], (1, GRID_NSIDE)))
    r = _ref()
    rx, ry = np.zeros_like(x), np.zeros_like(y)
    r.ref_locations_on_grid(GRID_NSIDE, GRID_NMAP, Noff, orc._p(rx), orc._p(ry))
    assert np.array_equal(rx, x) and np.array_equal(ry, y)


@pytest.mark.parametrize("seed", [1, 17, 123456])
def test_locations_random_matches_reference(seed):
    x, y = orc.locations_random(5000, 1000, 30, seed)
    assert x.min() >= 30 and x.max() <= 970 and y.min() >= 30 and y.max() <= 970
    r = _ref()
    rx, ry = np.zeros_like(x), np.zeros_like(y)
    r.ref_locations_random(5000, 1000, 30, seed, orc._p(rx), orc._p(ry))
    assert np.array_equal(rx, x) and np.array_equal(ry, y)


# ---------------------------------------------------------------- chi2 set-up
def test_setup_chi2_grid_input_hand_values():
    """cpu_functions.cpp:80-144 on grid_input.json with M = (0.4, 0.5),
    V = 1 (driver.cpp:195-196): indices derived by hand."""
    Dfac = np.sqrt(DS * DLS / DL)
    v = orc.setup_chi2((0.4, 0.5), 1.0, Dfac, GRID_D, GRID_T, GRID_S, GRID_SIZES)
    assert v["Njp"] == 5
    assert list(v["indA"]) == [1, 2, 3, 3, -1]
    assert list(v["indB"]) == [1, 2, 2, 3, 3]
    assert v["const_term"] == 0.0
    assert np.array_equal(v["d"], GRID_D) and np.array_equal(v["s"], GRID_S)
    x = np.array(GRID_SIZES)
    Rfac = 23.92865 * Dfac
    for ind, fac, M in [(v["indA"], v["facA"], 0.4), (v["indB"], v["facB"], 0.5)]:
        pos = np.array(GRID_T) * 8.64 / (Rfac * np.sqrt(M))
        for j, f, p in zip(ind, fac, pos):
            if j >= 0:
                assert 0.0 <= f < 1.0 and abs(x[j] + f * (x[j + 1] - x[j]) - p) < 1e-15


def test_find_indices_brute_force():
    rng = np.random.default_rng(5)
    x = np.sort(rng.random(7))
    t = np.concatenate([rng.uniform(-0.2, 1.2, 300), x])
    got = orc.find_indices(t, x)
    for ti, gi in zip(t, got):
        cand = [j for j in range(len(x) - 1) if x[j] <= ti < x[j + 1]]
        if ti == x[-1]:
            cand = [len(x) - 2]
        assert gi == (cand[0] if cand else -1), (ti, gi)


# ---------------------------------------------------------------- sort and bins
def _keys(rng, n):
    k = np.round(rng.lognormal(0, 2, n), 2)  # many ties
    k[rng.integers(0, n, 20)] = np.inf
    k[rng.integers(0, n, 20)] = np.nan
    k[rng.integers(0, n, 20)] = 0.0
    k[rng.integers(0, n, 20)] = -0.0
    return k


def test_stable_argsort_matches_numpy_stable():
    rng = np.random.default_rng(6)
    k = _keys(rng, 20000)
    assert np.array_equal(orc.stable_argsort(k), np.argsort(k, kind="stable"))


def test_bin_edges_are_upper_bounds():
    rng = np.random.default_rng(7)
    s = np.sort(np.round(rng.lognormal(0, 1, 5000), 3))
    bins = orc.mpd_bins(40, 0.01, 100.0)
    lo, npb = orc.bin_edges(s, bins)
    ub = np.searchsorted(s, bins, side="right")
    assert np.array_equal(lo, np.concatenate([[0], ub[:-1]]))
    assert np.array_equal(npb, np.diff(np.concatenate([[0], ub])))


def test_setup_integral_ranks_the_ratios():
    rng = np.random.default_rng(8)
    A, B = rng.random((40, 3)), rng.random((40, 3))
    A[3, 0] = A[5, 0]  # ties
    bins = orc.mpd_bins(30, 1e-2, 1e2)
    si, lo, npb = orc.setup_integral(A, B, bins)
    z = (A[:, 0][:, None] / B[:, 0][None, :]).ravel()
    assert np.array_equal(orc.zratio(A, B), z)
    perm = np.argsort(z, kind="stable")
    rank = np.empty_like(perm)
    rank[perm] = np.arange(z.size)
    assert np.array_equal(si, rank)
    ub = np.searchsorted(z[perm], bins, side="right")
    assert np.array_equal(npb, np.diff(np.concatenate([[0], ub])))


def test_chi2_pairs_interpolates_linear_light_curves_exactly():
    """LC linear in the size index: the interpolated flux is exact, so every
    pair's chi2 has a closed form; -1 indices use flux 1."""
    Nloc, Nprof = 6, 5
    rng = np.random.default_rng(9)
    a0, a1 = rng.uniform(1, 2, Nloc), rng.uniform(0.1, 0.2, Nloc)
    b0, b1 = rng.uniform(1, 2, Nloc), rng.uniform(0.1, 0.2, Nloc)
    k = np.arange(Nprof)
    LCA = a0[:, None] + a1[:, None] * k
    LCB = b0[:, None] + b1[:, None] * k
    DLCA, DLCB = np.diff(LCA, axis=1), np.diff(LCB, axis=1)
    v = dict(Njp=4, indA=[0, 2, 3, -1], indB=[1, -1, 3, 0], facA=[0.25, 0.5, 0.9, 0.0],
             facB=[0.75, 0.0, 0.1, 0.3], d=[0.9, 1.1, 1.0, 0.8], s=[0.1, 0.2, 0.1, 0.3])
    si = np.arange(Nloc * Nloc, dtype=np.uint32)
    got = orc.chi2_pairs(v, LCA, DLCA, LCB, DLCB, si)
    for a in range(Nloc):
        for b in range(Nloc):
            c = 0.0
            for i in range(4):
                fa = 1.0 if v["indA"][i] < 0 else a0[a] + a1[a] * (v["indA"][i] + v["facA"][i])
                fb = 1.0 if v["indB"][i] < 0 else b0[b] + b1[b] * (v["indB"][i] + v["facB"][i])
                c += ((v["d"][i] - fa / fb) / v["s"][i]) ** 2
            assert abs(got[a * Nloc + b] - c) < 1e-12 * c


def test_bin_chi2_constant_and_empty_bins():
    chi2 = np.full(10, 3.0)
    lo = np.array([0, 4, 4], np.uint32)
    npb = np.array([4, 0, 6], np.uint32)
    bc, be = orc.bin_chi2(chi2, lo, npb, gpu_convention=True)
    assert bc[0] == 3.0 and bc[2] == 3.0 and abs(be[0] - np.exp(-1.5)) < 1e-16
    assert bc[1] == 2000.0 and be[1] == 0.0  # exp(-1000) underflows (gpu_functions.cu:789-792)
    bc, be = orc.bin_chi2(chi2, lo, npb, gpu_convention=False)
    assert bc[1] == 0.0 and be[1] == 0.0  # cpu_functions.cpp:269-281


# ---------------------------------------------------------------- MPDs and integrals
@pytest.mark.parametrize("log_scale", [True, False])
def test_mpd_bins_match_reference(log_scale):
    r = _ref()
    for (N, lo, hi) in [(500, 1e-4, 400.0), (300, 1e-4, 300.0), (7, 0.5, 3.0)]:
        b = np.zeros(N)
        r.ref_mpd_bins(N, lo, hi, 1 if log_scale else 0, orc._p(b))
        assert np.array_equal(orc.mpd_bins(N, lo, hi, log_scale), b)


def test_mpd_normalize_divide_and_integrals_match_reference():
    r = _ref()
    rng = np.random.default_rng(10)
    bA, bB, bR = orc.mpd_bins(50, 0.01, 50.0), orc.mpd_bins(60, 0.01, 60.0), orc.mpd_bins(40, 0.01, 10.0)
    cA, cB = rng.random(50), rng.random(60)
    want = cA.copy()
    r.ref_mpd_normalize(50, orc._p(bA), orc._p(want))
    assert np.array_equal(orc.mpd_normalize(bA, cA), want)
    nA, nB = orc.mpd_normalize(bA, cA), orc.mpd_normalize(bB, cB)
    wR = np.zeros(40)
    r.ref_mpd_divide_by(50, orc._p(bA), orc._p(nA), 60, orc._p(bB), orc._p(nB), 40, orc._p(bR), orc._p(wR))
    assert np.array_equal(orc.mpd_divide_by(bA, nA, bB, nB, bR), wR)
    x = np.sort(rng.random(33))
    f1, f2 = rng.random(33), rng.random(33)
    assert orc.trapezium(f1, f2, x) == r.ref_trapezium(33, orc._p(f1), orc._p(f2), orc._p(x))
    assert orc.simpson(f1, f2, x) == r.ref_simpson(33, orc._p(f1), orc._p(f2), orc._p(x))


def test_trapezium_closed_form():
    x = np.linspace(0.0, 2.0, 101)
    # integral of (1 + x)(2 - x/2) is exact for the trapezium up to the x^2 term error
    f1, f2 = 1 + x, np.ones_like(x)
    assert abs(orc.trapezium(f1, f2, x) - 4.0) < 1e-14


def test_binned_mpd_histogram_brute_force():
    """magnification_map_gpu.cu:45-87: counts on log right-edge bins within
    [min, max), empty bins floored at 1/N^2, normalised by N^2."""
    rng = np.random.default_rng(11)
    m = rng.lognormal(0.5, 1.0, (60, 50))
    N, lo, hi = 25, 0.1, 30.0
    bins, counts = orc.binned_mpd(m, N, lo, hi)
    assert np.array_equal(bins, orc.mpd_bins(N, lo, hi))
    v = np.sort(m.ravel())
    v = v[(v >= lo) & (v < hi)]
    ub = np.searchsorted(v, bins, side="right")
    raw = np.diff(np.concatenate([[0], ub])).astype(float)
    raw[raw == 0] = 1.0
    assert np.allclose(counts, raw / m.size, rtol=1e-15, atol=0)


def test_gerlumph_normalise_closed_form():
    """magnification_map.cpp:7-40: counts * |avgmu/avgN| / |1/((1-k)^2-g^2)|."""
    counts = np.array([[137, 274], [0, 1]], np.int32)
    out = orc.gerlumph_normalise(counts, 2.0, 137.0, 0.4, 0.4)
    mu_th = 1.0 / ((1 - 0.4) ** 2 - 0.4 ** 2)
    assert abs(out[0, 0] - 2.0 / mu_th) < 1e-15
    assert abs(out[0, 1] - 4.0 / mu_th) < 1e-15 and out[1, 0] == 0.0


# ---------------------------------------------------------------- chi2_driver
def test_multiband_known_answer():
    """chi2_driver.cu:35-56 constant fake data: chi2 = 70 * ((0.97 - 0.42/1.1)/1.1)^2 / 2,
    float bits in the GPU (per-filter partials) and CPU (one running sum)
    orders; the float likelihood stagnates at 1024, the exact one is 4508.1."""
    N, F, T = 10000, 7, 10
    f_obs = np.full((F, T), 0.97, np.float32)
    df = np.full((F, T), 1.1, np.float32)
    mu1 = np.full((N, F, T), 0.42, np.float32)
    mu2 = np.full((N, F, T), 1.1, np.float32)
    exact = 35 * ((0.97 - 0.42 / 1.1) / 1.1) ** 2
    small = slice(0, 50)
    c_gpu, _, _ = orc.chi2_multiband(f_obs, df, mu1[small], mu2[small], gpu_quirk=True)
    c_cpu, _, _ = orc.chi2_multiband(f_obs, df, mu1[small], mu2[small], gpu_quirk=False)
    assert np.all(c_gpu.view(np.uint32) == 0x41201CDC)
    assert np.all(c_cpu.view(np.uint32) == 0x41201CD6)
    assert abs(float(c_gpu[0]) - exact) < 1e-5 and abs(float(c_cpu[0]) - exact) < 1e-5
    _, l32, l64 = orc.chi2_multiband(f_obs, df, mu1, mu2, gpu_quirk=False, want_all=False)
    assert l32 == 1024.0
    assert abs(l64 - N * N * np.exp(-float(c_cpu[0]))) < 1e-9 * l64
    assert abs(l64 - 4508.1) < 0.1


# ---------------------------------------------------------------- light curves (magnitude chi2 model)
def test_bilinear_point_samples_and_affine_exactness():
    rng = np.random.default_rng(12)
    m = rng.random((9, 11))
    for _ in range(50):
        x, y = int(rng.integers(0, 11)), int(rng.integers(0, 9))
        assert orc.bilinear(m, float(x), float(y)) == m[y, x]
    yy, xx = np.mgrid[0:9, 0:11]
    aff = 0.3 + 1.7 * xx - 0.6 * yy
    for _ in range(50):
        x, y = rng.uniform(0, 10), rng.uniform(0, 8)
        assert abs(orc.bilinear(aff, x, y) - (0.3 + 1.7 * x - 0.6 * y)) < 1e-13


def _obs(rng, B, E):
    t = [np.sort(rng.uniform(0, 40, E)) for _ in range(B)]
    return t, [rng.normal(-1.5, 0.2, E) for _ in range(B)], [rng.uniform(0.03, 0.08, E) for _ in range(B)]


def _run(maps, B, off, cnt, rad, bs, prm, t, mo, sg):
    obs_off = np.cumsum([0] + [len(a) for a in t]).astype(np.int32)
    n = prm.shape[0]
    return orc.lc_chi2_grad(maps, B, off, cnt, rad, None, bs, prm, np.zeros(n, np.int32),
                            np.zeros(n, np.int32), obs_off, np.concatenate(t), np.concatenate(mo),
                            np.concatenate(sg))


def test_lc_chi2_constant_map_closed_form():
    """mu constant, dmu/dr = 0: chi2 = sum ((mobs - dm + 2.5 log10 mu)/sig)^2,
    only the dm gradient is non-zero."""
    rng = np.random.default_rng(13)
    Ny, Nx, mu0 = 40, 48, 3.7
    maps = np.zeros((2, Ny, Nx, 2))
    maps[..., 0] = mu0
    t, mo, sg = _obs(rng, 2, 30)
    prm = np.array([[10.0, 12.0, 0.4, 0.5, 2.0, 0.01, 0.13]])
    chi2, grad = _run(maps, 2, [0, 1], [1, 1], [2.0, 4.0], [1.0, 2.0], prm, t, mo, sg)
    m = 0.13 - 2.5 * np.log10(mu0)
    rho = [(a - m) / s for a, s in zip(mo, sg)]
    want = sum((r * r).sum() for r in rho)
    assert abs(chi2[0] - want) < 1e-12 * want
    assert np.all(grad[0, :6] == 0.0)
    assert abs(grad[0, 6] - sum((-2 * r / s).sum() for r, s in zip(rho, sg))) < 1e-10 * abs(grad[0, 6])


def test_lc_chi2_hermite_reproduces_cubic_radius_law():
    """Node maps uniform in space with (f(r_k), f'(r_k)) of a cubic f: the
    Hermite interpolant is exact between nodes, the Taylor step outside."""
    f = lambda r: 2.0 + 0.3 * r - 0.05 * r * r + 0.002 * r ** 3
    df = lambda r: 0.3 - 0.1 * r + 0.006 * r * r
    rad = np.array([1.0, 2.5, 4.0, 7.0])
    maps = np.zeros((4, 16, 16, 2))
    for k, r in enumerate(rad):
        maps[k, ..., 0], maps[k, ..., 1] = f(r), df(r)
    t = [np.linspace(0.0, 30.0, 61)]
    mo, sg = [np.full(61, -0.7)], [np.full(61, 0.05)]
    r0, w = 0.5, 0.25  # r from 0.5 (below the first node) to 8 (above the last)
    prm = np.array([[5.0, 6.0, 0.0, 0.0, r0, w, 0.0]])
    chi2, grad = _run(maps, 1, [0], [4], rad, [1.0], prm, t, mo, sg)
    r = r0 + w * t[0]
    mu = np.where(r <= rad[0], f(rad[0]) + df(rad[0]) * (r - rad[0]),
                  np.where(r >= rad[-1], f(rad[-1]) + df(rad[-1]) * (r - rad[-1]), f(r)))
    dmu = np.where(r <= rad[0], df(rad[0]), np.where(r >= rad[-1], df(rad[-1]), df(r)))
    rho = (mo[0] + 2.5 * np.log10(mu)) / sg[0]
    assert abs(chi2[0] - (rho * rho).sum()) < 1e-11 * chi2[0]
    dc = 2 * rho * 2.5 / (np.log(10) * mu * sg[0])
    assert abs(grad[0, 4] - (dc * dmu).sum()) < 1e-9 * abs(grad[0, 4])
    assert abs(grad[0, 5] - (dc * dmu * t[0]).sum()) < 1e-9 * abs(grad[0, 5])


def test_lc_chi2_gradient_matches_central_differences():
    rng = np.random.default_rng(14)
    Ny, Nx, B = 64, 72, 2
    base = np.exp(0.3 * rng.standard_normal((Ny, Nx)))
    smooth = np.fft.irfft2(np.fft.rfft2(base) * np.exp(-0.5 * (np.fft.fftfreq(Ny)[:, None] ** 2
                                                              + np.fft.rfftfreq(Nx)[None, :] ** 2)
                                                       * (2 * np.pi * 2.0) ** 2), s=base.shape) + 2.0
    rad = [1.5, 3.0, 2.0, 5.0]
    maps = np.zeros((4, Ny, Nx, 2))
    for k, rk in enumerate(rad):
        maps[k, ..., 0] = smooth * (1 + 0.05 * rk)
        maps[k, ..., 1] = 0.05 * smooth + 0.01 * np.roll(smooth, k + 1, 1)
    t, mo, sg = _obs(rng, B, 25)
    prm = np.array([[20.3, 25.7, 0.7, 0.6, 2.1, 0.02, 0.05],
                    [40.1, 30.2, 2.9, 0.45, 2.4, 0.01, -0.1]])
    args = (maps, B, [0, 2], [2, 2], rad, [1.0, 1.6])
    chi2, grad = _run(*args, prm, t, mo, sg)
    for q in range(7):
        h = 1e-6 * max(1.0, abs(prm[0, q]))
        pp, pm = prm.copy(), prm.copy()
        pp[:, q] += h
        pm[:, q] -= h
        fd = (_run(*args, pp, t, mo, sg)[0] - _run(*args, pm, t, mo, sg)[0]) / (2 * h)
        assert np.all(np.abs(fd - grad[:, q]) < 2e-5 * np.maximum(1.0, np.abs(grad[:, q]))), (q, fd, grad[:, q])


def test_lc_chi2_leaving_the_map_is_flagged():
    maps = np.ones((1, 16, 16, 2))
    t, mo, sg = [np.linspace(0, 10, 11)], [np.zeros(11)], [np.full(11, 0.1)]
    prm = np.array([[14.0, 8.0, 0.0, 0.5, 2.0, 0.0, 0.0],   # x reaches 19 > 15
                    [8.0, 8.0, 0.0, 0.5, 2.0, 0.0, 0.0]])
    chi2, grad = _run(maps, 1, [0], [1], [2.0], [1.0], prm, t, mo, sg)
    assert chi2[0] == np.inf and np.all(grad[0] == 0.0)
    assert np.isfinite(chi2[1]) and chi2[1] == 0.0


def test_expanding_source_composes_profile_kernel_and_sampling():
    rng = np.random.default_rng(16)
    N, width = 64, 4.0
    m = rng.random((N, N)) + 1.0
    lx, ly = orc.locations_on_grid(5, N, 8)
    LC, DLC = orc.expanding_source(m, width, [0.0, 0.15], "gaussian", lx, ly)
    assert np.array_equal(LC[:, 0], m[ly, lx])
    pix = orc.parity_round(width / N)
    k = orc.set_kernel(orc.profile("gaussian", orc.parity_round(0.15), pix), N, N)
    conv = np.fft.irfft2(np.fft.rfft2(m) * np.fft.rfft2(k), s=m.shape)
    assert np.max(np.abs(LC[:, 1] - conv[ly, lx])) < 1e-12
    assert np.array_equal(DLC[:, 0], LC[:, 1] - LC[:, 0])
