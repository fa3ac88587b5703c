"""GPU parity of the reference's pair mode, sort/binning, MPD histogram,
sampling and the chi2_driver kernel against the FP64 / float oracle.

Integer outputs (orderings, ranks, bin tables, counts) must be identical;
FP64 values computed with the same operation order (z, pairwise chi2,
GERLUMPH normalisation, sampling) must be bit-identical; sums whose order
differs (bin means, likelihood) agree to a few ulps."""
import numpy as np
import pytest
import torch

from microsn_amd import api

pytestmark = pytest.mark.gpu


def _cuda(a, dt=None):
    return torch.tensor(a, device="cuda") if dt is None else torch.tensor(a, dtype=dt, device="cuda")


def test_gerlumph_normalise_bitexact(ctx, orc):
    rng = np.random.default_rng(0)
    counts = rng.integers(0, 5000, size=(300, 300), dtype=np.int32)
    out = api.gerlumph_normalise(ctx, _cuda(counts), 0.915, 137.3, 0.4, 0.35)
    ref = orc.gerlumph_normalise(counts, 0.915, 137.3, 0.4, 0.35)
    assert np.array_equal(out.cpu().numpy(), ref)


def test_sample_points_bitexact(ctx, orc):
    rng = np.random.default_rng(1)
    conv = rng.random((50, 70))
    lx = rng.integers(0, 70, 333).astype(np.int32)
    ly = rng.integers(0, 50, 333).astype(np.int32)
    LC = api.SimLC(ctx, 333, 3)
    for k in range(3):
        api.sample_points(ctx, _cuda(conv * (k + 1)), k, _cuda(lx), _cuda(ly), LC)
    ref_LC = np.stack([conv[ly, lx] * (k + 1) for k in range(3)], 1)
    assert np.array_equal(LC.LC.cpu().numpy(), ref_LC)
    assert np.array_equal(LC.DLC.cpu().numpy(), np.diff(ref_LC, axis=1))


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_expanding_source_matches_oracle(ctx, orc, prec):
    rng = np.random.default_rng(2)
    N, width = 200, 8.0
    m = 5.0 * np.exp(0.5 * rng.standard_normal((N, N)))
    sizes = [0.0, 0.05, 0.1, 0.2]
    lx, ly = orc.locations_on_grid(9, N, 30)
    LC = api.SimLC(ctx, lx.size, len(sizes))
    api.expanding_source(ctx, _cuda(m), width, sizes, "gaussian", _cuda(lx), _cuda(ly), LC, prec)
    rLC, rDLC = orc.expanding_source(m, width, sizes, "gaussian", lx, ly)
    got = LC.LC.cpu().numpy()
    assert np.array_equal(got[:, 0], rLC[:, 0])  # point source: raw map
    tol = 1e-12 if prec == "f64" else 3e-6
    assert np.max(np.abs(got - rLC) / np.abs(rLC)) < tol
    assert np.max(np.abs(LC.DLC.cpu().numpy() - rDLC)) < tol * np.max(np.abs(rLC))


def _special_keys(rng, n):
    k = rng.choice(np.array([0.5, 1.0, 2.0, -1.0, 0.0, -0.0, np.inf, -np.inf, np.nan, 3.25]), n)
    r = rng.random(n) < 0.5
    k[r] = np.round(rng.standard_normal(r.sum()) * 4) / 4  # many ties
    return k


@pytest.mark.parametrize("n", [1, 2, 5, 1023, 1024, 1025, 4095, 4096, 4097, 100_000, 1_000_003,
                               2_097_152, 2_097_153])
def test_sort_matches_stable_sort(ctx, orc, n):
    rng = np.random.default_rng(n)
    keys = _special_keys(rng, n) if n < 100_000 else rng.standard_normal(n) * 10.0 ** rng.integers(-3, 3, n)
    if n >= 100_000:
        keys[rng.random(n) < 0.01] = 1.0  # ties
        keys[:3] = [np.nan, -0.0, 0.0]
    perm, rank, sk = api.sort_f64(ctx, _cuda(keys), rank=True, sorted_keys=True)
    ref = orc.stable_argsort(keys)
    p = perm.cpu().numpy().view(np.uint32)
    assert np.array_equal(p, ref)
    r = rank.cpu().numpy().view(np.uint32)
    assert np.array_equal(r[ref], np.arange(n, dtype=np.uint32))
    assert np.array_equal(sk.cpu().numpy().view(np.uint64), keys[ref].view(np.uint64))


@pytest.mark.parametrize("kind", ["all_equal", "small_ints", "negative_zero"])
def test_sort_trivial_passes(ctx, orc, kind):
    """Keys whose digits are constant in some or all passes (those passes copy)."""
    n = 70_001
    rng = np.random.default_rng(3)
    if kind == "all_equal":
        keys = np.full(n, 2.5)
    elif kind == "small_ints":
        keys = rng.integers(0, 7, n).astype(np.float64)
    else:
        keys = np.where(rng.random(n) < 0.5, 0.0, -0.0)
    perm, rank, sk = api.sort_f64(ctx, _cuda(keys), rank=True, sorted_keys=True)
    ref = orc.stable_argsort(keys)
    assert np.array_equal(perm.cpu().numpy().view(np.uint32), ref)
    assert np.array_equal(sk.cpu().numpy().view(np.uint64), keys[ref].view(np.uint64))


def test_sort_large_properties(ctx):
    """2^25 + 7 keys: a sorted, stable permutation (checked on the device)."""
    n = (1 << 25) + 7
    g = torch.Generator(device="cuda").manual_seed(5)
    keys = torch.floor(torch.rand(n, device="cuda", dtype=torch.float64, generator=g) * 1000.0)
    perm, rank, sk = api.sort_f64(ctx, keys, rank=True, sorted_keys=True)
    p = perm.long()
    assert torch.all(sk[1:] >= sk[:-1])
    assert torch.equal(sk, keys[p])
    # stability: equal keys keep increasing original indices
    same = sk[1:] == sk[:-1]
    assert torch.all(p[1:][same] > p[:-1][same])
    # rank is the inverse permutation
    assert torch.equal(rank.long()[p], torch.arange(n, device="cuda"))


def test_bin_edges_match_upper_bound(ctx, orc):
    rng = np.random.default_rng(7)
    keys = np.sort(np.round(rng.lognormal(0, 1, 50_000), 2))
    keys = np.concatenate([keys, [np.nan] * 5])
    bins = orc.mpd_bins(300, 1e-4, 300, True)
    bins[10] = keys[1000]  # an edge equal to keys
    lo, npb = api.bin_edges(ctx, _cuda(keys), bins)
    rlo, rnpb = orc.bin_edges(keys, bins)
    assert np.array_equal(lo, rlo) and np.array_equal(npb, rnpb)


@pytest.mark.parametrize("n", [0, 7, 50_000, 300_000])
def test_bin_table_device_equals_host_path(ctx, orc, n):
    """mlc_bin_table_device against mlc_bin_edges + mlc_bin_chi2 bit for bit:
    empty bins, an edge equal to keys, bins of more than one 4096-key chunk,
    +inf and NaN keys past the last edge, and no keys at all."""
    rng = np.random.default_rng(n + 5)
    keys = np.sort(np.round(rng.lognormal(3, 1.5, n), 3))
    if n:
        keys[-3:] = np.inf
        keys = np.concatenate([keys, [np.nan] * 2])
    bins = orc.mpd_bins(60, 1e-2, 1e4, True)
    if n > 1000:  # an edge equal to a key, the edges still ascending
        k = int(np.searchsorted(bins, keys[1000]))
        bins[k] = keys[1000]
    assert np.all(np.diff(bins) >= 0)
    d = _cuda(keys)
    lo, npb, bc, be = api.bin_table_device(ctx, d, bins)
    hlo, hnpb = api.bin_edges(ctx, d, bins)
    assert np.array_equal(lo.cpu().numpy().view(np.uint32), hlo)
    assert np.array_equal(npb.cpu().numpy().view(np.uint32), hnpb)
    sb = api.Chi2SortBins(ctx, 1, bins.size)
    sb.lower_ind[:], sb.n_per_bin[:] = hlo, hnpb
    hbc, hbe = api.bin_chi2_GPU(ctx, sb, d)
    assert np.array_equal(bc.cpu().numpy(), hbc) and np.array_equal(be.cpu().numpy(), hbe)
    if n == 300_000:
        assert hnpb.max() > 4096  # a bin of several 4096-key chunks
    # descending edges are rejected before any launch
    with pytest.raises(Exception, match="non-decreasing"):
        api.bin_table_device(ctx, d, bins[::-1].copy())


def _pair_case(orc, Nloc, Nprof, rng):
    LCA = 0.5 + rng.random((Nloc, Nprof)) * 4
    LCB = 0.5 + rng.random((Nloc, Nprof)) * 4
    LCA[rng.random(Nloc) < 0.1, 0] = 2.0  # z ties
    LCB[rng.random(Nloc) < 0.1, 0] = 1.0
    return LCA, np.diff(LCA, axis=1), LCB, np.diff(LCB, axis=1)


@pytest.mark.parametrize("Nloc", [1, 37, 300])
def test_pair_mode_matches_oracle(ctx, orc, Nloc):
    rng = np.random.default_rng(Nloc)
    Nprof = 5
    LCA, DLCA, LCB, DLCB = _pair_case(orc, Nloc, Nprof, rng)
    A, B = api.SimLC(ctx, Nloc, Nprof), api.SimLC(ctx, Nloc, Nprof)
    A.LC.copy_(_cuda(LCA)); A.DLC.copy_(_cuda(DLCA))
    B.LC.copy_(_cuda(LCB)); B.DLC.copy_(_cuda(DLCB))
    bins = orc.mpd_bins(300, 1e-4, 300, True)
    sb = api.Chi2SortBins(ctx, Nloc, bins.size)
    api.setup_integral_GPU(ctx, A, B, bins, sb)
    rsi, rlo, rnpb = orc.setup_integral(LCA, LCB, bins)
    assert np.array_equal(sb.d_sorted_ind.cpu().numpy().view(np.uint32), rsi)
    assert np.array_equal(sb.lower_ind, rlo) and np.array_equal(sb.n_per_bin, rnpb)
    x = np.array([0.0, 0.05, 0.1, 0.2, 0.3])
    v = orc.setup_chi2([0.4, 0.5], 1.0, np.sqrt(1750 * 850 / 1570), [0.7, 1.2, 1.4, 1.1, 0.9],
                       [5, 7, 12, 14, 17], [0.1, 0.15, 0.1, 0.1, 0.15], x)
    chi2 = torch.empty(Nloc * Nloc, dtype=torch.float64, device="cuda")
    api.calculate_chi2_GPU(ctx, v, chi2, sb, A, B)
    ref = orc.chi2_pairs(v, LCA, DLCA, LCB, DLCB, rsi)
    assert np.array_equal(chi2.cpu().numpy(), ref)  # same operations, same order
    bc, be = api.bin_chi2_GPU(ctx, sb, chi2)
    rbc, rbe = orc.bin_chi2(ref, rlo, rnpb, gpu_convention=True)
    assert np.allclose(bc, rbc, rtol=1e-13, atol=0) and np.allclose(be, rbe, rtol=1e-12, atol=1e-300)


def test_bin_chi2_rejects_ranges_outside(ctx):
    sb = api.Chi2SortBins(ctx, 3, 2)
    sb.lower_ind[:] = [0, 5]
    sb.n_per_bin[:] = [5, 5]
    chi2 = torch.zeros(9, dtype=torch.float64, device="cuda")
    with pytest.raises(Exception, match="outside"):
        api.bin_chi2_GPU(ctx, sb, chi2)


def test_binned_mpd_counts_exact(ctx, orc):
    rng = np.random.default_rng(9)
    m = rng.lognormal(1.0, 1.2, (512, 512))
    m[0, :10] = [0.02, 200.0, 0.019, 199.9, 1e-9, 1e9, np.nextafter(0.02, 0), 5.0, 5.0, 5.0]
    bins, counts = api.binned_mpd(ctx, _cuda(m), 500, 0.02, 200.0)
    rbins, rcounts = orc.binned_mpd(m, 500, 0.02, 200.0)
    assert np.array_equal(bins, rbins)
    assert np.array_equal(counts, rcounts)


def test_multiband_known_answer(ctx, orc):
    """chi2_driver's constant fake data (chi2_driver.cu:35-56), N reduced."""
    N, F, T = 300, 7, 10
    f_obs = np.full((F, T), 0.97, np.float32)
    df = np.full((F, T), 1.1, np.float32)
    mu1 = np.full((N, F, T), 0.42, np.float32)
    mu2 = np.full((N, F, T), 1.1, np.float32)
    for mode, bits in [(1, 0x41201CDC), (0, 0x41201CD6)]:
        c, like = api.getChi2GPU(ctx, _cuda(f_obs), _cuda(df), _cuda(mu1), _cuda(mu2), mode)
        c = c.cpu().numpy()
        assert np.all(c.view(np.uint32) == bits)
        assert abs(like - N * N * np.exp(-np.float64(c[0]))) < 1e-9 * like


@pytest.mark.parametrize("mode", [0, 1])
def test_multiband_random_bitexact(ctx, orc, mode):
    rng = np.random.default_rng(mode)
    N, F, T = 133, 3, 45  # F*T not a multiple of the 64-term staging chunk
    f_obs = rng.random((F, T)).astype(np.float32)
    df = (0.5 + rng.random((F, T))).astype(np.float32)
    mu1 = (0.2 + rng.random((N, F, T))).astype(np.float32)
    mu2 = (0.2 + rng.random((N, F, T))).astype(np.float32)
    c, like = api.getChi2GPU(ctx, _cuda(f_obs), _cuda(df), _cuda(mu1), _cuda(mu2), mode)
    rc, _, l64 = orc.chi2_multiband(f_obs, df, mu1, mu2, gpu_quirk=bool(mode))
    assert np.array_equal(c.cpu().numpy().view(np.uint32), rc.view(np.uint32))
    assert abs(like - l64) < 1e-12 * l64
