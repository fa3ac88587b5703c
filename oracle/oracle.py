"""ctypes wrapper over oracle/liborc.so (FP64 CPU oracle) and, when built,
oracle/_ref/libmlcref.so (the reference's own locations.cpp, mpd.cpp and
integration.cpp compiled from the reference tree).

TEST INFRASTRUCTURE ONLY -- see oracle/__init__.py.  Every wrapper names the
C function it calls; those cite the reference file:line they restate.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")
REF_PATH = os.path.join(HERE, "_ref", "libmlcref.so")

SHAPES = {"uniform": 0, "gaussian": 1, "gaussian_hole": 2, "thermal_hole": 3,
          "wavy": 4, "exponential": 5}

_lib = None
_ref = None


def build():
    """Compile liborc.so (and _ref when the reference tree is present)."""
    subprocess.check_call(["make", "-s", "-C", HERE, "all"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def ref():
    """The compiled reference pieces, or None where they were not built."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        _ref = C.CDLL(REF_PATH)
        d, i, p = C.c_double, C.c_int, C.c_void_p
        _ref.ref_trapezium.restype = d
        _ref.ref_simpson.restype = d
        _ref.ref_mpd_bins.argtypes = [i, d, d, i, p]
        _ref.ref_trapezium.argtypes = [i, p, p, p]
        _ref.ref_simpson.argtypes = [i, p, p, p]
    return _ref


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _declare(L):
    d, i, p, i64 = C.c_double, C.c_int, C.c_void_p, C.c_int64
    L.orc_parity_round.restype = d
    L.orc_parity_round.argtypes = [d]
    L.orc_profile_params.argtypes = [i, d, d, d, d, p]
    L.orc_profile_values_deriv.argtypes = [i, d, d, d, d, p, p, i]
    L.orc_set_kernel.argtypes = [p, i, i, i, p]
    L.orc_fft1d.argtypes = [p, i, i]
    L.orc_fft_convolve.argtypes = [p, p, i, i, p]
    L.orc_direct_convolve.argtypes = [p, p, i, i, p]
    L.orc_locations_on_grid.argtypes = [i, i, i, p, p]
    L.orc_locations_random.argtypes = [i, i, i, i, p, p]
    L.orc_sample.argtypes = [p, i, i, i, i, p, p, p, p]
    L.orc_expanding_source.argtypes = [p, i, i, d, p, i, i, i, p, p, p, p]
    L.orc_find_indices.argtypes = [p, p, i, i, p]
    L.orc_setup_chi2.argtypes = [d, d, d, d, i, p, p, p, i, p, i, p, p, p, p, p, p, p]
    L.orc_z.argtypes = [p, p, i, i, p]
    L.orc_stable_argsort.argtypes = [p, i64, p]
    L.orc_bin_edges.argtypes = [p, i64, p, i, p, p]
    L.orc_setup_integral.argtypes = [p, p, i, i, p, i, p, p, p]
    L.orc_chi2_pairs.argtypes = [i, p, p, p, p, p, p, p, p, p, p, i, i, p, p]
    L.orc_bin_chi2.argtypes = [p, p, p, i, i, p, p]
    L.orc_mpd_bins.argtypes = [i, d, d, i, p]
    L.orc_trapezium.restype = d
    L.orc_trapezium.argtypes = [i, p, p, p]
    L.orc_simpson.restype = d
    L.orc_simpson.argtypes = [i, p, p, p]
    L.orc_mpd_normalize.argtypes = [i, p, p]
    L.orc_mpd_divide_by.argtypes = [i, p, p, i, p, p, i, p, p]
    L.orc_binned_mpd.argtypes = [p, i64, i, d, d, p, p]
    L.orc_gerlumph_normalise.argtypes = [p, i64, d, d, d, d, p]
    L.orc_chi2_multiband.argtypes = [p, p, p, p, i, i, i, i, p, p, p]
    L.orc_lc_chi2_grad.argtypes = [p, i, i, i, p, p, p, p, p, i, p, p, p, p, p, p, p, p, p]
    L.orc_bilinear.restype = d
    L.orc_bilinear.argtypes = [p, i, i, d, d]


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ---------------------------------------------------------------- profiles
def parity_round(x):
    return lib().orc_parity_round(float(x))


def profile_params(shape, rhalf, pix, rin=0.0, a=1.0):
    out = np.zeros(2)
    n = lib().orc_profile_params(SHAPES[shape], rhalf, pix, rin, a, _p(out))
    return n, out


def profile(shape, rhalf, pix=1.0, rin=0.0, a=1.0, deriv=False):
    n, _ = profile_params(shape, rhalf, pix, rin, a)
    v = np.zeros((n, n))
    dv = np.zeros((n, n)) if deriv else None
    rc = lib().orc_profile_values_deriv(SHAPES[shape], rhalf, pix, rin, a, _p(v),
                                        _p(dv) if deriv else None, n)
    assert rc == n, rc
    return (v, dv) if deriv else v


def set_kernel(prof, Nx, Ny):
    prof = f64(prof)
    k = np.zeros((Ny, Nx))
    lib().orc_set_kernel(_p(prof), prof.shape[0], Nx, Ny, _p(k))
    return k


def fft1d(x, sign=-1):
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().orc_fft1d(_p(a), a.size, sign)
    return a


def fft_convolve(m, k):
    m, k = f64(m), f64(k)
    out = np.zeros_like(m)
    lib().orc_fft_convolve(_p(m), _p(k), m.shape[1], m.shape[0], _p(out))
    return out


def direct_convolve(m, k):
    m, k = f64(m), f64(k)
    out = np.zeros_like(m)
    lib().orc_direct_convolve(_p(m), _p(k), m.shape[1], m.shape[0], _p(out))
    return out


# ---------------------------------------------------------------- locations
def locations_on_grid(Nside, Nmap, Noff):
    x = np.zeros(Nside * Nside, np.int32)
    y = np.zeros_like(x)
    lib().orc_locations_on_grid(Nside, Nmap, Noff, _p(x), _p(y))
    return x, y


def locations_random(Nloc, Nmap, Noff, seed):
    x = np.zeros(Nloc, np.int32)
    y = np.zeros_like(x)
    lib().orc_locations_random(Nloc, Nmap, Noff, seed, _p(x), _p(y))
    return x, y


def expanding_source(m, width, sizes, shape, lx, ly):
    m, sizes, lx, ly = f64(m), f64(sizes), i32(lx), i32(ly)
    Nloc, Nprof = lx.size, sizes.size
    LC = np.zeros((Nloc, Nprof))
    DLC = np.zeros((Nloc, max(Nprof - 1, 1)))
    rc = lib().orc_expanding_source(_p(m), m.shape[1], m.shape[0], width, _p(sizes), Nprof,
                                    SHAPES[shape], Nloc, _p(lx), _p(ly), _p(LC), _p(DLC))
    assert rc == 0, rc
    return LC, DLC


# ---------------------------------------------------------------- chi2 setup
def find_indices(t, x):
    t, x = f64(t), f64(x)
    ind = np.zeros(t.size, np.int32)
    lib().orc_find_indices(_p(t), _p(x), t.size, x.size, _p(ind))
    return ind


def setup_chi2(M, V, Dfac, d, t, s, x, quirk_const_term=True):
    d, t, s, x = f64(d), f64(t), f64(s), f64(x)
    Nd = d.size
    indA = np.zeros(Nd, np.int32)
    indB = np.zeros(Nd, np.int32)
    facA, facB, nd, ns = (np.zeros(Nd) for _ in range(4))
    ct = np.zeros(1)
    Njp = lib().orc_setup_chi2(M[0], M[1], V, Dfac, Nd, _p(d), _p(t), _p(s), x.size, _p(x),
                               1 if quirk_const_term else 0, _p(indA), _p(indB), _p(facA),
                               _p(facB), _p(nd), _p(ns), _p(ct))
    return dict(Njp=Njp, indA=indA[:Njp], indB=indB[:Njp], facA=facA[:Njp], facB=facB[:Njp],
                d=nd[:Njp], s=ns[:Njp], const_term=float(ct[0]))


# ---------------------------------------------------------------- pair mode
def zratio(LCA, LCB):
    LCA, LCB = f64(LCA), f64(LCB)
    Nloc, Nprof = LCA.shape
    z = np.zeros(Nloc * Nloc)
    lib().orc_z(_p(LCA), _p(LCB), Nloc, Nprof, _p(z))
    return z


def stable_argsort(keys):
    keys = f64(keys)
    perm = np.zeros(keys.size, np.uint32)
    lib().orc_stable_argsort(_p(keys), keys.size, _p(perm))
    return perm


def bin_edges(sorted_keys, bins):
    sorted_keys, bins = f64(sorted_keys), f64(bins)
    lo = np.zeros(bins.size, np.uint32)
    npb = np.zeros(bins.size, np.uint32)
    lib().orc_bin_edges(_p(sorted_keys), sorted_keys.size, _p(bins), bins.size, _p(lo), _p(npb))
    return lo, npb


def setup_integral(LCA, LCB, bins):
    LCA, LCB, bins = f64(LCA), f64(LCB), f64(bins)
    Nloc, Nprof = LCA.shape
    si = np.zeros(Nloc * Nloc, np.uint32)
    lo = np.zeros(bins.size, np.uint32)
    npb = np.zeros(bins.size, np.uint32)
    lib().orc_setup_integral(_p(LCA), _p(LCB), Nloc, Nprof, _p(bins), bins.size, _p(si),
                             _p(lo), _p(npb))
    return si, lo, npb


def chi2_pairs(v, LCA, DLCA, LCB, DLCB, sorted_ind):
    LCA, DLCA, LCB, DLCB = f64(LCA), f64(DLCA), f64(LCB), f64(DLCB)
    Nloc, Nprof = LCA.shape
    out = np.zeros(Nloc * Nloc)
    si = np.ascontiguousarray(sorted_ind, dtype=np.uint32)
    iA, iB = i32(v["indA"]), i32(v["indB"])
    fA, fB, d, s = f64(v["facA"]), f64(v["facB"]), f64(v["d"]), f64(v["s"])
    lib().orc_chi2_pairs(v["Njp"], _p(iA), _p(iB), _p(fA), _p(fB), _p(d), _p(s), _p(LCA),
                         _p(DLCA), _p(LCB), _p(DLCB), Nloc, Nprof, _p(si), _p(out))
    return out


def bin_chi2(chi2, lower_ind, n_per_bin, gpu_convention=True):
    chi2 = f64(chi2)
    lo = np.ascontiguousarray(lower_ind, dtype=np.uint32)
    npb = np.ascontiguousarray(n_per_bin, dtype=np.uint32)
    bc = np.zeros(lo.size)
    be = np.zeros(lo.size)
    lib().orc_bin_chi2(_p(chi2), _p(lo), _p(npb), lo.size, 1 if gpu_convention else 0, _p(bc),
                       _p(be))
    return bc, be


# ---------------------------------------------------------------- MPDs
def mpd_bins(Nbins, lo, hi, log_scale=True):
    b = np.zeros(Nbins)
    lib().orc_mpd_bins(Nbins, lo, hi, 1 if log_scale else 0, _p(b))
    return b


def trapezium(f1, f2, x):
    f1, f2, x = f64(f1), f64(f2), f64(x)
    return lib().orc_trapezium(x.size, _p(f1), _p(f2), _p(x))


def simpson(f1, f2, x):
    f1, f2, x = f64(f1), f64(f2), f64(x)
    return lib().orc_simpson(x.size, _p(f1), _p(f2), _p(x))


def mpd_normalize(bins, counts):
    bins = f64(bins)
    c = f64(counts).copy()
    lib().orc_mpd_normalize(bins.size, _p(bins), _p(c))
    return c


def mpd_divide_by(bA, cA, bB, cB, bR):
    bA, cA, bB, cB, bR = f64(bA), f64(cA), f64(bB), f64(cB), f64(bR)
    out = np.zeros(bR.size)
    lib().orc_mpd_divide_by(bA.size, _p(bA), _p(cA), bB.size, _p(bB), _p(cB), bR.size, _p(bR),
                            _p(out))
    return out


def binned_mpd(m, Nbins, lo, hi):
    m = f64(m)
    bins = np.zeros(Nbins)
    counts = np.zeros(Nbins)
    lib().orc_binned_mpd(_p(m), m.size, Nbins, lo, hi, _p(bins), _p(counts))
    return bins, counts


def gerlumph_normalise(counts, avgmu, avgN, k, g):
    c = np.ascontiguousarray(counts, dtype=np.int32)
    out = np.zeros(c.shape)
    lib().orc_gerlumph_normalise(_p(c), c.size, avgmu, avgN, k, g, _p(out))
    return out


# ---------------------------------------------------------------- chi2_driver
def chi2_multiband(f_obs, df_obs, mu1, mu2, gpu_quirk=False, want_all=True):
    f_obs = np.ascontiguousarray(f_obs, np.float32)
    df_obs = np.ascontiguousarray(df_obs, np.float32)
    mu1 = np.ascontiguousarray(mu1, np.float32)
    mu2 = np.ascontiguousarray(mu2, np.float32)
    N, F, T = mu1.shape
    chi2_all = np.zeros(N * N, np.float32) if (want_all or gpu_quirk) else None
    l32 = np.zeros(1, np.float32)
    l64 = np.zeros(1)
    lib().orc_chi2_multiband(_p(f_obs), _p(df_obs), _p(mu1), _p(mu2), N, F, T,
                             1 if gpu_quirk else 0, _p(chi2_all) if chi2_all is not None else None,
                             _p(l32), _p(l64))
    return chi2_all, float(l32[0]), float(l64[0])


# ---------------------------------------------------------------- light-curve model
def bilinear(m, x, y):
    m = f64(m)
    return lib().orc_bilinear(_p(m), m.shape[1], m.shape[0], x, y)


def lc_chi2_grad(maps, B, node_off, node_cnt, node_rad, node_map, band_scale, params, prof,
                 obs, obs_off, t, mobs, sig):
    """maps: (nmaps, Ny, Nx, 2) float64 of (value, d/dr); node_map None means
    node k uses map k."""
    maps = f64(maps)
    nm, Ny, Nx, two = maps.shape
    assert two == 2
    if node_map is None:
        node_map = np.arange(len(node_rad), dtype=np.int32)
    params = f64(params).reshape(-1, 7)
    nc = params.shape[0]
    chi2 = np.zeros(nc)
    grad = np.zeros((nc, 7))
    args = [i32(node_off), i32(node_cnt), f64(node_rad), i32(node_map), f64(band_scale)]
    prof, obs, obs_off = i32(prof), i32(obs), i32(obs_off)
    t, mobs, sig = f64(t), f64(mobs), f64(sig)
    lib().orc_lc_chi2_grad(_p(maps), Nx, Ny, B, _p(args[0]), _p(args[1]), _p(args[2]),
                           _p(args[3]), _p(args[4]), nc, _p(params), _p(prof), _p(obs),
                           _p(obs_off), _p(t), _p(mobs), _p(sig), _p(chi2), _p(grad))
    return chi2, grad
