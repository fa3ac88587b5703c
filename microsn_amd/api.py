"""Python face of libmlc.so over torch device memory.

The functions mirror the reference's host interface for this path
(simulating_lcurves/include/gpu_functions.hpp:11-18, cpu_functions.hpp:9-15,
chi2/chi2_functions.cuh:93) with the same names, argument meaning and data
layouts; device arrays are torch tensors on the context's device.  torch
provides device memory and the stream; every computation runs in libmlc.so.
"""
import ctypes as C
import weakref

import numpy as np
import torch

from ._lib import F32, F64, SHAPES, MlcError, lib, ptr

PREC = {"f32": F32, "f64": F64, F32: F32, F64: F64}


def _prec(p):
    try:
        return PREC[p]
    except KeyError:
        raise ValueError(f"precision must be 'f32' or 'f64', not {p!r}") from None


def _shape(s):
    if isinstance(s, int):
        return s
    try:
        return SHAPES[s]
    except KeyError:
        raise ValueError(f"unknown profile shape {s!r}") from None


class Context:
    """An mlc_ctx bound to one device and (by default) torch's current stream."""

    def __init__(self, device=0, stream=None):
        self.device = torch.device("cuda", device) if isinstance(device, int) else torch.device(device)
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        h = C.c_void_p()
        rc = lib().mlc_ctx_create(self.device.index or 0, C.c_void_p(stream.cuda_stream), C.byref(h))
        if rc != 0:
            raise MlcError(rc, "mlc_ctx_create failed")
        self.h = h
        self.stream = stream
        self._maps = weakref.WeakSet()  # closed before the context (they point to it)

    def check(self, rc):
        if rc != 0:
            raise MlcError(rc, lib().mlc_last_error(self.h).decode())

    def set_stream(self, stream):
        """Run the library's work on another torch stream (e.g. one that a
        torch.cuda.graph capture records)."""
        self.check(lib().mlc_ctx_set_stream(self.h, C.c_void_p(stream.cuda_stream)))
        self.stream = stream

    def sync(self):
        self.check(lib().mlc_ctx_sync(self.h))

    def generation(self):
        """Counter bumped whenever the library reallocates memory or rewrites
        a device table that a captured graph may refer to: a graph captured
        at generation g replays correctly while this still returns g."""
        g = C.c_uint64()
        self.check(lib().mlc_ctx_generation(self.h, C.byref(g)))
        return g.value

    def reserve(self, nbytes):
        """Grow the workspace to at least nbytes now (before a capture)."""
        self.check(lib().mlc_ctx_reserve(self.h, int(nbytes)))

    # ---- launch-shape tuner of the band kernels (mlc_tune_*)
    TUNE_OFF, TUNE_AUTO, TUNE_FORCE = 0, 1, 2
    TUNE_KIND_BANDS, TUNE_KIND_BATCH = 0, 1
    TUNE_GENERAL_NODES = 1
    TUNE_BATCH_PER_BAND = 2

    def tune_mode(self, mode):
        self.check(lib().mlc_tune_mode(self.h, int(mode)))

    def tune_force(self, lanes, grid_multiple=1, flags=0):
        """Use this launch shape for every band launch (sets TUNE_FORCE)."""
        self.check(lib().mlc_tune_force(self.h, int(lanes), int(grid_multiple), int(flags)))
        self.tune_mode(self.TUNE_FORCE)

    def tune_lookup(self, kind, precision, epochs_per_band, pixels, with_flags=False):
        """(lanes, grid_multiple) kept for a key, or (0, 0); with_flags also
        the shape flags (TUNE_BATCH_PER_BAND)."""
        la, gm, fl = C.c_int(), C.c_int(), C.c_int()
        self.check(lib().mlc_tune_lookup(self.h, int(kind), _prec(precision), float(epochs_per_band),
                                         float(pixels), C.byref(la), C.byref(gm), C.byref(fl)))
        return (la.value, gm.value, fl.value) if with_flags else (la.value, gm.value)

    def tune_save(self, path):
        self.check(lib().mlc_tune_save(self.h, str(path).encode()))

    def tune_load(self, path):
        self.check(lib().mlc_tune_load(self.h, str(path).encode()))

    def close(self):
        if getattr(self, "h", None):
            for m in list(getattr(self, "_maps", ())):
                m.close()
            lib().mlc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- per-kernel timing (HIP events on the context stream)
    def timing(self, on=True):
        self.check(lib().mlc_timing_enable(self.h, int(on)))

    def timing_reset(self):
        self.check(lib().mlc_timing_reset(self.h))

    def timing_get(self, name):
        ms, n = C.c_double(), C.c_longlong()
        self.check(lib().mlc_timing_get(self.h, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def timing_all(self):
        buf = C.create_string_buffer(1 << 16)
        self.check(lib().mlc_timing_names(self.h, buf, len(buf)))
        return {k: self.timing_get(k) for k in buf.value.decode().split("\n") if k}


_default = {}


def context(device=0):
    """A cached Context for `device` on torch's current stream."""
    key = (device, torch.cuda.current_stream(device).cuda_stream)
    if key not in _default:
        _default[key] = Context(device)
    return _default[key]


def _dev(ctx, t, dtype, name):
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == dtype and t.is_contiguous()):
        raise TypeError(f"{name} must be a contiguous {dtype} CUDA tensor")
    if t.device != ctx.device:
        raise ValueError(f"{name} is on {t.device}, the context on {ctx.device}")
    return t


# ------------------------------------------------------------------ maps
class MapSpectrum:
    """mlc_map: the map's 2-D spectrum held on the device in one precision."""

    def __init__(self, ctx, dmap, precision="f32", subtract_mean=None):
        _dev(ctx, dmap, torch.float64, "map")
        self.ctx, self.prec = ctx, _prec(precision)
        self.Ny, self.Nx = dmap.shape
        if subtract_mean is None:
            subtract_mean = self.prec == F32
        h = C.c_void_p()
        ctx.check(lib().mlc_map_create(ctx.h, self.prec, ptr(dmap), self.Nx, self.Ny,
                                       int(bool(subtract_mean)), C.byref(h)))
        self.h = h
        ctx._maps.add(self)

    @property
    def dtype(self):
        return torch.float64 if self.prec == F64 else torch.float32

    def update(self, dmap):
        _dev(self.ctx, dmap, torch.float64, "map")
        self.ctx.check(lib().mlc_map_update(self.h, ptr(dmap)))

    def convolve(self, shape, rhalf, pix=1.0, rin=0.0, a=1.0, value=True, node=False, plain=False,
                 tiled=False):
        """Convolved map (Ny, Nx) and/or node map (Ny, Nx, 2) = (mu, dmu/drhalf).

        The library writes node maps in 4 x 4 pixel tiles; they are returned
        row-major unless tiled=True."""
        v = torch.empty((self.Ny, self.Nx), dtype=self.dtype, device=self.ctx.device) if value else None
        nd = torch.empty((self.Ny, self.Nx, 2), dtype=self.dtype, device=self.ctx.device) if node else None
        f = lib().mlc_map_convolve_plain if plain else lib().mlc_map_convolve
        self.ctx.check(f(self.h, _shape(shape), float(rhalf), float(pix), float(rin), float(a),
                         ptr(v), ptr(nd)))
        if node and not tiled:
            nd = untile(nd)
        if value and node:
            return v, nd
        return v if value else nd

    def close(self):
        if getattr(self, "h", None):
            lib().mlc_map_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def update_spectra(spectra, maps):
    """spectra[i].update(maps[i]) for every i in one call (mlc_maps_update:
    the same bits, the small per-map kernels run once for all maps).  The
    spectra share one context and precision."""
    spectra, maps = list(spectra), list(maps)
    if len(spectra) != len(maps):
        raise ValueError("update_spectra: one map per spectrum")
    if not spectra:
        return
    ctx = spectra[0].ctx
    for sp, m in zip(spectra, maps):
        if sp.ctx is not ctx:
            raise ValueError("update_spectra: spectra of different contexts")
        _dev(ctx, m, torch.float64, "map")
        if tuple(m.shape) != (sp.Ny, sp.Nx):
            raise ValueError(f"update_spectra: map of shape {tuple(m.shape)} for a {sp.Ny} x {sp.Nx} spectrum")
    n = len(spectra)
    hs = (C.c_void_p * n)(*[sp.h.value for sp in spectra])
    ds = (C.c_void_p * n)(*[m.data_ptr() for m in maps])
    ctx.check(lib().mlc_maps_update(n, hs, ds))


def untile(t):
    """Row-major (..., Ny, Nx, 2) copy of a node map stored in 4 x 4 pixel tiles."""
    *lead, Ny, Nx, two = t.shape
    v = t.reshape(*lead, Ny // 4, Nx // 4, 4, 4, two)
    nl = len(lead)
    perm = list(range(nl)) + [nl, nl + 2, nl + 1, nl + 3, nl + 4]
    return v.permute(*perm).reshape(*lead, Ny, Nx, two)


def gerlumph_normalise(ctx, counts, avgmu, avgN, k, g):
    """MagnificationMap FP64 normalisation of int32 ray counts (device)."""
    _dev(ctx, counts, torch.int32, "counts")
    out = torch.empty(counts.shape, dtype=torch.float64, device=ctx.device)
    ctx.check(lib().mlc_gerlumph_normalise(ctx.h, ptr(counts), counts.numel(), avgmu, avgN, k, g,
                                           ptr(out)))
    return out


def profile(ctx, shape, rhalf, pix=1.0, rin=0.0, a=1.0):
    """(profile, d profile/d rhalf), FP64 n x n device tensors."""
    n = C.c_int()
    ctx.check(lib().mlc_profile(ctx.h, _shape(shape), rhalf, pix, rin, a, C.byref(n), None, None, 0))
    p = torch.empty((n.value, n.value), dtype=torch.float64, device=ctx.device)
    dp = torch.empty_like(p)
    ctx.check(lib().mlc_profile(ctx.h, _shape(shape), rhalf, pix, rin, a, C.byref(n), ptr(p), ptr(dp),
                                p.numel()))
    return p, dp


def parity_round(x):
    return lib().mlc_parity_round(float(x))


def fft_r2c(ctx, x):
    """Unnormalised real 2-D FFT of a float32/float64 Ny x Nx device tensor."""
    if x.dtype not in (torch.float32, torch.float64):
        raise TypeError(f"fft_r2c: input must be float32 or float64, not {x.dtype}")
    if x.dim() != 2:
        raise ValueError(f"fft_r2c: input must be 2-D, not {tuple(x.shape)}")
    prec = F64 if x.dtype == torch.float64 else F32
    _dev(ctx, x, x.dtype, "input")
    Ny, Nx = x.shape
    out = torch.empty((Ny, Nx // 2 + 1), dtype=torch.complex128 if prec == F64 else torch.complex64,
                      device=ctx.device)
    ctx.check(lib().mlc_fft_r2c(ctx.h, prec, ptr(x), ptr(out), Nx, Ny))
    return out


def fft_c2r(ctx, X, Nx):
    """Unnormalised inverse of fft_r2c: complex64/complex128 Ny x (Nx/2+1) in."""
    if X.dtype not in (torch.complex64, torch.complex128):
        raise TypeError(f"fft_c2r: input must be complex64 or complex128, not {X.dtype}")
    if X.dim() != 2 or X.shape[1] != Nx // 2 + 1:
        raise ValueError(f"fft_c2r: input must be Ny x {Nx // 2 + 1}, not {tuple(X.shape)}")
    prec = F64 if X.dtype == torch.complex128 else F32
    _dev(ctx, X, X.dtype, "input")
    Ny = X.shape[0]
    out = torch.empty((Ny, Nx), dtype=torch.float64 if prec == F64 else torch.float32, device=ctx.device)
    ctx.check(lib().mlc_fft_c2r(ctx.h, prec, ptr(X), ptr(out), Nx, Ny))
    return out


# ------------------------------------------------------------------ light curves (reference mode)
class SimLC:
    """data_structs.hpp:56-63: LC (Nloc x Nprof) and DLC (Nloc x Nprof-1), FP64."""

    def __init__(self, ctx, Nloc, Nprof):
        self.Nloc, self.Nprof = Nloc, Nprof
        self.LC = torch.zeros((Nloc, Nprof), dtype=torch.float64, device=ctx.device)
        self.DLC = torch.zeros((Nloc, max(Nprof - 1, 1)), dtype=torch.float64, device=ctx.device)


def sample_points(ctx, conv, k, loc_x, loc_y, LC):
    prec = F64 if conv.dtype == torch.float64 else F32
    Ny, Nx = conv.shape
    ctx.check(lib().mlc_sample_points(ctx.h, prec, ptr(conv), Nx, Ny, k, LC.Nprof, LC.Nloc,
                                      ptr(loc_x), ptr(loc_y), ptr(LC.LC), ptr(LC.DLC)))


def expanding_source(ctx, dmap, width, sizes, shape, loc_x, loc_y, LC, precision="f64"):
    """gpu_functions.hpp:11 (expanding_source)."""
    _dev(ctx, dmap, torch.float64, "map")
    _dev(ctx, loc_x, torch.int32, "loc_x")
    _dev(ctx, loc_y, torch.int32, "loc_y")
    sizes = np.ascontiguousarray(sizes, dtype=np.float64)
    Ny, Nx = dmap.shape
    ctx.check(lib().mlc_expanding_source(ctx.h, _prec(precision), ptr(dmap), Nx, Ny, float(width),
                                         ptr(sizes), sizes.size, _shape(shape), LC.Nloc, ptr(loc_x),
                                         ptr(loc_y), ptr(LC.LC), ptr(LC.DLC)))


class Chi2SortBins:
    """data_structs.hpp:41-54: sorted_ind (device, uint32 stored as int32),
    lower_ind / n_per_bin (host uint32)."""

    def __init__(self, ctx, Nloc, Nbins):
        self.Nloc, self.Nbins = Nloc, Nbins
        self.d_sorted_ind = torch.empty(Nloc * Nloc, dtype=torch.int32, device=ctx.device)
        self.lower_ind = np.zeros(Nbins, np.uint32)
        self.n_per_bin = np.zeros(Nbins, np.uint32)


def setup_integral_GPU(ctx, LCA, LCB, bins, sort_struct):
    """gpu_functions.hpp:15: z ratio, stable sort, sorted_ind, bin table."""
    bins = np.ascontiguousarray(bins, np.float64)
    ctx.check(lib().mlc_setup_integral(ctx.h, ptr(LCA.LC), ptr(LCB.LC), LCA.Nloc, LCA.Nprof,
                                       ptr(bins), bins.size, ptr(sort_struct.d_sorted_ind),
                                       ptr(sort_struct.lower_ind), ptr(sort_struct.n_per_bin)))


def calculate_chi2_GPU(ctx, chi2_vars, chi2_values, sort_struct, LCA, LCB):
    """gpu_functions.hpp:16.  chi2_vars: dict of Njp, indA, indB, facA, facB, d, s."""
    v = chi2_vars
    arr = {k: np.ascontiguousarray(v[k], np.int32 if k.startswith("ind") else np.float64)
           for k in ("indA", "indB", "facA", "facB", "d", "s")}
    ctx.check(lib().mlc_calculate_chi2(
        ctx.h, int(v["Njp"]), ptr(arr["indA"]), ptr(arr["indB"]), ptr(arr["facA"]), ptr(arr["facB"]),
        ptr(arr["d"]), ptr(arr["s"]), ptr(sort_struct.d_sorted_ind), ptr(LCA.LC), ptr(LCA.DLC),
        ptr(LCB.LC), ptr(LCB.DLC), LCA.Nloc, LCA.Nprof, ptr(chi2_values)))


def bin_chi2_GPU(ctx, sort_struct, chi2_values):
    """gpu_functions.hpp:17 -> (binned_chi2, binned_exp) host arrays."""
    nb = sort_struct.Nbins
    bc, be = np.zeros(nb), np.zeros(nb)
    ctx.check(lib().mlc_bin_chi2(ctx.h, ptr(chi2_values), chi2_values.numel(), ptr(sort_struct.lower_ind),
                                 ptr(sort_struct.n_per_bin), nb, ptr(bc), ptr(be)))
    return bc, be


def sort_f64(ctx, keys, rank=False, sorted_keys=False):
    _dev(ctx, keys, torch.float64, "keys")
    n = keys.numel()
    perm = torch.empty(n, dtype=torch.int32, device=ctx.device)
    rk = torch.empty(n, dtype=torch.int32, device=ctx.device) if rank else None
    sk = torch.empty(n, dtype=torch.float64, device=ctx.device) if sorted_keys else None
    ctx.check(lib().mlc_sort_f64(ctx.h, ptr(keys), n, ptr(perm), ptr(rk), ptr(sk)))
    return perm, rk, sk


def bin_edges(ctx, sorted_keys, bins):
    bins = np.ascontiguousarray(bins, np.float64)
    lo = np.zeros(bins.size, np.uint32)
    npb = np.zeros(bins.size, np.uint32)
    ctx.check(lib().mlc_bin_edges(ctx.h, ptr(sorted_keys), sorted_keys.numel(), ptr(bins), bins.size,
                                  ptr(lo), ptr(npb)))
    return lo, npb


def bin_table_device(ctx, sorted_keys, bins):
    """(lower_ind, n_per_bin, binned_chi2, binned_exp) as device tensors
    (uint32 stored as int32, FP64) for ascending keys; the same values as
    bin_edges + bin_chi2_GPU, with no host round trip."""
    _dev(ctx, sorted_keys, torch.float64, "keys")
    bins = np.ascontiguousarray(bins, np.float64)
    nb = bins.size
    lo = torch.empty(nb, dtype=torch.int32, device=ctx.device)
    npb = torch.empty(nb, dtype=torch.int32, device=ctx.device)
    bc = torch.empty(nb, dtype=torch.float64, device=ctx.device)
    be = torch.empty(nb, dtype=torch.float64, device=ctx.device)
    ctx.check(lib().mlc_bin_table_device(ctx.h, ptr(sorted_keys), sorted_keys.numel(), ptr(bins), nb, ptr(lo),
                                         ptr(npb), ptr(bc), ptr(be)))
    return lo, npb, bc, be


def binned_mpd(ctx, dmap, Nbins, vmin=0.02, vmax=200.0):
    """MagnificationMap::getBinnedMpd -> (bins, counts) host arrays."""
    bins = mpd_bins(Nbins, vmin, vmax, True)
    counts = np.zeros(Nbins)
    ctx.check(lib().mlc_binned_mpd(ctx.h, ptr(dmap), dmap.numel(), Nbins, vmin, vmax, ptr(bins),
                                   ptr(counts)))
    return bins, counts


def getChi2GPU(ctx, f_obs, df_obs, mu1, mu2, mode=0):
    """chi2_functions.cuh:93 -> (chi2_all (N*N, [N*y+x]), likelihood FP64)."""
    N, F, T = mu1.shape
    out = torch.empty(N * N, dtype=torch.float32, device=ctx.device)
    like = C.c_double()
    ctx.check(lib().mlc_chi2_multiband(ctx.h, ptr(f_obs), ptr(df_obs), ptr(mu1), ptr(mu2), N, F, T,
                                       mode, ptr(out), C.byref(like)))
    return out, like.value


# ------------------------------------------------------------------ host-side helpers
def locations_on_grid(Nside, Nmap, Noff):
    x = np.zeros(Nside * Nside, np.int32)
    y = np.zeros_like(x)
    if lib().mlc_locations_on_grid(Nside, Nmap, Noff, ptr(x), ptr(y)):
        raise ValueError("bad grid arguments")
    return x, y


def locations_random(Nloc, Nmap, Noff, seed):
    x = np.zeros(Nloc, np.int32)
    y = np.zeros_like(x)
    if lib().mlc_locations_random(Nloc, Nmap, Noff, seed, ptr(x), ptr(y)):
        raise ValueError("bad random-location arguments")
    return x, y


def setup_chi2_calculation(M, V, Dfac, d, t, s, x, quirk_const_term=True):
    """cpu_functions.hpp:10 -> dict(Njp, indA, indB, facA, facB, d, s, const_term)."""
    d, t, s, x = (np.ascontiguousarray(a, np.float64) for a in (d, t, s, x))
    Nd = d.size
    indA, indB = np.zeros(Nd, np.int32), np.zeros(Nd, np.int32)
    facA, facB, nd, ns = (np.zeros(Nd) for _ in range(4))
    Njp, ct = C.c_int(), C.c_double()
    rc = lib().mlc_setup_chi2(M[0], M[1], V[0], Dfac, Nd, ptr(d), ptr(t), ptr(s), x.size, ptr(x),
                              int(quirk_const_term), ptr(indA), ptr(indB), ptr(facA), ptr(facB),
                              ptr(nd), ptr(ns), C.byref(Njp), C.byref(ct))
    if rc:
        raise ValueError("bad chi2 set-up arguments")
    n = Njp.value
    return dict(Njp=n, indA=indA[:n], indB=indB[:n], facA=facA[:n], facB=facB[:n], d=nd[:n],
                s=ns[:n], const_term=ct.value)


def find_indices(t, x):
    t, x = np.ascontiguousarray(t, np.float64), np.ascontiguousarray(x, np.float64)
    ind = np.zeros(t.size, np.int32)
    lib().mlc_find_indices(ptr(t), ptr(x), t.size, x.size, ptr(ind))
    return ind


def mpd_bins(Nbins, vmin, vmax, log_scale=True):
    b = np.zeros(Nbins)
    if lib().mlc_mpd_bins(Nbins, vmin, vmax, int(log_scale), ptr(b)):
        raise ValueError("bad MPD bins")
    return b


def mpd_normalize(bins, counts):
    bins = np.ascontiguousarray(bins, np.float64)
    c = np.array(counts, np.float64)
    lib().mlc_mpd_normalize(bins.size, ptr(bins), ptr(c))
    return c


def mpd_divide_by(bA, cA, bB, cB, bR):
    bA, cA, bB, cB, bR = (np.ascontiguousarray(a, np.float64) for a in (bA, cA, bB, cB, bR))
    out = np.zeros(bR.size)
    if lib().mlc_mpd_divide_by(bA.size, ptr(bA), ptr(cA), bB.size, ptr(bB), ptr(cB), bR.size, ptr(bR),
                               ptr(out)):
        raise ValueError("bad MPDs")
    return out


def trapezium(f1, f2, x):
    f1, f2, x = (np.ascontiguousarray(a, np.float64) for a in (f1, f2, x))
    return lib().mlc_trapezium(x.size, ptr(f1), ptr(f2), ptr(x))


# ------------------------------------------------------------------ source sizes
def _stof(v):
    """std::stof of the string form of a parameter (profile.cpp:36-48)."""
    return float(np.float32(float(v)))


def get_size(pars, lrest=None, wavelength=None, throughput=None):
    """BaseProfile::getSize (profile.cpp:31-84): half-light radius from the
    size law in pars["type"] ("parametric": r0, l0, nu; "ss_disc": mbh, fedd,
    eta; "vector"; "custom"), at rest wavelength lrest or averaged over a
    filter curve (wavelength, throughput; trapezoid rule).  Parameters are
    read through std::stof as in the reference.  An unknown type gives 0."""
    t = pars.get("type")
    band = wavelength is not None
    if band:
        wl = np.ascontiguousarray(wavelength, np.float64)
        tp = np.ascontiguousarray(throughput, np.float64)
        if wl.shape != tp.shape or wl.size < 2:
            raise ValueError("wavelength and throughput must be equal-length arrays of >= 2 points")
    if t == "parametric":
        r0, l0, nu = (_stof(pars[k]) for k in ("r0", "l0", "nu"))
        return (lib().mlc_size_parametric_band(r0, l0, nu, ptr(wl), ptr(tp), wl.size) if band
                else lib().mlc_size_parametric(r0, l0, nu, float(lrest)))
    if t == "ss_disc":
        mbh, fedd, eta = (_stof(pars[k]) for k in ("mbh", "fedd", "eta"))
        return (lib().mlc_size_ss_band(mbh, fedd, eta, ptr(wl), ptr(tp), wl.size) if band
                else lib().mlc_size_ss(mbh, fedd, eta, float(lrest)))
    if t == "vector":
        return lib().mlc_size_vector_band(ptr(wl), ptr(tp), wl.size) if band else float(lrest)
    if t == "custom":
        return 1.0
    return 0.0
