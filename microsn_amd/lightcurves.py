"""Light curves along straight trajectories with a per-curve magnitude chi2
and its adjoint (libmlc.so: mlc_map_convolve node maps + mlc_lc_chi2_grad),
and the torch.autograd.Function that exposes them to a Python sampler (the
caller slot of driver_code/sampling.ipynb).

Curve parameters, one row of 7 per curve:
    x0, y0      trajectory origin (pixels; pixel centres at integer coords)
    theta       direction (radians)
    v           effective velocity (pixels per time unit)
    r0, w       source half-light radius at t = 0 and its growth rate:
                radius of band b at time t = band_scale[b] * (r0 + w t)
    dm          magnitude offset of the image
"""
import ctypes as C

import numpy as np
import torch

from ._lib import F32, F64, lib, ptr
from .api import MapSpectrum, _dev, _prec

NPARAM = 7


class NodeSet:
    """Convolved (mu, dmu/drhalf) maps at a few radii per (profile, band).

    nodes[k] is node map k, shape (Ny, Nx, 2), stored in the library's 4 x 4
    pixel tiles (rowmajor() gives the plain layout); node_off/node_cnt give
    the nodes of profile p in band b with ascending radii node_rad."""

    def __init__(self, ctx, spectrum: MapSpectrum, shapes, radii, band_scale, pix=1.0, plain=False,
                 compute=True):
        """shapes: P profile shapes; radii[p][b]: ascending list of node radii
        (half-light radius in pixels) for profile p, band b.  With
        compute=False the maps are only allocated (see compute())."""
        self.ctx, self.spectrum, self.pix = ctx, spectrum, float(pix)
        self.shapes = [_shape_id(s) for s in shapes]
        self.P, self.B = len(shapes), len(band_scale)
        self.Ny, self.Nx = spectrum.Ny, spectrum.Nx
        self.prec = spectrum.prec
        rad, off, cnt, shp = [], [], [], []
        for p in range(self.P):
            for b in range(self.B):
                rs = list(radii[p][b])
                off.append(len(rad))
                cnt.append(len(rs))
                rad.extend(float(r) for r in rs)
                shp.extend([self.shapes[p]] * len(rs))
        self.node_off = np.array(off, np.int32)
        self.node_cnt = np.array(cnt, np.int32)
        self.node_rad = np.array(rad, np.float64)
        self.node_shape = shp
        self.band_scale = np.ascontiguousarray(band_scale, np.float64)
        self.nodes = torch.empty((len(rad), self.Ny, self.Nx, 2), dtype=spectrum.dtype, device=ctx.device)
        if compute:
            self.compute(plain)

    def rowmajor(self):
        """All node maps as a row-major (nnodes, Ny, Nx, 2) tensor."""
        from .api import untile
        return untile(self.nodes)

    def compute(self, plain=False):
        """(Re)convolve every node map from the spectrum's current map: one
        mlc_map_convolve_nodes call, or one mlc_map_convolve_plain call per
        node on the plain path."""
        K = len(self.node_rad)
        if plain:
            for k, (s, r) in enumerate(zip(self.node_shape, self.node_rad)):
                self.ctx.check(lib().mlc_map_convolve_plain(self.spectrum.h, s, float(r), self.pix, 0.0, 1.0,
                                                            None, ptr(self.nodes[k])))
            return
        if not hasattr(self, "_batch"):
            shp = np.ascontiguousarray(self.node_shape, np.int32)
            rad = np.ascontiguousarray(self.node_rad, np.float64)
            dst = (C.c_void_p * K)(*[self.nodes[k].data_ptr() for k in range(K)])
            self._batch = (shp, rad, dst)
        shp, rad, dst = self._batch
        self.ctx.check(lib().mlc_map_convolve_nodes(self.spectrum.h, K, shp.ctypes.data, rad.ctypes.data,
                                                    self.pix, None, None, C.addressof(dst)))


def _shape_id(s):
    from ._lib import SHAPES
    return s if isinstance(s, int) else SHAPES[s]


class Observations:
    """Observed epochs of nobs observation sets: for set o and band b the
    epochs are t[off[o*B+b] : off[o*B+b+1]] (FP64 on the device)."""

    def __init__(self, ctx, B, sets):
        """sets: list of per-set lists of B (t, mobs, sig) triples of arrays."""
        t, m, s, off = [], [], [], [0]
        for bands in sets:
            assert len(bands) == B
            for (tb, mb, sb) in bands:
                tb, mb, sb = (np.atleast_1d(np.asarray(a, np.float64)) for a in (tb, mb, sb))
                assert tb.shape == mb.shape == sb.shape
                t.append(tb)
                m.append(mb)
                s.append(sb)
                off.append(off[-1] + tb.size)
        self.B, self.nobs = B, len(sets)
        cat = lambda xs: torch.tensor(np.concatenate(xs) if xs else np.zeros(0), dtype=torch.float64,
                                      device=ctx.device)
        self.t, self.mobs, self.sig = cat(t), cat(m), cat(s)
        self.off = torch.tensor(np.array(off, np.int32), device=ctx.device)
        self.nepoch = int(off[-1])


def chi2_grad(ctx, nodes: NodeSet, obs: Observations, params, prof, obs_idx, plain=False, convolve=False):
    """(chi2 [n], grad [n, 7]) FP64 device tensors for n curves.

    convolve=True first recomputes the node maps from the spectrum's current
    map, in the same call (mlc_lc_chi2_grad_convolve); the result equals
    nodes.compute() followed by chi2_grad(...) bit for bit."""
    _dev(ctx, params, torch.float64, "params")
    _dev(ctx, prof, torch.int32, "prof")
    _dev(ctx, obs_idx, torch.int32, "obs")
    n = params.shape[0]
    assert params.shape == (n, NPARAM) and prof.shape == (n,) and obs_idx.shape == (n,)
    chi2 = torch.empty(n, dtype=torch.float64, device=ctx.device)
    grad = torch.empty((n, NPARAM), dtype=torch.float64, device=ctx.device)
    if convolve:
        assert not plain, "convolve=True runs the fused path"
        if not hasattr(nodes, "_shp"):
            nodes._shp = np.ascontiguousarray(nodes.node_shape, np.int32)
        shp = nodes._shp
        ctx.check(lib().mlc_lc_chi2_grad_convolve(
            nodes.spectrum.h, shp.ctypes.data, nodes.pix, ptr(nodes.nodes), nodes.nodes.shape[0], nodes.P,
            nodes.B, ptr(nodes.node_off), ptr(nodes.node_cnt), ptr(nodes.node_rad), ptr(nodes.band_scale), n,
            ptr(params), ptr(prof), ptr(obs_idx), obs.nobs, ptr(obs.off), obs.nepoch, ptr(obs.t),
            ptr(obs.mobs), ptr(obs.sig), ptr(chi2), ptr(grad)))
        return chi2, grad
    f = lib().mlc_lc_chi2_grad_plain if plain else lib().mlc_lc_chi2_grad
    ctx.check(f(ctx.h, nodes.prec, ptr(nodes.nodes), nodes.nodes.shape[0], nodes.Nx, nodes.Ny, nodes.P,
                nodes.B, ptr(nodes.node_off), ptr(nodes.node_cnt), ptr(nodes.node_rad),
                ptr(nodes.band_scale), n, ptr(params), ptr(prof), ptr(obs_idx), obs.nobs, ptr(obs.off),
                obs.nepoch, ptr(obs.t), ptr(obs.mobs), ptr(obs.sig), ptr(chi2), ptr(grad)))
    return chi2, grad


class NodeBatch:
    """The node maps of several maps as one batched call sees them (host
    tables of mlc_lc_chi2_grad_batch): map m is nodesets[m].  Every NodeSet
    has the same profile count P, bands B, band scales and precision; the
    maps may differ in size.  Holds references to the node sets."""

    def __init__(self, nodesets):
        self.nodesets = list(nodesets)
        first = self.nodesets[0]
        self.P, self.B, self.prec = first.P, first.B, first.prec
        for ns in self.nodesets:
            if (ns.P, ns.B, ns.prec) != (self.P, self.B, self.prec) or \
                    not np.array_equal(ns.band_scale, first.band_scale):
                raise ValueError("NodeBatch: node sets differ in profiles, bands, band scales or precision")
        self.Nx = np.array([ns.Nx for ns in self.nodesets], np.int32)
        self.Ny = np.array([ns.Ny for ns in self.nodesets], np.int32)
        ptrs, rad, off, cnt = [], [], [], []
        for ns in self.nodesets:
            base = len(ptrs)
            ptrs.extend(ns.nodes[k].data_ptr() for k in range(ns.nodes.shape[0]))
            rad.extend(ns.node_rad.tolist())
            off.extend((ns.node_off + base).tolist())
            cnt.extend(ns.node_cnt.tolist())
        self.nnodes = len(ptrs)
        self.dptr = (C.c_void_p * len(ptrs))(*ptrs)
        self.node_rad = np.array(rad, np.float64)
        self.node_off = np.array(off, np.int32)
        self.node_cnt = np.array(cnt, np.int32)
        self.band_scale = np.ascontiguousarray(first.band_scale, np.float64)

    def compute(self):
        """(Re)convolve every node map of every map from its spectrum's
        current map: one mlc_maps_convolve_nodes call (the same bits as
        NodeSet.compute() on each set); per set when their pixel sizes
        differ."""
        sets = self.nodesets
        if len({ns.pix for ns in sets}) > 1:
            for ns in sets:
                ns.compute()
            return
        if not hasattr(self, "_conv"):
            nm = len(sets)
            hs = (C.c_void_p * nm)(*[ns.spectrum.h.value for ns in sets])
            counts = np.array([len(ns.node_rad) for ns in sets], np.int32)
            shp = np.ascontiguousarray(np.concatenate([np.asarray(ns.node_shape, np.int32) for ns in sets]))
            rad = np.ascontiguousarray(np.concatenate([ns.node_rad for ns in sets]), np.float64)
            dst = (C.c_void_p * len(rad))(*[ns.nodes[k].data_ptr() for ns in sets for k in range(ns.nodes.shape[0])])
            self._conv = (nm, hs, counts, shp, rad, dst)
        nm, hs, counts, shp, rad, dst = self._conv
        sets[0].ctx.check(lib().mlc_maps_convolve_nodes(nm, hs, counts.ctypes.data, shp.ctypes.data, rad.ctypes.data,
                                                        sets[0].pix, None, None, dst))


def chi2_grad_batch(ctx, batch: NodeBatch, obs: Observations, params, map_idx, prof, obs_idx):
    """(chi2 [n], grad [n, 7]) for curves on several maps in one call
    (mlc_lc_chi2_grad_batch): curve c lies on map map_idx[c] of the batch."""
    _dev(ctx, params, torch.float64, "params")
    _dev(ctx, map_idx, torch.int32, "map_idx")
    _dev(ctx, prof, torch.int32, "prof")
    _dev(ctx, obs_idx, torch.int32, "obs")
    n = params.shape[0]
    assert params.shape == (n, NPARAM) and map_idx.shape == prof.shape == obs_idx.shape == (n,)
    b = batch
    chi2 = torch.empty(n, dtype=torch.float64, device=ctx.device)
    grad = torch.empty((n, NPARAM), dtype=torch.float64, device=ctx.device)
    ctx.check(lib().mlc_lc_chi2_grad_batch(
        ctx.h, b.prec, len(b.nodesets), b.Nx.ctypes.data, b.Ny.ctypes.data, b.nnodes, C.addressof(b.dptr),
        b.node_rad.ctypes.data, b.P, b.B, b.node_off.ctypes.data, b.node_cnt.ctypes.data, b.band_scale.ctypes.data,
        n, ptr(params), ptr(map_idx), ptr(prof), ptr(obs_idx), obs.nobs, ptr(obs.off), obs.nepoch, ptr(obs.t),
        ptr(obs.mobs), ptr(obs.sig), ptr(chi2), ptr(grad)))
    return chi2, grad


class LightCurveChi2(torch.autograd.Function):
    """chi2 per curve as a differentiable function of the curve parameters.

    forward runs the fused chi2 + adjoint kernel once and keeps the gradient;
    backward scales it by the incoming gradient."""

    @staticmethod
    def forward(fctx, params, ctx, nodes, obs, prof, obs_idx):
        chi2, grad = chi2_grad(ctx, nodes, obs, params.detach().contiguous(), prof, obs_idx)
        fctx.save_for_backward(grad)
        return chi2

    @staticmethod
    def backward(fctx, g):
        (grad,) = fctx.saved_tensors
        return g[:, None] * grad, None, None, None, None, None


def lightcurve_chi2(params, ctx, nodes, obs, prof, obs_idx):
    return LightCurveChi2.apply(params, ctx, nodes, obs, prof, obs_idx)
