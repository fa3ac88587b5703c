"""Map input and output: GERLUMPH map directories (the reference's
MagnificationMap constructor, simulating_lcurves/src/magnification_map.cpp:7-40)
and the FITS / raw binary image writers of Image
(simulating_lcurves/src/image.cpp:4-62), and FITS images as an optional map
input.  Reading and writing run on the host
inside libmlc.so (mlc_host.h); the normalisation to magnifications runs on
the device (mlc_gerlumph_normalise) or, for host arrays, here in FP64 with
the same formula."""
import ctypes as C
import os

import numpy as np

from ._lib import lib, ptr

_READ_ERR = {1: "bad arguments", 2: "missing or malformed mapmeta.dat", 3: "missing or short map.bin"}


def read_gerlumph_meta(map_dir):
    """mapmeta.dat of a GERLUMPH map directory as a dict
    (avgmu, avgN, Nx, width, k, g, s)."""
    d = [C.c_double() for _ in range(6)]
    nx = C.c_int()
    rc = lib().mlc_gerlumph_read(os.fsencode(map_dir), C.byref(d[0]), C.byref(d[1]), C.byref(nx),
                                 C.byref(d[2]), C.byref(d[3]), C.byref(d[4]), C.byref(d[5]), None, 0)
    if rc:
        raise OSError(f"{map_dir}: {_READ_ERR.get(rc, rc)}")
    return dict(avgmu=d[0].value, avgN=d[1].value, Nx=nx.value, width=d[2].value, k=d[3].value,
                g=d[4].value, s=d[5].value)


def read_gerlumph(map_dir):
    """(meta, counts): the metadata and the Nx x Nx int32 ray counts."""
    meta = read_gerlumph_meta(map_dir)
    n = meta["Nx"]
    counts = np.empty((n, n), np.int32)
    d = [C.c_double() for _ in range(6)]
    nx = C.c_int()
    rc = lib().mlc_gerlumph_read(os.fsencode(map_dir), C.byref(d[0]), C.byref(d[1]), C.byref(nx),
                                 C.byref(d[2]), C.byref(d[3]), C.byref(d[4]), C.byref(d[5]),
                                 ptr(counts), counts.size)
    if rc:
        raise OSError(f"{map_dir}: {_READ_ERR.get(rc, rc)}")
    return meta, counts


def write_gerlumph(map_dir, counts, avgmu, avgN, width, k, g, s=0.0):
    """Write a square int32 count map as a GERLUMPH directory (created if
    needed)."""
    counts = np.ascontiguousarray(counts, np.int32)
    if counts.ndim != 2 or counts.shape[0] != counts.shape[1]:
        raise ValueError("GERLUMPH maps are square")
    os.makedirs(map_dir, exist_ok=True)
    rc = lib().mlc_gerlumph_write(os.fsencode(map_dir), avgmu, avgN, counts.shape[0], width, k, g, s,
                                  ptr(counts))
    if rc:
        raise OSError(f"{map_dir}: cannot write map ({rc})")


def magnification(meta, counts):
    """Host FP64 magnifications of the counts (magnification_map.cpp:34-37):
    counts * |avgmu/avgN| / |1/((1-k)^2 - g^2)|."""
    factor = abs(meta["avgmu"] / meta["avgN"])
    muth = abs(1.0 / ((1.0 - meta["k"]) ** 2 - meta["g"] ** 2))
    return np.asarray(counts, np.float64) * factor / muth


def load_map(ctx, map_dir):
    """MagnificationMap(id, Rein=1) on the device: (meta, FP64 Ny x Nx map)
    normalised by mlc_gerlumph_normalise."""
    import torch

    from .api import gerlumph_normalise

    meta, counts = read_gerlumph(map_dir)
    dmap = gerlumph_normalise(ctx, torch.tensor(counts, device=ctx.device), meta["avgmu"], meta["avgN"],
                              meta["k"], meta["g"])
    return meta, dmap


def write_fits(path, image, sampling=1):
    """Image::writeImageFITS: every sampling-th pixel, rows flipped, FLOAT_IMG
    in the primary HDU and in the extension NEW-EXTENSION."""
    a = np.ascontiguousarray(image, np.float64)
    if lib().mlc_write_fits(os.fsencode(path), ptr(a), a.shape[1], a.shape[0], int(sampling)):
        raise OSError(f"{path}: cannot write FITS")


def read_fits_image(path, hdu=0):
    """A 2-D FITS image HDU (0 = primary) as an (Ny, Nx) FP64 array in the
    orientation write_fits was given (the optional FITS map input)."""
    nx, ny = C.c_int(), C.c_int()
    rc = lib().mlc_read_fits(os.fsencode(path), int(hdu), None, 0, C.byref(nx), C.byref(ny))
    if rc not in (0, 3):
        raise OSError(f"{path}: no readable 2-D image in HDU {hdu} (status {rc})")
    a = np.empty((ny.value, nx.value))
    if lib().mlc_read_fits(os.fsencode(path), int(hdu), ptr(a), a.size, C.byref(nx), C.byref(ny)):
        raise OSError(f"{path}: cannot read HDU {hdu}")
    return a


def load_fits_map(ctx, path, hdu=0):
    """A magnification map from a FITS image, as an FP64 device tensor."""
    import torch

    return torch.tensor(read_fits_image(path, hdu), device=ctx.device)


def write_bin(path, image, sampling=1):
    """Image::writeImageBIN: every sampling-th pixel as native FP64."""
    a = np.ascontiguousarray(image, np.float64)
    if lib().mlc_write_bin(os.fsencode(path), ptr(a), a.shape[1], a.shape[0], int(sampling)):
        raise OSError(f"{path}: cannot write image")


_FITS_DTYPE = {8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}


def read_fits(path):
    """The 2-D image HDUs of a FITS file (basic FITS: fixed-format headers,
    BITPIX 8/16/32/64/-32/-64, no scaling keywords) as a list of
    (header dict, array) in file order, arrays in stored row order."""
    with open(path, "rb") as f:
        raw = f.read()
    out, pos = [], 0
    while pos + 2880 <= len(raw):
        hdr = {}
        while True:
            block = raw[pos:pos + 2880]
            pos += 2880
            done = False
            for i in range(0, 2880, 80):
                card = block[i:i + 80].decode("ascii")
                key = card[:8].strip()
                if key == "END":
                    done = True
                    break
                if card[8:10] == "= ":
                    val = card[10:].split("/")[0].strip() if not card[10:].strip().startswith("'") \
                        else card[10:].strip().split("'")[1].strip()
                    hdr[key] = val
            if done:
                break
        naxis = int(hdr.get("NAXIS", 0))
        shape = [int(hdr[f"NAXIS{i}"]) for i in range(naxis, 0, -1)]
        n = int(np.prod(shape)) if naxis else 0
        dt = np.dtype(_FITS_DTYPE[int(hdr["BITPIX"])])
        data = np.frombuffer(raw, dt, n, pos).reshape(shape) if n else None
        pos += (n * dt.itemsize + 2879) // 2880 * 2880
        out.append((hdr, data))
    return out
