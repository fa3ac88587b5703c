"""Host-side inputs and outputs around the path, CPU only: GERLUMPH map
directories (magnification_map.cpp:7-40), the FITS and raw image writers
(image.cpp:4-62), the source size laws (profile.cpp:31-124) and the
observed-data loader for data.csv (the reference's own file, kept as the
fixture tests/golden/data.csv)."""
import os

import numpy as np
import pytest

from microsn_amd import api, mapio, observed

HERE = os.path.dirname(os.path.abspath(__file__))


# ---------------------------------------------------------------- GERLUMPH maps
def test_gerlumph_round_trip_and_magnification(tmp_path):
    rng = np.random.default_rng(0)
    counts = rng.integers(0, 1000, (37, 37)).astype(np.int32)
    d = str(tmp_path / "12123")
    mapio.write_gerlumph(d, counts, avgmu=1.25, avgN=211.5, width=25.0, k=0.4, g=0.35, s=0.0)
    meta, back = mapio.read_gerlumph(d)
    assert np.array_equal(back, counts)
    assert meta == dict(avgmu=1.25, avgN=211.5, Nx=37, width=25.0, k=0.4, g=0.35, s=0.0)
    # metadata as the reference writes it: "avgmu avgN / Nx / width / k g s"
    with open(os.path.join(d, "mapmeta.dat")) as f:
        assert f.read().split() == ["1.25", "211.5", "37", "25", "0.40000000000000002",
                                    "0.34999999999999998", "0"]
    mu = mapio.magnification(meta, counts)
    mu_th = 1.0 / ((1 - 0.4) ** 2 - 0.35 ** 2)
    assert np.array_equal(mu, counts * abs(1.25 / 211.5) / abs(mu_th))


def test_gerlumph_read_reports_missing_and_short_files(tmp_path):
    with pytest.raises(OSError, match="mapmeta"):
        mapio.read_gerlumph(str(tmp_path / "none"))
    d = tmp_path / "m"
    d.mkdir()
    (d / "mapmeta.dat").write_text("1 2\n8\n10\n0.3 0.3 0\n")
    with pytest.raises(OSError, match="map.bin"):
        mapio.read_gerlumph(str(d))
    (d / "map.bin").write_bytes(np.zeros(63, np.int32).tobytes())  # one short
    with pytest.raises(OSError, match="map.bin"):
        mapio.read_gerlumph(str(d))
    (d / "map.bin").write_bytes(np.arange(64, dtype=np.int32).tobytes())
    meta, c = mapio.read_gerlumph(str(d))
    assert meta["Nx"] == 8 and np.array_equal(c.ravel(), np.arange(64))
    (d / "mapmeta.dat").write_text("1 2\n")
    with pytest.raises(OSError, match="mapmeta"):
        mapio.read_gerlumph(str(d))


# ---------------------------------------------------------------- images
def test_fits_layout_flip_and_sampling(tmp_path):
    img = np.arange(6 * 10, dtype=np.float64).reshape(6, 10) / 7.0
    p = str(tmp_path / "m.fits")
    mapio.write_fits(p, img, sampling=2)
    raw = open(p, "rb").read()
    assert len(raw) % 2880 == 0
    hdus = mapio.read_fits(p)
    assert len(hdus) == 2
    (h0, a0), (h1, a1) = hdus
    assert h0["SIMPLE"] == "T" and h0["BITPIX"] == "-32" and h0["NAXIS1"] == "5" and h0["NAXIS2"] == "3"
    assert h1["XTENSION"] == "IMAGE" and h1["EXTNAME"] == "NEW-EXTENSION"
    sub = img[::2, ::2][:3, :5].astype(np.float32)
    assert a0.dtype == np.dtype(">f4") and np.array_equal(a0, sub[::-1])  # first stored row = last row
    assert np.array_equal(a1, a0)


def _fits_file(path, bitpix, arr, extra=(), naxis3=None):
    """A one-HDU FITS file written by hand: fixed-format cards, big-endian
    data, both padded to 2880 bytes."""
    def card(k, v):
        return (f"{k:<8}= {v:>20}").ljust(80).encode()
    ny, nx = arr.shape
    h = card("SIMPLE", "T") + card("BITPIX", str(bitpix))
    h += card("NAXIS", "3" if naxis3 else "2") + card("NAXIS1", str(nx)) + card("NAXIS2", str(ny))
    if naxis3:
        h += card("NAXIS3", str(naxis3))
    for k, v in extra:
        h += card(k, v)
    h += b"END".ljust(80)
    h += b" " * (-len(h) % 2880)
    dt = {16: ">i2", 32: ">i4", -32: ">f4", -64: ">f8", 8: ">u1"}[bitpix]
    d = np.ascontiguousarray(arr).astype(dt).tobytes()
    d += b"\0" * (-len(d) % 2880)
    with open(path, "wb") as f:
        f.write(h + d)


def test_fits_reader_round_trip_and_formats(tmp_path):
    """mlc_read_fits: both HDUs of the writer's files come back in the
    writer's orientation; hand-made files with integer BITPIX and
    BSCALE/BZERO, FP64 data, and malformed files."""
    img = np.random.default_rng(2).random((12, 20)) * 50
    p = str(tmp_path / "m.fits")
    mapio.write_fits(p, img, sampling=2)
    want = img[::2, ::2][:6, :10].astype(np.float32).astype(np.float64)
    assert np.array_equal(mapio.read_fits_image(p, 0), want)
    assert np.array_equal(mapio.read_fits_image(p, 1), want)
    with pytest.raises(OSError):
        mapio.read_fits_image(p, 2)
    stored = np.arange(35, dtype=np.int16).reshape(5, 7) - 17
    q = str(tmp_path / "i16.fits")
    _fits_file(q, 16, stored, extra=(("BSCALE", "0.5"), ("BZERO", "100.0")))
    assert np.array_equal(mapio.read_fits_image(q), (100.0 + 0.5 * stored.astype(np.float64))[::-1])
    d64 = np.random.default_rng(3).standard_normal((4, 9))
    q = str(tmp_path / "f64.fits")
    _fits_file(q, -64, d64[::-1])
    assert np.array_equal(mapio.read_fits_image(q), d64)
    q = str(tmp_path / "cube.fits")
    _fits_file(q, -32, np.zeros((4, 4)), naxis3=1)
    with pytest.raises(OSError):
        mapio.read_fits_image(q)
    raw = open(p, "rb").read()
    q = str(tmp_path / "short.fits")
    open(q, "wb").write(raw[:2880 + 100])
    with pytest.raises(OSError):
        mapio.read_fits_image(q)
    with pytest.raises(OSError):
        mapio.read_fits_image(str(tmp_path / "missing.fits"))


def test_bin_writer(tmp_path):
    img = np.random.default_rng(1).random((9, 12))
    p = str(tmp_path / "m.bin")
    mapio.write_bin(p, img, sampling=3)
    back = np.fromfile(p, np.float64).reshape(3, 4)
    assert np.array_equal(back, img[::3, ::3])


# ---------------------------------------------------------------- source sizes
def test_size_laws_closed_forms():
    f32 = lambda v: float(np.float32(v))
    assert api.get_size({"type": "parametric", "r0": "2.5", "l0": "400", "nu": "1.333"}, 400.0) == f32(2.5)
    r = api.get_size({"type": "parametric", "r0": "2.5", "l0": "400", "nu": "1.333"}, 800.0)
    assert abs(r - f32(2.5) * 2.0 ** f32(1.333)) < 1e-12 * r
    r = api.get_size({"type": "ss_disc", "mbh": "1e9", "fedd": "0.1", "eta": "0.1"}, 500.0)
    want = 0.0097 * (500.0 ** 4 * f32(1e9) ** 2 * (f32(0.1) / f32(0.1))) ** (1.0 / 3.0)
    assert abs(r - want) < 1e-12 * want
    assert api.get_size({"type": "vector"}, 3.7) == 3.7
    assert api.get_size({"type": "custom"}, 3.7) == 1.0
    assert api.get_size({"type": "other"}, 3.7) == 0.0


def test_size_laws_over_a_filter_curve():
    wl = np.linspace(350.0, 450.0, 41)
    tp = np.exp(-0.5 * ((wl - 400.0) / 20.0) ** 2)
    trap = lambda y: float(np.sum((y[1:] + y[:-1]) * np.diff(wl)) / 2.0)
    r = api.get_size({"type": "vector"}, wavelength=wl, throughput=tp)
    assert abs(r - trap(tp * wl) / trap(tp)) < 1e-12 * r
    f32 = lambda v: float(np.float32(v))
    pars = {"type": "parametric", "r0": "1.2", "l0": "400", "nu": "1.5"}
    r = api.get_size(pars, wavelength=wl, throughput=tp)
    law = f32(1.2) * (wl / f32(400.0)) ** f32(1.5)
    assert abs(r - trap(tp * law) / trap(tp)) < 1e-12 * r
    pars = {"type": "ss_disc", "mbh": "3e8", "fedd": "0.2", "eta": "0.1"}
    r = api.get_size(pars, wavelength=wl, throughput=tp)
    law = 0.0097 * (wl ** 4 * f32(3e8) ** 2 * (f32(0.2) / f32(0.1))) ** (1.0 / 3.0)
    assert abs(r - trap(tp * law) / trap(tp)) < 1e-12 * r


# ---------------------------------------------------------------- observed data
def test_data_csv_fixture():
    d = observed.load_csv(os.path.join(HERE, "golden", "data.csv"))
    assert d.nimages == 4
    assert d.filters == ["F390W", "F475W", "F625W", "F814W", "F110W", "F160W"]
    assert sum(len(d.t[f]) for f in d.filters) == 41
    assert [len(d.t[f]) for f in d.filters] == [1, 2, 9, 9, 10, 10]
    # first data row: 57685.952,F390W,0.724,0.126,0.352,0.126,0.123,0.126,0.01,0.126
    assert d.t["F390W"][0] == 57685.952
    assert [d.flux[i]["F390W"][0] for i in range(4)] == [0.724, 0.352, 0.123, 0.01]
    assert all(d.sigma[i]["F390W"][0] == 0.126 for i in range(4))
    for f in d.filters:
        assert np.all(np.diff(d.t[f]) >= 0)
    m, sm = d.magnitudes(0, "F625W")
    fl, sg = d.flux[0]["F625W"], d.sigma[0]["F625W"]
    assert np.allclose(m, -2.5 * np.log10(fl), rtol=0, atol=1e-14)
    assert np.allclose(sm, 2.5 / np.log(10) * sg / fl, rtol=1e-14)


def test_data_csv_text_forms():
    text = "﻿MJD,Filter,Flux1,σ1\r\n2.0,B,1.5,0.1\r\n1.0,B,2.0,0.2\r\n3.0,A,-1.0,0.3\r\n"
    d = observed.parse_csv(text)
    assert d.filters == ["B", "A"] and d.nimages == 1
    assert np.array_equal(d.t["B"], [1.0, 2.0]) and np.array_equal(d.flux[0]["B"], [2.0, 1.5])
    m, sm = d.magnitudes(0, "A")
    assert np.isnan(m[0]) and np.isnan(sm[0])
    with pytest.raises(ValueError):
        observed.parse_csv("Time,Filter,F1,s1\n1,B,1,1\n")
    with pytest.raises(ValueError):
        observed.parse_csv("MJD,Filter,Flux1,s1\n1,B,1\n")
