"""The C ABI without a GPU: libmlc.so loads, exports every function that
include/mlc.h and include/mlc_host.h declare, the ctypes signature table
matches the headers, and the host-only entry points validate their input.

No device work runs here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from microsn_amd import _lib, api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("mlc.h", "mlc_host.h")]

_DECL = re.compile(r"^(?:const\s+)?(?:int|void|double|char|int64_t)\s*\*?\s*(mlc_\w+)\s*\(", re.M)


def declared():
    names = []
    for h in HEADERS:
        with open(h) as f:
            names += _DECL.findall(f.read())
    return names


def test_headers_declare_functions():
    names = declared()
    assert len(names) == len(set(names)), "duplicate declarations"
    assert "mlc_lc_chi2_grad" in names and "mlc_locations_on_grid" in names
    assert len(names) >= 40


def test_library_loads_and_exports_every_declared_symbol():
    L = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_signature_table_matches_headers():
    assert set(_lib.SIGNATURES) == set(declared())


def test_argument_counts_match_headers():
    """Each ctypes argtypes list has as many entries as the C prototype."""
    src = "".join(open(h).read() for h in HEADERS)
    for name, (_, args) in _lib.SIGNATURES.items():
        m = re.search(r"\b" + name + r"\s*\(([^)]*)\)", src, re.S)
        assert m, name
        params = m.group(1).strip()
        n = 0 if params in ("", "void") else params.count(",") + 1
        assert n == len(args), (name, n, len(args))


def test_version_and_host_only_helpers():
    L = _lib.lib()
    assert L.mlc_version() >= 1
    # std::stof(std::to_string(x)): six decimals, then float
    assert L.mlc_parity_round(0.05) == float(np.float32(0.05))
    assert L.mlc_parity_round(1.23456789) == float(np.float32(1.234568))
    assert L.mlc_parity_round(4e-7) == 0.0  # "0.000000"
    # profile sizes of grid_input.json (sizes 0.05..0.3 R_E, 25 R_E over 10000 px)
    pix = L.mlc_parity_round(25.0 / 10000)
    for shape, want in [(1, [82, 162, 322, 482]), (0, [64, 120, 234, 346]),
                        (5, [82, 162, 322, 482])]:
        got = [L.mlc_profile_size(shape, L.mlc_parity_round(s), pix, 0.0, 1.0)
               for s in (0.05, 0.1, 0.2, 0.3)]
        assert got == want, (shape, got)
    assert L.mlc_profile_size(17, 1.0, 1.0, 0.0, 1.0) < 0  # unknown shape
    assert L.mlc_profile_size(1, -1.0, 1.0, 0.0, 1.0) < 0  # negative radius


def test_last_error_without_context():
    msg = _lib.lib().mlc_last_error(None)
    assert isinstance(msg, bytes)


def test_host_entry_points_reject_bad_arguments():
    with pytest.raises(ValueError):
        api.locations_on_grid(0, 100, 10)
    with pytest.raises(ValueError):
        api.locations_on_grid(10, 100, 60)  # Nmap - 2 Noff <= 0
    with pytest.raises(ValueError):
        api.locations_random(10, 100, 60, 1)
    with pytest.raises(ValueError):
        api.mpd_bins(0, 1.0, 2.0)
    with pytest.raises(ValueError):
        api.mpd_bins(10, 0.0, 2.0, True)  # log bins need min > 0
    with pytest.raises(ValueError):
        api.mpd_bins(10, 3.0, 2.0)
    with pytest.raises(ValueError):  # no size grid
        api.setup_chi2_calculation((0.4, 0.5), (1.0,), 1.0, [1.0], [1.0], [0.1], [])
