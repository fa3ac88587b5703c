"""Observed multi-band, multi-image light curves: the reference's data.csv
(repository root; header "MJD,Filter,Flux1,σ1,...,Flux4,σ4" with a UTF-8 byte
order mark, CRLF lines, uneven epochs per filter), turned into the per-image
magnitude observations that mlc_lc_chi2_grad takes (one observation set per
image, one band per filter)."""
import csv
import io

import numpy as np


class ObservedData:
    """Fluxes per image and filter.

    filters: filter names in order of first appearance;
    t[f]: MJD of the epochs of filter f, ascending;
    flux[i][f], sigma[i][f]: image i's fluxes and errors at those epochs."""

    def __init__(self, filters, t, flux, sigma):
        self.filters, self.t, self.flux, self.sigma = filters, t, flux, sigma
        self.nimages = len(flux)

    def magnitudes(self, i, f):
        """(mag, sigma_mag) of image i in filter f: m = -2.5 log10(flux),
        sigma_m = 2.5/ln(10) sigma/flux.  Non-positive fluxes give NaN."""
        fl = np.asarray(self.flux[i][f], np.float64)
        sg = np.asarray(self.sigma[i][f], np.float64)
        with np.errstate(divide="ignore", invalid="ignore"):
            m = np.where(fl > 0, -2.5 * np.log10(np.where(fl > 0, fl, 1.0)), np.nan)
            sm = np.where(fl > 0, 2.5 / np.log(10.0) * sg / np.where(fl > 0, fl, 1.0), np.nan)
        return m, sm


def load_csv(path):
    """Parse a data.csv file."""
    with open(path, "rb") as f:
        return parse_csv(f.read().decode("utf-8-sig"))


def parse_csv(text):
    """Parse the text of a data.csv file (a leading byte order mark is ignored)."""
    text = text.lstrip("\ufeff")
    rows = [r for r in csv.reader(io.StringIO(text)) if r and any(c.strip() for c in r)]
    head = [c.strip() for c in rows[0]]
    if len(head) < 4 or head[0] != "MJD" or head[1] != "Filter" or (len(head) - 2) % 2:
        raise ValueError(f"unexpected header {head}")
    nimg = (len(head) - 2) // 2
    filters, per = [], {}
    for r in rows[1:]:
        if len(r) != len(head):
            raise ValueError(f"row with {len(r)} fields, header has {len(head)}: {r}")
        mjd, flt = float(r[0]), r[1].strip()
        vals = [float(v) for v in r[2:]]
        if flt not in per:
            filters.append(flt)
            per[flt] = []
        per[flt].append((mjd, vals))
    t, flux, sigma = {}, [dict() for _ in range(nimg)], [dict() for _ in range(nimg)]
    for flt in filters:
        recs = sorted(per[flt], key=lambda e: e[0])  # stable: equal MJDs keep file order
        t[flt] = np.array([e[0] for e in recs])
        for i in range(nimg):
            flux[i][flt] = np.array([e[1][2 * i] for e in recs])
            sigma[i][flt] = np.array([e[1][2 * i + 1] for e in recs])
    return ObservedData(filters, t, flux, sigma)


def observations(ctx, data, images=None, filters=None, t0=None):
    """lightcurves.Observations with one observation set per image and one
    band per filter, magnitudes from the fluxes, times in days from t0
    (default: the earliest MJD).  Epochs with non-positive flux are dropped."""
    from .lightcurves import Observations

    images = list(range(data.nimages)) if images is None else list(images)
    filters = list(data.filters) if filters is None else list(filters)
    if t0 is None:
        t0 = min(float(data.t[f].min()) for f in filters)
    sets = []
    for i in images:
        bands = []
        for f in filters:
            m, sm = data.magnitudes(i, f)
            keep = np.isfinite(m)
            bands.append((data.t[f][keep] - t0, m[keep], sm[keep]))
        sets.append(bands)
    return Observations(ctx, len(filters), sets)
