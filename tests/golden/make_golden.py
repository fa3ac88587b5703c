"""Writes tests/golden/locations_random.json: locations_random
(locations.cpp:17-25) of the reference, compiled from its own sources into
oracle/_ref/libmlcref.so by oracle/Makefile, on a few seeds.  Its integer
distribution is libstdc++'s std::uniform_int_distribution, which only a
build of the reference's own source pins; the fixture keeps that pin where
the reference tree is absent.

Usage: make -C oracle ref && python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as orc  # noqa: E402

CASES = [(64, 10000, 120, 1), (64, 10000, 120, 42), (200, 1000, 30, 7), (16, 50, 0, 123456)]


def main():
    r = orc.ref()
    if r is None:
        sys.exit("oracle/_ref/libmlcref.so is not built (needs the reference tree)")
    out = []
    for Nloc, Nmap, Noff, seed in CASES:
        x = np.zeros(Nloc, np.int32)
        y = np.zeros(Nloc, np.int32)
        r.ref_locations_random(Nloc, Nmap, Noff, seed, orc._p(x), orc._p(y))
        out.append(dict(Nloc=Nloc, Nmap=Nmap, Noff=Noff, seed=seed, x=x.tolist(), y=y.tolist()))
    with open(os.path.join(HERE, "locations_random.json"), "w") as f:
        json.dump({"source": "locations.cpp:17-25 built from the reference sources", "cases": out}, f)


if __name__ == "__main__":
    main()
