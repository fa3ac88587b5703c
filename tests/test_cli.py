"""The microlc driver (driver.cpp:19-270 on top of libmlc.so): argument and
input errors, all of which are caught before any device work.  CPU only;
the end-to-end run is in test_gpu_cli.py."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "microsn_amd", "microlc")

GRID_INPUT = {
    "sizes": [0, 0.05, 0.1, 0.2, 0.3], "shape": "gaussian",
    "locations": {"type": "grid", "Nmap": 10000, "ss": 25, "Nside": 99},
    "mpds": {"mpd": {"Nbins": 500, "min": 0.0001, "max": 400},
             "ratio": {"Nbins": 300, "min": 0.0001, "max": 300}},
    "DL": 1570, "DS": 1750, "DLS": 850,
    "data": [{"d": [0.7, 1.2, 1.4, 1.1, 0.9], "s": [0.1, 0.15, 0.1, 0.1, 0.15], "t": [5, 7, 12, 14, 17]}],
}


def _run(args, cwd):
    if not os.path.exists(EXE):
        pytest.skip("microlc not built (make -C microsn_amd/csrc)")
    return subprocess.run([EXE] + args, cwd=cwd, capture_output=True, text=True, timeout=60)


def test_usage_message(tmp_path):
    r = _run([], tmp_path)
    assert r.returncode == 1
    assert "You need to provide the name of a .json file" in r.stdout


def test_missing_and_malformed_input(tmp_path):
    r = _run([str(tmp_path / "none.json")], tmp_path)
    assert r.returncode == 1 and "cannot open" in r.stderr
    bad = tmp_path / "bad.json"
    bad.write_text('{"sizes": [0, 0.1,], }')
    r = _run([str(bad)], tmp_path)
    assert r.returncode == 1 and "JSON" in r.stderr and "line 1" in r.stderr


@pytest.mark.parametrize("edit, message", [
    (lambda j: j["data"][0].update(d=[1.0] * 50, s=[0.1] * 50, t=list(range(50))),
     "too large to fit in register memory"),
    (lambda j: j.update(shape="square"), "unknown profile shape"),
    (lambda j: j.update(precision="f16"), "precision"),
    (lambda j: j.pop("DL"), 'missing key "DL"'),
    (lambda j: j["locations"].update(Nside=2.5), "integer expected"),
])
def test_input_errors_exit_nonzero(tmp_path, edit, message):
    j = json.loads(json.dumps(GRID_INPUT))
    edit(j)
    p = tmp_path / "in.json"
    p.write_text(json.dumps(j))
    r = _run([str(p)], tmp_path)
    assert r.returncode == 1, r.stdout + r.stderr
    assert message in r.stderr
    # the reference's progress lines come first, in its format
    if message != 'missing key "DL"' and message != "integer expected":
        assert r.stdout.startswith("Number of profiles set to: ") or message == "unknown profile shape"
