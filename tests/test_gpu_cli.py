"""GPU runs of the command-line host: `fbls search` (raw and --prepared) and
`fbls batch` on NPZ files written by numpy, against the library's own
search calls on the same observations.

The report prints every double as its shortest round-trip text, so the
candidates, single events and statistics read back from the JSON equal the
library's values exactly; the period-table CSV holds one row per
qualifying period.  Batch mode writes one report per star (the same as a
single search with that star's pmax) and an .error file for a star whose
input is missing."""
import json
import math
import os
import subprocess

import numpy as np
import pytest

import fbls_amd as F

pytestmark = pytest.mark.gpu


def _star(seed, period=3.7):
    t, f, e = F.synth_star(seed=seed, sectors=1, sigma=1e-3, period=period, depth=2e-3,
                           duration=2.5 / 24)
    slow = 1 + 0.004 * np.sin(2 * np.pi * (t - t[0]) / 6.0)
    return t, 12000.0 * f * slow, 12000.0 * e


def _run(args, timeout=300):
    print("[cli]", " ".join(os.path.basename(a) for a in args), flush=True)
    r = subprocess.run([F.CLI_PATH] + args, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stderr
    return r


def _same_peak(j, p):
    for k, v in j.items():
        want = p[k]
        if v is None:
            assert isinstance(want, float) and math.isnan(want), (k, v, want)
        else:
            assert v == want, (k, v, want)


def test_cli_search_raw_matches_library(gpu, tmp_path):
    t, raw, err = _star(31)
    a = str(tmp_path / "a.npz")
    np.savez(a, time_bjd=t, flux=raw, flux_err=err, sector=np.int64(7))
    out, csv = str(tmp_path / "a.json"), str(tmp_path / "a.csv")
    _run(["search", "--input", a, "--output", out, "--save-table", csv, "--pmax", "8"])
    rep = json.load(open(out))
    assert rep["schema_version"] == 3 and rep["input_mode"] == "raw"
    assert rep["input_files"] == [a] and "period_table" not in rep
    cfg = F.SearchConfig(period_max=8.0)
    peaks, passes, lib = F.search_transits(gpu, t, raw, err, np.zeros(len(t), np.int64), cfg,
                                           report=True)
    assert [j["period_days"] for j in rep["signal_candidates"]] == [
        p["period_days"] for p in F.by_signal(peaks)]
    for j, p in zip(rep["candidates"], F.by_model(peaks)):
        _same_peak(j, p)
    assert rep["single_events"] == lib["events"]
    st = rep["statistics"]
    assert st["resolved_peaks"] == len(peaks) and st["observations"] == lib["observations"]
    assert st["periods_tested"] == lib["periods_tested"]
    assert [p["rows"] for p in st["passes"]] == [p["rows"] for p in passes]
    assert len(st["passes"][0]["families"]) == lib["nfam"] > 1
    assert set(st["single_event_scan"]) == {str(k) for k in range(lib["nfam"])}
    lines = open(csv).read().splitlines()
    assert lines[0] == "period_days,epoch_bjd,duration_days,snr,depth"
    assert len(lines) - 1 == st["qualifying_periods"] == passes[0]["rows"]
    assert abs(rep["candidates"][0]["period_days"] / 3.7 - 1) < 1e-3


def test_cli_search_prepared(gpu, tmp_path):
    t, f, e = F.synth_star(seed=32, sectors=1, sigma=8e-4, period=2.9, depth=2.5e-3,
                           duration=2.0 / 24)
    a = str(tmp_path / "p.npz")
    np.savez_compressed(a, time_bjd=t, flux=f - 1.0, flux_err=e)
    r = _run(["search", "--input", a, "--prepared", "--pmin", "0.5", "--pmax", "6"])
    rep = json.loads(r.stdout)
    cfg = F.SearchConfig(period_min=0.5, period_max=6.0)
    peaks, _, lib = F.search_prepared(gpu, t, f - 1.0, e, cfg, report=True)
    for j, p in zip(rep["candidates"], F.by_model(peaks)):
        _same_peak(j, p)
    assert rep["input_mode"] == "prepared"
    fam = rep["statistics"]["passes"][0]["families"][0]
    assert fam["window_days"] is None and fam["observations"] == len(t)


def test_cli_batch(gpu, tmp_path):
    files = []
    for k, seed in enumerate((33, 34)):
        t, raw, err = _star(seed, period=2.6 + k)
        p = str(tmp_path / f"s{k}.npz")
        np.savez(p, time_bjd=t, flux=raw, flux_err=err)
        files.append(p)
    # an archive whose time array declares 2^61 + 1 doubles (its byte count
    # wraps in 64 bits): that star fails alone, the batch goes on
    import io
    import zipfile
    h = b"{'descr': '<f8', 'fortran_order': False, 'shape': (2305843009213693953,), }"
    h += b" " * (63 - (10 + len(h)) % 64) + b"\n"
    huge = str(tmp_path / "huge.npz")
    with zipfile.ZipFile(huge, "w") as z:
        buf = io.BytesIO()
        np.lib.format.write_array(buf, np.arange(10.0))
        z.writestr("flux.npy", buf.getvalue())
        z.writestr("time_bjd.npy", b"\x93NUMPY\x01\x00" + len(h).to_bytes(2, "little") + h
                   + b"\0" * 8)
    lst = tmp_path / "stars.txt"
    lst.write_text(f"# name pmax files\nstar0 7 {files[0]}\nstar1 - {files[1]}\n"
                   f"broken 5 {tmp_path / 'missing.npz'}\nhuge 5 {huge}\nshort 5\n")
    outdir = str(tmp_path / "out")
    r = _run(["batch", "--list", str(lst), "--output-dir", outdir, "--workers", "2"], 600)
    assert "4 stars in" in r.stderr
    errors = {f: open(os.path.join(outdir, f)).read() for f in os.listdir(outdir)
              if f.endswith(".error")}
    assert sorted(os.listdir(outdir)) == ["broken.error", "huge.error", "star0.json",
                                          "star1.json"], (r.stderr, errors)
    assert "too large" in errors["huge.error"]
    for name, pmax, path in (("star0", "7", files[0]), ("star1", "20", files[1])):
        got = json.load(open(os.path.join(outdir, name + ".json")))
        assert got["config"]["period_max"] == float(pmax)
        assert len(got["period_table"]) == got["statistics"]["qualifying_periods"]
        single = str(tmp_path / (name + ".single.json"))
        _run(["search", "--input", path, "--pmax", pmax, "--output", single])
        want = json.load(open(single))
        for key in ("candidates", "signal_candidates", "single_events", "config"):
            assert got[key] == want[key], key
