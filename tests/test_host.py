"""CPU tests of the product's host side: the C ABI loads and exports every
declared symbol, and the host helpers (grid, bands, durations, prepared-curve
noise model, fixed point, synthetic generator) agree with the oracle."""
import json
import math
import os
import re

import numpy as np
import pytest

import fbls_amd as F
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def test_abi_exports_every_declared_symbol():
    text = open(os.path.join(ROOT, "include", "fbls.h")).read()
    declared = set(re.findall(r"\b(fbls_[a-z_0-9]+)\s*\(", text))
    assert declared, "no declarations found"
    lib = F.lib()
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert declared == set(F.EXPORTS)


def test_period_grid_matches_oracle():
    for args in [(0.5, 20.0, 50.0, 0.02, 3.0), (0.5, 13.7, 27.4, 0.5 / 24, 3.0),
                 (0.3, 100.0, 730.0, 0.5 / 24, 3.0)]:
        np.testing.assert_array_equal(F.period_grid(*args), O.period_grid(*args))


def test_period_grid_concurrent_callers():
    """The grid is computed in pieces on the library's shared host threads;
    many callers at once (as the batch pipeline's threads are) each get the
    oracle's grid, element for element."""
    import concurrent.futures as cf

    cases = [(0.5, 20.0, 109.6, 0.5 / 24, 3.0), (0.3, 13.7, 27.4, 1.0 / 24, 5.0),
             (0.5, 100.0, 1077.0, 0.5 / 24, 3.0), (1.0, 2.0, 30.0, 0.1, 1.0)]
    want = [O.period_grid(*c) for c in cases]

    def one(k):
        return np.array_equal(F.period_grid(*cases[k % len(cases)]), want[k % len(cases)])

    with cf.ThreadPoolExecutor(12) as ex:
        assert all(ex.map(one, range(96)))


def test_durations_days():
    cfg = F.SearchConfig()
    np.testing.assert_array_equal(cfg.durations_days(),
                                  O.durations_days(cfg.durations_hours))
    assert len(cfg.durations_days()) == 13


def test_band_plan_matches_oracle_and_report():
    d = json.load(open(os.path.join(GOLDEN, "search_bands_72668830.json")))
    cfg_hours = d["config"]["durations_hours"]
    for f in d["passes"][0]["families"]:
        hrs = sorted(min(cfg_hours, key=lambda c: abs(c - h))
                     for h in f["durations_hours"])
        days = F.durations_days(hrs)
        pmin = max(0.3, 4.0 * days[0] * (1.0 + 1e-9))
        step = math.log(100.0 / pmin) / (f["periods"] - 1.5)
        baseline = days[0] / (3.0 * math.expm1(step))
        grid = F.period_grid(pmin, 100.0, baseline, days[0], 3.0)
        assert len(grid) == f["periods"]
        mine = F.plan_bands(grid, days, 4, 2)
        ref = O.plan_bands(grid, days, 4, 2)
        assert mine == ref
        assert [b["count"] for b in mine] == [b["periods"] for b in f["bands"]]


def test_bank_batches_formula():
    for np_, nb in [(289901, 58), (254845, 14746), (1, 19200), (0, 10)]:
        assert F.bank_batches(np_, nb) == O.bank_batches(np_, nb, 4096, 1024.0)
    assert F.bank_batches(10, 1 << 20, 4096, 1.0) == 0  # over budget


def test_family_state_matches_oracle():
    from fbls_amd import synth

    for maker in (synth.single_sector, synth.four_sector_red):
        t, f, e = maker()
        d = F.SearchConfig().durations_days()
        s = F.family_state(t, f - 1.0, e, d)
        o = O.family_state(t, f - 1.0, e, d)
        np.testing.assert_array_equal(s.t, o["t"])
        np.testing.assert_array_equal(s.v, o["v"])
        np.testing.assert_array_equal(s.w, o["w"])
        np.testing.assert_array_equal(s.rho, o["rho"])
        assert s.baseline_norm == o["baseline_norm"]
        assert s.origin == o["origin"]
        assert s.inflation == o["inflation"] and s.segments == o["segments"]


def test_family_state_kat_and_errors():
    """curve.rs:636-661 through the product's prepared-curve entry, plus the
    reference's data errors (curve.rs:114-126, 429-431)."""
    n = 500
    t = 1000.0 + np.arange(n) * 0.01
    f = 2000.0 + 3.0 * np.sin(np.arange(n) * 1.7)
    s = F.family_state(t, f - 1.0, np.ones(n), [0.05])
    assert s.origin == 1000.0 and s.t[0] == 0.0
    assert s.inflation > 2.0 and s.segments == 1
    with pytest.raises(F.FblsError) as ei:
        F.family_state(t[:5], f[:5] - 1.0, np.ones(5), [0.05])
    assert ei.value.code == 6
    with pytest.raises(F.FblsError):
        F.family_state(t, -np.ones(n) * 3.0, np.ones(n), [0.05])  # median <= 0


def test_family_states_many_equals_one_at_a_time():
    """fbls_family_state_many (the library's preparation threads) gives each
    star exactly fbls_family_state's result: normalised fluxes with the -1
    offset, unsorted times with duplicates and non-finite points, the
    uncorrelated path, and a failing star that names itself while the others
    are still prepared."""
    from fbls_amd import synth

    d = F.SearchConfig().durations_days()
    rng = np.random.default_rng(3)
    curves = [synth.single_sector(), synth.four_sector_red()]
    for k in range(6):
        n = 300 + 700 * k
        t = np.sort(rng.uniform(0.0, 20.0 + k, n))
        if k % 2:
            t = rng.permutation(np.concatenate([t, t[: n // 10]]))  # unsorted, duplicates
        f = 1.0 + 1e-3 * rng.standard_normal(len(t))
        e = np.full(len(t), 1e-3 * (1 + k))
        if k == 3:
            f[::17] = np.nan
        curves.append((t, f, e))
    for corr in (True, False):
        many = F.family_states(curves, d, corr, flux_offset=-1.0)
        for (t, f, e), b in zip(curves, many):
            a = F.family_state(t, f - 1.0, e, d, corr)
            for k in ("t", "v", "w", "rho"):
                np.testing.assert_array_equal(getattr(b, k), getattr(a, k))
            for k in ("baseline_norm", "origin", "inflation", "segments", "residual_energy"):
                assert getattr(b, k) == getattr(a, k), k
    o = O.family_state(*curves[0][:1], curves[0][1] - 1.0, curves[0][2], d)
    np.testing.assert_array_equal(many[0].v, F.family_state(curves[0][0], curves[0][1] - 1.0,
                                                            curves[0][2], d, False).v)
    assert np.array_equal(F.family_states(curves[:1], d, True, -1.0)[0].rho, o["rho"])
    bad = (np.arange(4.0), np.ones(4), np.ones(4))
    with pytest.raises(F.FblsError) as ei:
        F.family_states([curves[2], bad, curves[3]], d, True, -1.0)
    assert ei.value.code == 6 and "star 1" in str(ei.value)


def test_quantize_matches_oracle():
    rng = np.random.default_rng(9)
    # extreme scales: the exponent leaves the normal range (1e-300) and
    # scaled values underflow (1e300 with a few tiny entries)
    for scale in (1e-300, 1e-3, 1.0, 1e6, 1e300):
        for spike in (0.0, 1e3, 1e7):
            x = rng.normal(0, scale, 5000)
            x[17] += spike * scale
            x[23] = 1e-310
            q1, e1, b1 = F.quantize(x)
            q2, e2, b2 = O.quantize(x)
            assert e1 == e2 and b1 == b2 == (spike > 0)
            np.testing.assert_array_equal(q1, q2)


def test_synth_uses_reference_rng():
    """The generator's white noise is the reference's Rng stream
    (survey/numeric.rs:345-364): first point of a noise-only curve."""
    t, f, e = F.synth_star(seed=77, sectors=1, sigma=1e-3, period=0.0,
                           depth=0.0, epoch=0.0)
    r = O.Rng(77)
    assert f[0] == 1.0 + r.normal(0.0, 1e-3)
    assert f[1] == 1.0 + r.normal(0.0, 1e-3)
    assert np.all(e == 1e-3)
    assert abs(len(t) - (27.4 - 1.0) * 720) < 2


def test_synth_injects_box():
    t, f, e = F.synth_star(seed=1, sectors=1, sigma=0.0, period=3.7,
                           depth=2e-3, duration=2.5 / 24, epoch=2458001.0)
    in_tr = np.abs(((t - 2458001.0 + 1.85) % 3.7) - 1.85) <= 2.5 / 48
    assert np.all(f[in_tr] == 1.0 - 2e-3) and np.all(f[~in_tr] == 1.0)


def test_config_validation():
    F.SearchConfig().validate()
    with pytest.raises(ValueError):
        F.SearchConfig(period_min=0.01).validate()  # dmin >= pmin / 4
    with pytest.raises(ValueError):
        F.SearchConfig(prescan_fraction=0.0).validate()
    with pytest.raises(ValueError):
        F.SearchConfig(ingress_fractions=[0.0, 0.5]).validate()
    assert F.SearchConfig().shape_gate == 0.6 * 7.1
    assert F.SearchConfig(ingress_fractions=[0.1]).shape_gate == 0.0


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: asking for more exits with a message
    before any GPU work."""
    import subprocess
    import sys

    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr or "one process per GPU" in r.stderr


def test_ragged_bins_argument_checks():
    """fbls_signal_search_bins_ragged refuses offsets that do not start at
    zero, a star without bins and edges that do not ascend, before any
    device work (a null context is then reported as a null argument)."""
    import ctypes

    L = F.lib()
    P = ctypes.POINTER
    two = np.array([0.5, 1.0])
    bank = F.make_bank(F.SearchConfig(), two, np.ones(2))
    cfg = F.SearchCfgC(0.5, 13.7, 3.0)
    roff = np.zeros(3, dtype=np.uintp)

    def call(offs, keep, edges):
        offs = np.asarray(offs, dtype=np.uintp)
        keep = np.asarray(keep, dtype=np.uint8)
        edges = np.asarray(edges, dtype=np.float64)
        return L.fbls_signal_search_bins_ragged(
            None, None, 2, two.ctypes.data_as(P(ctypes.c_double)),
            np.ones(4).ctypes.data_as(P(ctypes.c_double)), 2, ctypes.byref(bank),
            ctypes.byref(cfg), None, 0, roff.ctypes.data_as(P(ctypes.c_size_t)), None,
            keep.ctypes.data_as(P(ctypes.c_uint8)), offs.ctypes.data_as(P(ctypes.c_size_t)),
            edges.ctypes.data_as(P(ctypes.c_double)))

    def message():
        return L.fbls_last_error().decode()

    assert call([1, 2, 3], [1, 1, 1], [0, 1, 2, 3, 4]) != 0
    assert "start at 0" in message()
    assert call([0, 2, 2], [1, 1], [0, 1, 2, 3]) != 0
    assert "has no bins" in message()
    assert call([0, 2, 3], [1, 1, 1], [0, 2, 1, 0, 1]) != 0
    assert "must ascend" in message()
    # valid bins: the null context is the first thing refused
    assert call([0, 2, 3], [1, 1, 1], [0, 1, 2, 0, 1]) != 0
    assert "null argument" in message()


def test_search_config_checked_before_device_work():
    """A period range or grid density that would make the grid count
    meaningless (NaN, negative, inverted range) is refused with FBLS_E_ARG
    before any device work, for a batch of 64 stars as for one; the grid
    count itself returns 0 for such arguments instead of casting NaN."""
    import ctypes

    L = F.lib()
    P = ctypes.POINTER
    d = np.array([0.02, 0.04])
    bank = F.make_bank(F.SearchConfig(), d, np.ones(2))
    roff = np.zeros(65, dtype=np.uintp)

    def batch(pmin, pmax, over):
        cfg = F.SearchCfgC(pmin, pmax, over)
        return L.fbls_signal_search_batch(
            None, None, 64, d.ctypes.data_as(P(ctypes.c_double)),
            np.ones(128).ctypes.data_as(P(ctypes.c_double)), 2, ctypes.byref(bank),
            ctypes.byref(cfg), None, 0, roff.ctypes.data_as(P(ctypes.c_size_t)), None, None)

    for over in (float("nan"), -3.0, 0.0, float("inf")):
        assert batch(0.5, 13.7, over) == 4
        assert "period_oversample" in L.fbls_last_error().decode()
    for pmin, pmax in ((13.7, 0.5), (-1.0, 5.0), (float("nan"), 5.0), (0.5, float("inf"))):
        assert batch(pmin, pmax, 3.0) == 4
        assert "period_min < period_max" in L.fbls_last_error().decode()
    assert batch(0.5, 13.7, 3.0) == 4  # a valid configuration: the null context
    assert "null argument" in L.fbls_last_error().decode()
    assert L.fbls_period_grid_count(0.5, 13.7, 27.4, 0.02, float("nan")) == 0
    assert L.fbls_period_grid_count(13.7, 0.5, 27.4, 0.02, 3.0) == 0
    assert L.fbls_period_grid_count(0.5, 13.7, 27.4, 0.02, 1e300) == 1 << 62


def test_missing_library_raises():
    """The product path has no CPU fallback: without the built library every
    entry into it raises, naming the build command."""
    import subprocess
    import sys

    code = ("import fbls_amd as F\n"
            "F.LIB_PATH = '/nonexistent/libfbls_hip.so'\n"
            "try:\n"
            "    F.lib()\n"
            "except ImportError as e:\n"
            "    print('raised:', e)\n"
            "else:\n"
            "    raise SystemExit('no error')\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "raised:" in r.stdout and "build" in r.stdout


def test_header_is_plain_c_and_links(tmp_path):
    """include/fbls.h compiles as strict C99 and C11 (the language cgo and
    bindgen read), and a C program linked against the library calls the
    host entry points: the grid's first and last periods are exactly the
    bounds, and a double prints as its shortest text."""
    import shutil
    import subprocess

    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    src = tmp_path / "probe.c"
    src.write_text(r'''
#include <stdio.h>
#include <stdlib.h>
#include "fbls.h"
int main(void) {
  size_t n = fbls_period_grid_count(0.5, 20.0, 50.0, 0.02, 3.0);
  double* g = (double*)malloc(n * sizeof(double));
  char text[32];
  if (n < 2 || !g) return 2;
  if (fbls_period_grid(0.5, 20.0, 50.0, 0.02, 3.0, g, n) != FBLS_OK) return 3;
  fbls_format_f64(0.1, text, sizeof text);
  printf("%zu %.17g %.17g %s\n", n, g[0], g[n - 1], text);
  free(g);
  return 0;
}
''')
    inc = os.path.join(ROOT, "include")
    libdir = os.path.dirname(F.LIB_PATH)
    for std in ("c99", "c11"):
        r = subprocess.run([gcc, f"-std={std}", "-Wall", "-Wextra", "-Werror", "-pedantic",
                            "-fsyntax-only", "-I", inc, str(src)],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
    exe = tmp_path / "probe"
    r = subprocess.run([gcc, "-std=c99", "-I", inc, str(src), "-o", str(exe),
                        "-L", libdir, "-lfbls_hip", f"-Wl,-rpath,{libdir}"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stderr)
    n, first, last, text = r.stdout.split()
    assert int(n) == len(F.period_grid(0.5, 20.0, 50.0, 0.02, 3.0))
    assert float(first) == 0.5 and float(last) == 20.0 and text == "0.1"


def test_median_selection_matches_sort(tmp_path):
    """The preparation's lower median (stats.rs:4-11) by bucket selection
    equals the element a full sort places at (m - 1) / 2 on 20,000 arrays
    built to break it (tests/native/median_check.cpp)."""
    import subprocess

    here = os.path.dirname(os.path.abspath(__file__))
    exe = str(tmp_path / "median_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off",
                    os.path.join(here, "native", "median_check.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "mismatches 0" in out.stdout

