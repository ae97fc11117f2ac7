"""Pins the on-disk formats of the command-line host (no GPU needed).

* Number text: every number with a fraction or exponent in the reference's
  own committed reports (tests/golden/float_text_72668830.json, written by
  serde_json) prints identically through fbls_format_f64, and random
  doubles (powers of two, subnormals, integers, wide exponents) print as
  the shortest text that reads back (Python's repr gives the same digits).
* Layout: the reference's committed search report re-prints byte for byte
  (sorted keys, two-space indent, "[]" / "{}", null).
* NPZ input (npz.rs:17-113, main.rs:50-132): archives written by numpy
  (stored and deflated, zip64 records) in every supported dtype and byte
  order read back to numpy's values; groups follow (file, sector) first
  appearance; the reference's error rules (missing columns, mismatched
  shapes, flux_err in some files only, fortran order, unsupported dtype).
* The command line rejects bad arguments before touching a GPU.
"""
import json
import math
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import fbls_amd as F  # noqa: E402

GOLDEN = os.path.join(ROOT, "tests", "golden")


def test_number_text_matches_reference_reports():
    texts = json.load(open(os.path.join(GOLDEN, "float_text_72668830.json")))
    assert len(texts) > 5000
    bad = [(t, F.format_f64(float(t))) for t in texts if F.format_f64(float(t)) != t]
    assert not bad, bad[:10]


def _digits(s):
    m = s.lstrip("-").split("e")[0].replace(".", "").lstrip("0").rstrip("0")
    return len(m) or 1


def test_number_text_is_shortest_round_trip():
    rng = np.random.default_rng(3)
    vals = [float(v) for v in rng.standard_normal(3000) * 10.0 ** rng.integers(-30, 30, 3000)]
    vals += [2.0 ** k for k in range(-1074, 1024, 7)] + [5e-324, 2.2250738585072014e-308,
                                                          1.7976931348623157e308]
    vals += [float(i) for i in range(0, 100001, 997)] + [1e15, 1e16, 1e17, 123456789012345678.0]
    vals += [0.1, 0.2, 0.3, 1 / 3, 2 / 3, 1e-5, 1e-4, 1e21, 1e22, 9007199254740993.0]
    for x in vals + [-v for v in vals[:500]]:
        s = F.format_f64(x)
        assert float(s) == x, (x, s)
        assert _digits(s) == _digits(repr(x)), (x, s, repr(x))
    for x, want in ((0.0, "0.0"), (-0.0, "-0.0"), (1.0, "1.0"), (1024.0, "1024.0"),
                    (1e-12, "1e-12"), (0.01, "0.01"), (1e-5, "0.00001"), (1e-6, "1e-6"),
                    (1e16, "1e16"), (1234567890123456.0, "1234567890123456.0"),
                    (1.5e-7, "1.5e-7"), (2.5e20, "2.5e20"), (math.inf, "null"),
                    (math.nan, "null")):
        assert F.format_f64(x) == want, (x, F.format_f64(x), want)


def test_report_layout_reprints_byte_for_byte():
    text = open(os.path.join(GOLDEN, "report_72668830.json")).read()
    assert F.json_reformat(text) == text.rstrip("\n")
    # and an object built in another key order prints sorted
    assert F.json_reformat('{"b": [], "a": {"y": null, "x": [1, 2.5]}, "c": {}}') == (
        '{\n  "a": {\n    "x": [\n      1,\n      2.5\n    ],\n    "y": null\n  },\n'
        '  "b": [],\n  "c": {}\n}')



def test_report_reader_rejects_malformed_json():
    """The report reader follows RFC 8259 and serde_json's limits: nesting
    deeper than 128 and malformed numbers or surrogate escapes are errors;
    integers beyond 64 bits, and -0, read as doubles."""
    deep = "[" * 128 + "]" * 128
    assert F.json_reformat(deep).count("[") == 128
    with pytest.raises(RuntimeError, match="recursion limit exceeded"):
        F.json_reformat("[" * 129 + "]" * 129)
    with pytest.raises(RuntimeError, match="recursion limit exceeded"):
        F.json_reformat('{"a": ' * 200_000 + "1" + "}" * 200_000)
    for bad in ("01", "1.", "-", "1e", "1e+", ".5", "1-2", "[1.5.5]", "--1"):
        with pytest.raises(RuntimeError):
            F.json_reformat(bad)
    assert F.json_reformat("[-0, 0.5e-3, 1E+2, 18446744073709551615]") == (
        "[\n  -0.0,\n  0.0005,\n  100.0,\n  18446744073709551615\n]")
    assert F.json_reformat("[18446744073709551616, -9223372036854775809]") == (
        "[\n  1.8446744073709552e19,\n  -9.223372036854776e18\n]")
    assert json.loads(F.json_reformat('"\\ud83d\\ude00"')) == "\U0001F600"
    for bad in ('"\\ud83d"', '"\\ud83d\\u0041"', '"\\ude00"'):
        with pytest.raises(RuntimeError, match="surrogate"):
            F.json_reformat(bad)


def _npz(path, compressed=False, **arrays):
    (np.savez_compressed if compressed else np.savez)(path, **arrays)
    return str(path)


@pytest.mark.parametrize("compressed", [False, True])
def test_npz_dtypes_and_groups(tmp_path, compressed):
    rng = np.random.default_rng(1)
    n = 1000
    t = np.sort(rng.uniform(0, 27, n)) + 2458000.0
    a = _npz(tmp_path / "a.npz", compressed, time_bjd=t, flux=rng.normal(1, 1e-3, n),
             flux_err=np.full(n, 1e-3, dtype=np.float32), sector=np.int64(14))
    sec = np.where(np.arange(n) < 400, 15, 16).astype(np.int16)
    b = _npz(tmp_path / "b.npz", compressed, time_bjd=(t + 30).astype(">f8"),
             flux=rng.integers(-1000, 1000, n).astype(np.int32),
             flux_err=rng.integers(1, 100, n).astype(np.uint8), sector=sec)
    c = _npz(tmp_path / "c.npz", compressed, time_bjd=(t + 60).astype(np.float32),
             flux=rng.integers(-5, 5, n).astype(">i8"), flux_err=(np.arange(n) % 2).astype(bool))
    got = F.read_inputs([a, b, c])
    ld = [np.load(p) for p in (a, b, c)]
    for key, col in (("time", "time_bjd"), ("flux", "flux"), ("err", "flux_err")):
        want = np.concatenate([x[col].astype(np.float64) for x in ld])
        np.testing.assert_array_equal(got[key], want)
    want_g = np.concatenate([np.zeros(n), np.where(np.arange(n) < 400, 1, 2), np.full(n, 3)])
    np.testing.assert_array_equal(got["group"], want_g)


def test_npz_without_errors_and_zip64(tmp_path):
    import zipfile

    t = np.linspace(0, 10, 50)
    p = str(tmp_path / "z.npz")
    with zipfile.ZipFile(p, "w", compression=zipfile.ZIP_DEFLATED) as z:
        for name, arr in (("time_bjd", t), ("flux", np.ones(50))):
            with z.open(name + ".npy", "w", force_zip64=True) as f:
                np.lib.format.write_array(f, arr)
    got = F.read_inputs([p, p])
    assert got["err"] is None
    np.testing.assert_array_equal(got["time"], np.concatenate([t, t]))
    np.testing.assert_array_equal(got["group"], np.repeat([0, 1], 50))


def _fails(paths, text):
    with pytest.raises(RuntimeError) as e:
        F.read_inputs(paths)
    assert text in str(e.value), str(e.value)


def test_npz_error_rules(tmp_path):
    t = np.arange(10.0)
    ok = _npz(tmp_path / "ok.npz", time_bjd=t, flux=t, flux_err=t)
    noerr = _npz(tmp_path / "noerr.npz", time_bjd=t, flux=t)
    _fails([_npz(tmp_path / "m.npz", time_bjd=t)], "required columns are time_bjd and flux")
    _fails([_npz(tmp_path / "s.npz", time_bjd=t, flux=t[:5])],
           "time_bjd and flux must be matching vectors")
    _fails([_npz(tmp_path / "e.npz", time_bjd=t, flux=t, flux_err=t[:3])],
           "flux_err must match time_bjd")
    _fails([ok, noerr], "all input files must either provide flux_err or omit it")
    _fails([noerr, ok], "all input files must either provide flux_err or omit it")
    _fails([_npz(tmp_path / "sec.npz", time_bjd=t, flux=t, sector=np.arange(3))],
           "sector must be a scalar or an observation vector")
    _fails([_npz(tmp_path / "f.npz", time_bjd=np.asfortranarray(np.ones((2, 3))), flux=t)],
           "fortran_order arrays are not supported")
    _fails([_npz(tmp_path / "c.npz", time_bjd=t.astype(np.complex128), flux=t)],
           "unsupported NPY dtype: c16")
    _fails([str(tmp_path / "missing.npz")], "missing.npz")
    _fails([str(tmp_path / "x.fits")], "FITS input is not supported")


def test_cli_rejects_bad_arguments(tmp_path, capfd):
    assert F.cli(["search"]) == 1
    assert "fbls: --input is required" in capfd.readouterr().err
    assert F.cli(["search", "--input", "a.npz", "--bogus"]) == 1
    assert "unknown argument --bogus" in capfd.readouterr().err
    assert F.cli(["search", "--input", "a.npz", "--pmin", "1", "--pmax", "0.5"]) == 1
    assert "require 0 < period_min < period_max" in capfd.readouterr().err
    assert F.cli(["search", "--input", str(tmp_path / "none.npz")]) == 1
    assert "none.npz" in capfd.readouterr().err
    assert F.cli(["batch", "--output-dir", str(tmp_path)]) == 1
    assert "--list FILE is required" in capfd.readouterr().err
    assert F.cli(["frobnicate"]) == 1
    assert "usage: fbls search | batch" in capfd.readouterr().err


def _zip64_tail(data, cd_off):
    """Rewrites a plain zip's end record as zip64 records whose central
    directory offset is cd_off (the entry count and size are kept)."""
    import struct

    e = data.rindex(b"PK\x05\x06")
    (entries,) = struct.unpack_from("<H", data, e + 10)
    (cd_size,) = struct.unpack_from("<I", data, e + 12)
    body = data[:e]
    z64_at = len(body)
    rec = struct.pack("<IQHHIIQQQQ", 0x06064b50, 44, 45, 45, 0, 0,
                      entries, entries, cd_size, cd_off)
    loc = struct.pack("<IIQI", 0x07064b50, 0, z64_at, 1)
    end = struct.pack("<IHHHHIIH", 0x06054b50, 0, 0, 0xffff, 0xffff,
                      0xffffffff, 0xffffffff, 0)
    return body + rec + loc + end


def test_npz_corrupt_sizes_fail_cleanly(tmp_path):
    """Sizes and offsets read from a damaged archive are checked without
    overflow: a shape whose byte count wraps, a deflated member claiming
    more than deflate can expand to, and a zip64 directory offset near
    2^64 each give an error rather than a huge allocation or a read past
    the file."""
    import io
    import zipfile

    t = np.arange(10.0)

    def npy(header, body):
        h = header.encode()
        h += b" " * (63 - (10 + len(h)) % 64) + b"\n"
        return b"\x93NUMPY\x01\x00" + len(h).to_bytes(2, "little") + h + body

    # shape (2^61 + 1,): its byte count wraps to 8 in 64 bits
    bad = npy("{'descr': '<f8', 'fortran_order': False, 'shape': (2305843009213693953,), }",
              b"\0" * 8)
    p = str(tmp_path / "shape.npz")
    with zipfile.ZipFile(p, "w") as z:
        buf = io.BytesIO()
        np.lib.format.write_array(buf, t)
        z.writestr("flux.npy", buf.getvalue())
        z.writestr("time_bjd.npy", bad)
    _fails([p], "NPY shape too large")

    # a deflated member whose central record claims ~4 GiB uncompressed
    p = str(tmp_path / "ratio.npz")
    _npz(tmp_path / "ratio0.npz", compressed=True, time_bjd=t, flux=t)
    with zipfile.ZipFile(str(tmp_path / "ratio0.npz")) as z:
        assert all(i.compress_type == zipfile.ZIP_DEFLATED for i in z.infolist())
    data = bytearray(open(str(tmp_path / "ratio0.npz"), "rb").read())
    c = data.index(b"PK\x01\x02")
    data[c + 24:c + 28] = (0xfffffff0).to_bytes(4, "little")
    open(p, "wb").write(bytes(data))
    _fails([p], "corrupt deflate stream")

    # zip64 end records: the true offset reads back; one near 2^64 is refused
    src = str(tmp_path / "plain.npz")
    with zipfile.ZipFile(src, "w") as z:
        for name, arr in (("time_bjd", t), ("flux", t)):
            buf = io.BytesIO()
            np.lib.format.write_array(buf, arr)
            z.writestr(name + ".npy", buf.getvalue())
    plain = open(src, "rb").read()
    good = _zip64_tail(plain, plain.index(b"PK\x01\x02"))
    p = str(tmp_path / "z64_ok.npz")
    open(p, "wb").write(good)
    np.testing.assert_array_equal(F.read_inputs([p])["time"], t)
    wrapped = _zip64_tail(plain, (1 << 64) - 10)
    p = str(tmp_path / "z64_wrap.npz")
    open(p, "wb").write(wrapped)
    _fails([p], "invalid central directory")
