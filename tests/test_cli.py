"""The command-line driver (the reference's main, src/main.cpp:16-159) and the
library's FITS I/O (src/fitsUtil.cpp:8-64).

CPU: the FITS reader against an independent reader (xbach.read_fits, numpy) on
the reference's own test-1 images, BITPIX -64 and BSCALE/BZERO inputs, the
writer's round trip, refusal of integer images with the reference's message,
and the CLI's argument errors (src/argsUtil.cpp:23-81).
GPU: the CLI on test 1 end to end, as tools/run_test.py drives the reference:
outputs named <op><o> and <op>sub.fits, nine -vt timing lines in the order
tools/run_measurements.py parses, images bit-identical to the library's own
run and within the HOTPANTS-golden tolerances of tools/run_test.py:13.  The
difference image carries HOTPANTS's quality cards: KSUM00 equals the golden
header's '-0.1658', CONVOL00 names the convolved image as the golden does, and
NX2NRM00, DMEAN00, DSIGE00, DSIG00 and X2NRM00 equal the same figures computed
here with numpy from the output images and mask (to the 4 written decimals).
"""
import gzip
import pathlib
import re
import shutil
import subprocess

import numpy as np
import pytest

import xbach

ROOT = pathlib.Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
CLI = ROOT / "thesis-emx-bach_amd" / "xbach"


def _fits_bytes(data: np.ndarray, bitpix: int, extra=()) -> bytes:
    cards = ["SIMPLE  =                    T", f"BITPIX  = {bitpix:20d}", "NAXIS   =                    2",
             f"NAXIS1  = {data.shape[1]:20d}", f"NAXIS2  = {data.shape[0]:20d}", *extra, "END"]
    hdr = "".join(c.ljust(80) for c in cards).encode()
    hdr += b" " * ((2880 - len(hdr) % 2880) % 2880)
    dt = {-32: ">f4", -64: ">f8", 16: ">i2"}[bitpix]
    body = np.asarray(data).astype(dt).tobytes()
    body += b"\0" * ((2880 - len(body) % 2880) % 2880)
    return hdr + body


def test_fits_reader_matches_independent_reader_on_test01():
    for name in ("test01_science.fits.gz", "test01_template.fits.gz"):
        _, ref = xbach.read_fits(GOLDEN / name)
        got = xbach.fits_read_native(GOLDEN / name)
        assert got.shape == ref.shape
        assert np.array_equal(got, ref.astype(np.float32))


def test_fits_double_and_scaled_inputs(tmp_path):
    rng = np.random.default_rng(1)
    d = rng.normal(100.0, 30.0, (37, 53))
    p = tmp_path / "d.fits"
    p.write_bytes(_fits_bytes(d, -64))
    assert np.array_equal(xbach.fits_read_native(p), d.astype(np.float32))  # (float) narrowing
    q = tmp_path / "s.fits.gz"
    raw = rng.normal(0.0, 1.0, (20, 30)).astype(np.float32)
    with gzip.open(q, "wb") as f:
        f.write(_fits_bytes(raw, -32, ["BSCALE  =                  2.0", "BZERO   =               1000.0"]))
    expect = (raw.astype(np.float64) * 2.0 + 1000.0).astype(np.float32)
    assert np.array_equal(xbach.fits_read_native(q), expect)


def test_fits_writer_round_trip_and_overwrite(tmp_path):
    rng = np.random.default_rng(2)
    a = rng.normal(0.0, 1e3, (41, 29)).astype(np.float32)
    a[3, 4] = np.nan
    a[5, 6] = np.float32(1e-30)
    p = tmp_path / "o.fits"
    p.write_bytes(b"junk that must be replaced")
    xbach.fits_write_native(p, a)
    assert p.stat().st_size % 2880 == 0
    _, back = xbach.read_fits(p)  # independent reader
    assert np.array_equal(back.astype(np.float32), a, equal_nan=True)
    assert np.array_equal(xbach.fits_read_native(p), a, equal_nan=True)


def test_fits_integer_image_is_refused(tmp_path):
    p = tmp_path / "i.fits"
    p.write_bytes(_fits_bytes(np.zeros((4, 4)), 16))
    with pytest.raises(xbach.XbachError, match="fits image of type16 is not supported."):
        xbach.fits_read_native(p)
    with pytest.raises(xbach.XbachError, match="Unable to open file"):
        xbach.fits_read_native(tmp_path / "missing.fits")


def _cli(*args, **kw):
    return subprocess.run([str(CLI), *args], capture_output=True, text=True, timeout=300, **kw)


@pytest.mark.skipif(not CLI.exists(), reason="xbach CLI not built (make -C thesis-emx-bach_amd)")
def test_cli_argument_errors():
    r = _cli()
    assert r.returncode == 1 and "Template file Input is required!" in r.stdout
    r = _cli("-t", "a.fits")
    assert r.returncode == 1 and "Science file input is required!" in r.stdout
    r = _cli("-t", "a.fits", "-s", "b.fits", "-sss", "gpu")
    assert r.returncode == 1 and "sss mode should be mp or cl" in r.stdout


@pytest.mark.gpu
@pytest.mark.skipif(not CLI.exists(), reason="xbach CLI not built (make -C thesis-emx-bach_amd)")
def test_cli_test01_end_to_end(tmp_path, test01_pair):
    tmpl, sci, gconv, gsub = test01_pair
    inp = tmp_path / "res"
    inp.mkdir()
    shutil.copy(GOLDEN / "test01_template.fits.gz", inp / "test1.fits.gz")
    shutil.copy(GOLDEN / "test01_science.fits.gz", inp / "test0.fits.gz")
    out = tmp_path / "out"
    out.mkdir()
    # tools/run_test.py: -t test1 -s test0 -o <name> -op <dir>/ -ip <dir>; -vt for the timings
    r = _cli("-t", "test1.fits.gz", "-s", "test0.fits.gz", "-o", "test01_diff.fits",
             "-ip", str(inp), "-op", str(out) + "/", "-vt")
    assert r.returncode == 0, r.stdout + r.stderr
    stages = re.findall(r"^(\w+) took (\d+) ms$", r.stdout, re.M)
    assert [s for s, _ in stages] == ["Ini", "SSS", "CMV", "CD", "KSC", "Conv", "Sub", "Fin", "BACH"]
    assert "BACH finished." in r.stdout
    conv = xbach.fits_read_native(out / "test01_diff.fits")
    diff = xbach.fits_read_native(out / "sub.fits")
    ctx = xbach.Context(0)
    try:
        c2, d2, _ = ctx.run(xbach.Params.default(), tmpl, sci)
    finally:
        ctx.close()
    for a, b in ((conv, c2), (diff, d2)):
        assert np.array_equal(np.isnan(a), np.isnan(b))
        assert np.array_equal(a[~np.isnan(a)].view(np.uint32), b[~np.isnan(b)].view(np.uint32))
    c = xbach.diff_fits_stats(gconv, conv)
    s = xbach.diff_fits_stats(gsub, diff)
    assert c["max_abs"] < 2e-4 and c["max_rel"] < 5e-3 and c["wrong_nans"] == 0, c
    assert s["max_abs"] < 5e-4 and s["max_rel"] < 4e-3 and s["wrong_nans"] == 0, s
    # HOTPANTS quality cards
    hdr, _ = xbach.read_fits(out / "sub.fits")
    ghdr, _ = xbach.read_fits(GOLDEN / "test01_sub_hotpants.fits.gz")
    assert hdr["KSUM00"] == ghdr["KSUM00"] == "-0.1658"
    assert hdr["CONVOL00"] == ghdr["CONVOL00"] == "IMAGE"
    assert hdr["PHOTNORM"] == ghdr["PHOTNORM"] == "t"
    ctx = xbach.Context(0)
    try:
        ctx.run(xbach.Params.default(), tmpl, sci)
        mask = ctx.mask()
        diag = ctx.diagnostics()
    finally:
        ctx.close()
    h, w = diff.shape
    yy, xx = np.mgrid[0:h, 0:w]
    good = (xx >= 10) & (xx < w - 10) & (yy >= 10) & (yy < h - 10) & ((mask & (1 << 12)) == 0)
    d = diff[good].astype(np.float64)
    noise = np.sqrt(np.abs(conv[good].astype(np.float64)) + np.abs(tmpl[good].astype(np.float64)))
    assert int(hdr["NX2NRM00"]) == diag["nx2nrm"] == int(good.sum())
    for key, val in (("DMEAN00", d.mean()), ("DSIGE00", d.std(ddof=1)), ("DSIG00", noise.mean()),
                     ("X2NRM00", np.mean((d / noise) ** 2))):
        assert abs(float(hdr[key]) - val) <= 1e-4 * max(1.0, abs(val)) + 5e-5, (key, hdr[key], val)
    print("quality cards:", {k: hdr[k] for k in ("KSUM00", "SSSIG00", "FSIG00", "X2NRM00",
                                                 "NX2NRM00", "DMEAN00", "DSIGE00", "DSIG00")},
          "HOTPANTS:", {k: ghdr[k] for k in ("SSSIG00", "FSIG00", "X2NRM00", "NX2NRM00")})


@pytest.mark.gpu
@pytest.mark.skipif(not CLI.exists(), reason="xbach CLI not built (make -C thesis-emx-bach_amd)")
def test_cli_size_mismatch(tmp_path):
    # src/bach.cpp:20-24: the reference's message, exit status 1, no output
    (tmp_path / "t.fits").write_bytes(_fits_bytes(np.full((64, 64), 100.0), -32))
    (tmp_path / "s.fits").write_bytes(_fits_bytes(np.full((80, 64), 100.0), -32))
    r = _cli("-t", "t.fits", "-s", "s.fits", "-ip", str(tmp_path), "-op", str(tmp_path) + "/")
    assert r.returncode == 1, r.stdout + r.stderr
    assert "Template image and science image must be the same size!" in r.stdout
    assert not (tmp_path / "diff.fits").exists() and not (tmp_path / "sub.fits").exists()
