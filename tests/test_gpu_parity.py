"""GPU parity of libxbach against the CPU oracle (needs an MI355X).

Bar: everything bit-identical — stamp sky/sigma, substamp coordinates and
values, kept and test stamps, merits, direction, the cd and ksc solutions (298
FP64 terms), refit iterations, final substamp indices, the 16-bit pixel mask,
and the FP32 convolved and difference images.  (tests/_compare.py's looser
image tolerance, 1e-5 |ref| + 1e-3 sigma_pix for conv and 1e-3 sigma_pix for
diff, is checked first so a failure reports its size.)
"""
import pathlib

import numpy as np
import pytest

import _compare
import _oracle
import xbach

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = xbach.Context(0)
    yield c
    c.close()


def check_case(ctx, tmpl, sci, params=None):
    params = params or xbach.Params.default()
    conv, diff, _ = ctx.run(params, tmpl, sci)
    tr = ctx.trace().as_dict()
    mask = ctx.mask()
    cr, dr, mr, trr, _ = _oracle.run(params, tmpl, sci)
    bad = _compare.trace_mismatches(tr, trr.as_dict())
    assert not bad, bad[:8]
    assert np.array_equal(mask, mr), int(np.count_nonzero(mask != mr))
    assert np.array_equal(np.isnan(conv), np.isnan(cr))
    assert np.array_equal(np.isnan(diff), np.isnan(dr))
    err = _compare.image_errors(conv, diff, cr, dr, tmpl, sci)
    assert err["conv_err"] <= 1.0 and err["diff_err"] <= 1.0, err
    assert err["masked_mismatch"] == 0, err
    # the plain form follows the reference's FP64 term order: bit-identical
    # (NaN payload bits differ between the GPU and x86; positions must match)
    for got, ref in ((conv, cr), (diff, dr)):
        nan = np.isnan(ref)
        assert np.array_equal(got[~nan].view(np.uint32), ref[~nan].view(np.uint32)), err
    return conv, diff, tr


def test_test01_bit_exact_and_hotpants(ctx, test01_pair):
    tmpl, sci, gconv, gsub = test01_pair
    conv, diff, tr = check_case(ctx, tmpl, sci)
    c = xbach.diff_fits_stats(gconv, conv)
    s = xbach.diff_fits_stats(gsub, diff)
    # tools/run_test.py:13 and :149-151
    assert c["max_abs"] < 2e-4 and c["max_rel"] < 5e-3 and c["wrong_nans"] == 0, c
    assert s["max_abs"] < 5e-4 and s["max_rel"] < 4e-3 and s["wrong_nans"] == 0, s
    assert round(tr["kernSum"], 4) == -0.1658 and tr["convTemplate"] == 0


def test_pair512(ctx):
    t, s = xbach.synth_pair(xbach.synth_cfg("pair512"))
    check_case(ctx, t, s)


def test_pair2048(ctx):
    t, s = xbach.synth_pair(xbach.synth_cfg("pair2048"))
    conv, diff, tr = check_case(ctx, t, s)
    # size-independent property: sub is the elementwise relation on good pixels
    ks = tr["kernSum"]
    good = diff != np.float32(1e-30)
    expect = (conv.astype(np.float64) * ks - s) * (-1.0 / ks)
    assert np.max(np.abs(diff[good] - expect[good]) / (1e-6 * np.abs(expect[good]) + 1e-3)) < 1


def test_stamp_grid_20x20_ragged(ctx):
    cfg = xbach.synth_cfg("pair2048", w=1100, h=900, nstars=3000, seed=77)
    t, s = xbach.synth_pair(cfg)
    check_case(ctx, t, s, xbach.Params.default(stampsx=20, stampsy=20))


def test_wide_stamps_unstaged_search(ctx):
    # 6x6 stamps on 1600x1600: 245-px stamps do not fit in LDS, so the
    # substamp search reads the image and mask from global memory
    cfg = xbach.synth_cfg("pair2048", w=1600, h=1600, nstars=2500, seed=1601)
    t, s = xbach.synth_pair(cfg)
    check_case(ctx, t, s, xbach.Params.default(stampsx=6, stampsy=6))


def test_science_convolved_direction(ctx):
    # science sharper than the template: the science image is convolved
    cfg = xbach.synth_cfg("pair512", fwhmT=3.2, fwhmS0=1.9, seed=5)
    t, s = xbach.synth_pair(cfg)
    _, _, tr = check_case(ctx, t, s)
    assert tr["convTemplate"] == 0


def test_saturation_cosmics_and_nans(ctx):
    cfg = xbach.synth_cfg("pair512", nstars=500, seed=9, fluxMax=3e5)
    t, s = xbach.synth_pair(cfg)
    rng = np.random.default_rng(4)
    for _ in range(40):  # cosmic-ray hits
        y, x = rng.integers(30, 480, 2)
        s[y:y + rng.integers(1, 3), x:x + rng.integers(1, 3)] += rng.uniform(5e3, 5e4)
    t[100, 200] = np.nan
    s[300, 310] = np.nan
    s[200:203, 40:44] = 0.0  # bad (zero) pixels
    check_case(ctx, t, s)


def test_nans_in_both_images(ctx):
    # NaN pixels scattered through both images (some at the same pixel): each
    # image's statistics and substamp search see the other image's NaN marks
    # in the reference's order (calcStats T, calcStats S, then the searches)
    t, s = xbach.synth_pair(xbach.synth_cfg("pair512", seed=41))
    rng = np.random.default_rng(41)
    both = rng.integers(20, 492, (150, 2))
    for arr, seed in ((t, 1), (s, 2)):
        pts = np.random.default_rng(seed).integers(20, 492, (1500, 2))
        arr[pts[:, 0], pts[:, 1]] = np.nan
        arr[both[:, 0], both[:, 1]] = np.nan
    check_case(ctx, t, s)


def test_stage_api_equals_pipeline(ctx):
    t, s = xbach.synth_pair(xbach.synth_cfg("pair512", seed=21))
    p = xbach.Params.default()
    conv, diff, _ = ctx.run(p, t, s)
    ctx.init(p, t, s)
    ctx.sss()
    ctx.cmv()
    ct = ctx.cd()
    ctx.ksc()
    ks = ctx.conv()
    ctx.sub()
    assert ct == bool(ctx.trace().convTemplate)
    assert ks == ctx.trace().kernSum
    assert np.array_equal(ctx.conv_image(), conv)
    assert np.array_equal(ctx.diff_image(), diff)


def test_device_resident_run_and_determinism(ctx):
    t, s = xbach.synth_pair(xbach.synth_cfg("pair512", seed=33))
    p = xbach.Params.default()
    conv, diff, _ = ctx.run(p, t, s)
    d_t = ctx.buf_alloc(t.nbytes)
    d_s = ctx.buf_alloc(s.nbytes)
    try:
        ctx.buf_upload(d_t, t)
        ctx.buf_upload(d_s, s)
        for _ in range(2):
            ctx.run_device(p, 512, 512, d_t, d_s)
            dc, dd = ctx.device_outputs()
            c2 = np.empty_like(t)
            d2 = np.empty_like(t)
            ctx.buf_download(c2, dc)
            ctx.buf_download(d2, dd)
            assert np.array_equal(c2, conv) and np.array_equal(d2, diff)
    finally:
        ctx.buf_free(d_t)
        ctx.buf_free(d_s)


def test_no_substamps_is_an_error(ctx):
    z = np.zeros((256, 256), np.float32)
    with pytest.raises(xbach.XbachError, match="NO_SUBSTAMPS"):
        ctx.run(xbach.Params.default(), z, z)
    with pytest.raises(RuntimeError, match="NO_SUBSTAMPS"):
        _oracle.run(xbach.Params.default(), z, z)


def test_unsupported_method_constants_rejected(ctx):
    t = np.ones((128, 128), np.float32)
    with pytest.raises(xbach.XbachError, match="ERR_ARG"):
        ctx.run(xbach.Params.default(fKernelWidth=19, hKernelWidth=9), t, t)


def test_sparse_field_runs_the_refill_retry(ctx):
    # 20 stars on 512x512: fewer than 10 % of the stamps hold a substamp, so
    # sss clears the SKIP marks and searches again (src/bach.cpp:142-161)
    t, s = xbach.synth_pair(xbach.synth_cfg("pair512", nstars=20, seed=1))
    _, _, tr = check_case(ctx, t, s)
    assert tr["retried"] == 1


def test_template_without_substamps_has_merit_666(ctx):
    # a star-free template keeps no stamp: its direction's merit is the
    # reference's 666 (src/cdkscUtil.cpp:101-103, bach.cpp:282-300), the science
    # direction wins, and the clip-refit loop runs on a template with no stamps
    tn, _ = xbach.synth_pair(xbach.synth_cfg("pair512", nstars=0, seed=8))
    _, s = xbach.synth_pair(xbach.synth_cfg("pair512", seed=7))
    _, _, tr = check_case(ctx, tn, s)
    assert tr["stampCount"][0] == 0 and tr["merit"][0] == 666.0
    assert tr["convTemplate"] == 1 and tr["retried"] == 1


@pytest.mark.parametrize("tmpl_name,sci_name", [
    ("test01_template", "test0_bg"),
    ("test01_template", "test2"),
    ("test0_bg", "test2"),
    ("test0_bg", "test01_science"),
])
def test_reference_held_images(ctx, tmpl_name, sci_name):
    # the reference's other 330x330 images (res/test0_bg.fits, res/test2.fits;
    # no goldens): negative merits, 1 to 7 clip-refit passes, both directions
    g = pathlib.Path(__file__).resolve().parent / "golden"
    _, t = xbach.read_fits(g / f"{tmpl_name}.fits.gz")
    _, s = xbach.read_fits(g / f"{sci_name}.fits.gz")
    check_case(ctx, t.astype(np.float32), s.astype(np.float32))


def test_image_too_small_for_a_stamp(ctx):
    # 40x40 holds no 51-pixel stamp (the reference's fallback grid is 0x0)
    z = np.full((40, 40), 100.0, np.float32)
    with pytest.raises(xbach.XbachError, match="CAPACITY"):
        ctx.run(xbach.Params.default(), z, z)
    with pytest.raises(RuntimeError, match="CAPACITY"):
        _oracle.run(xbach.Params.default(), z, z)
