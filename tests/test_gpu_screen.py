"""The learned substamp screen in the pipeline (XB_FORM_RECORD, XB_FORM_SCREEN).

Bars: recording the screen's samples changes nothing — images and every
decision bit-identical to the plain run; the recorded samples cover every
kept stamp of the convolved image in every recorded pass, with labels in
{-1, 0, 1, 2, 3}.  With the committed model (screen_model.npy)
on held-out contaminated crowded pairs (seeds never used in training), every
pair's difference-image residual RMS stays within 1 % of the plain form's (two
pairs also within 1 % of the FP64 oracle's), the screen replaces substamps,
every substamp it replaced is judged by the chi2 re-check after the loop
converged (restored or confirmed).  The re-check costs passes: about half of
the screened substamps pass the chi2 test under the converged fit, are
restored, and the loop runs again (six held-out pairs: 59 passes plain, 72
screened, 88 screened substamps of which 39 restored); the test bounds that
cost at 1.5x.  A model whose threshold screens nothing gives the plain form's
results bit for bit.
"""
import pathlib
import sys

import numpy as np
import pytest

import _compare
import _oracle
import xbach

sys.path.insert(0, str(pathlib.Path(xbach.__file__).resolve().parent))
import screen  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = xbach.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def held_out(ctx):
    p = xbach.Params.default(stampsx=20, stampsy=20)
    cfgs = screen.pair_cfgs(6, 777000, side=(2048, 2560))
    return p, list(screen.contaminated_pairs(ctx, p, cfgs))


def test_record_changes_nothing(ctx, held_out):
    p, pairs = held_out
    t, s = pairs[0]
    ctx.set_form(xbach.FORM_PLAIN)
    c0, d0, _ = ctx.run(p, t, s)
    tr0 = ctx.trace().as_dict()
    ctx.set_form(xbach.FORM_RECORD)
    try:
        c1, d1, _ = ctx.run(p, t, s)
        tr1 = ctx.trace().as_dict()
        f, lab, pr, scr = ctx.screen_samples()
    finally:
        ctx.set_form(xbach.FORM_PLAIN)
    assert not _compare.trace_mismatches(tr1, tr0)
    for a, b in ((c1, c0), (d1, d0)):
        nan = np.isnan(b)
        assert np.array_equal(a[~nan].view(np.uint32), b[~nan].view(np.uint32))
    conv_im = 0 if tr0["convTemplate"] else 1
    npass = min(tr0["kscIterations"], 16)
    assert len(lab) == npass * tr0["stampCount"][conv_im]
    assert set(np.unique(lab)) <= {-1, 0, 1, 2, 3}
    assert f.shape == (len(lab), 32) and np.all(np.isfinite(f)) and scr == 0


@pytest.mark.skipif(not screen.MODEL_PATH.exists(), reason="no trained screen model")
def test_screen_rechecked_by_chi2(ctx, held_out):
    p, pairs = held_out
    model = np.load(screen.MODEL_PATH)
    ctx.screen_load(model)
    it_plain, it_scr, n_scr, n_res, n_conf = [], [], 0, 0, 0
    for i, (t, s) in enumerate(pairs):
        ctx.set_form(xbach.FORM_PLAIN)
        _, dp, _ = ctx.run(p, t, s)
        it_plain.append(ctx.trace().kscIterations)
        ctx.set_form(xbach.FORM_SCREEN)
        try:
            _, ds, _ = ctx.run(p, t, s)
            tr = ctx.trace()
            it_scr.append(tr.kscIterations)
            n_scr += ctx.screen_samples()[3]
            assert tr.screened == ctx.screen_samples()[3]
            n_res += tr.screenRestored
            n_conf += tr.screenConfirmed
        finally:
            ctx.set_form(xbach.FORM_PLAIN)
        rp, rs = _compare.residual_rms(dp), _compare.residual_rms(ds)
        assert abs(rs - rp) <= 0.01 * rp, (i, rs, rp)
        if i < 2:
            _, d64, _, _, _ = _oracle.run(p, t, s, mode=_oracle.MODE_FP64)
            r64 = _compare.residual_rms(d64)
            assert abs(rs - r64) <= 0.01 * r64, (i, rs, r64)
    print(f"clip-refit passes plain {it_plain} screened {it_scr}; substamps screened {n_scr}, "
          f"restored by the chi2 re-check {n_res}, confirmed {n_conf}")
    assert n_scr > 0
    # every screened substamp was judged by the chi2 re-check
    assert n_res + n_conf == n_scr
    # what the screen costs once every screened substamp is re-judged: about
    # half of them pass the chi2 test under the converged fit and re-open the
    # loop (measured: 72 passes against 59 on these six pairs); bounded here
    assert sum(it_scr) <= 1.5 * sum(it_plain)


@pytest.mark.skipif(not screen.MODEL_PATH.exists(), reason="no trained screen model")
def test_screen_that_screens_nothing_equals_plain(ctx, held_out):
    # threshold above 1: no substamp is screened, so the screened form's
    # decisions and images equal the plain form's bit for bit
    p, pairs = held_out
    t, s = pairs[1]
    model = np.load(screen.MODEL_PATH).copy()
    model[screen.TAU] = 2.0
    ctx.set_form(xbach.FORM_PLAIN)
    c0, d0, _ = ctx.run(p, t, s)
    tr0 = ctx.trace().as_dict()
    ctx.screen_load(model)
    ctx.set_form(xbach.FORM_SCREEN)
    try:
        c1, d1, _ = ctx.run(p, t, s)
        tr = ctx.trace()
        tr1 = tr.as_dict()
    finally:
        ctx.set_form(xbach.FORM_PLAIN)
    assert tr.screened == 0 and tr.screenRestored == 0 and tr.screenConfirmed == 0
    assert not _compare.trace_mismatches(tr1, tr0)
    for a, b in ((c1, c0), (d1, d0)):
        nan = np.isnan(b)
        assert np.array_equal(a[~nan].view(np.uint32), b[~nan].view(np.uint32))
