"""The path at the real shapes of the precision, batching and screening
workloads (the 2048x2048 pair is covered in test_gpu_parity.py).

* 4096x4096, 20000 stars, 20x20 stamps (400 stamps, 3 substamps each): the plain
  form bit-identical to the oracle; the mixed form's difference image, with W
  in BF16 and in FP16 split precision, within 1 % of the FP64 oracle's residual
  RMS, with its coefficient and pixel errors reported.
* 64 independent pairs, sides U[512, 3072], densities U[1e-3, 5e-3], FWHM
  U[1.5, 4.0] (xbach.batch_cfgs): one batch, every pair bit-identical to its
  own single-pair run.
* 3072x3072 crowded field, 30000 stars, 20 % of the selected substamps
  contaminated (cosmic rays, saturation at 6e4, variable sources): the
  mixed + screened form within 1 % of the FP64 oracle's residual RMS.
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


def _same(a, b):
    nan = np.isnan(b)
    return np.array_equal(np.isnan(a), nan) and np.array_equal(a[~nan].view(np.uint32),
                                                               b[~nan].view(np.uint32))


@pytest.fixture(scope="module")
def ctx():
    c = xbach.Context(0)
    yield c
    c.close()


def test_precision_workload_4096(ctx):
    p = xbach.Params.default(stampsx=20, stampsy=20)
    t, s = xbach.synth_pair(xbach.synth_cfg("pair4096"))
    ctx.set_form(xbach.FORM_PLAIN)
    conv, diff, _ = ctx.run(p, t, s)
    tr = ctx.trace().as_dict()
    cr, dr, mr, trr, _ = _oracle.run(p, t, s)
    assert not _compare.trace_mismatches(tr, trr.as_dict())
    assert np.array_equal(ctx.mask(), mr)
    assert _same(conv, cr) and _same(diff, dr)
    _, d64, _, tr64, _ = _oracle.run(p, t, s, mode=_oracle.MODE_FP64)
    r64 = _compare.residual_rms(d64)
    sp = _compare.sigma_pix(t, s)
    x64 = np.asarray(tr64.as_dict()["solution"])
    for leg, form in (("bf16", xbach.FORM_MIXED),
                      ("fp16", xbach.FORM_MIXED | xbach.FORM_FP16)):
        ctx.set_form(form)
        try:
            _, dm, _ = ctx.run(p, t, s)
            trm = ctx.trace().as_dict()
        finally:
            ctx.set_form(xbach.FORM_PLAIN)
        rm = _compare.residual_rms(dm)
        good = (d64 != np.float32(1e-30)) & (dm != np.float32(1e-30))
        rel = np.abs(dm[good].astype(np.float64) - d64[good]) / sp[good]
        x = np.asarray(trm["solution"])
        dsol = np.max(np.abs(x - x64)) / np.max(np.abs(x64))
        nrm = np.linalg.norm(x - x64) / np.linalg.norm(x64)
        print(f"4096 {leg}: stamps {tr['stampCount']}, passes plain {tr['kscIterations']} "
              f"mixed {trm['kscIterations']}; residual RMS fp64 {r64:.5f} mixed {rm:.5f} "
              f"({(rm - r64) / r64:+.4%}); |d diff|/sigma median {np.median(rel):.2e} max "
              f"{rel.max():.2e}; coefficients {dsol:.2e} of the largest, {nrm:.2e} normwise")
        assert abs(rm - r64) <= 0.01 * r64, leg


def test_batching_workload_64_pairs(ctx):
    p = xbach.Params.default()
    cfgs = xbach.batch_cfgs(64)
    pairs = xbach.synth_pairs(cfgs, threads=16)
    with xbach.Batch(0, workers=8) as b:
        outs, res = b.run(p, pairs)
    sides = [c.w for c in cfgs]
    assert min(sides) < 800 and max(sides) > 2800
    for i, (t, s) in enumerate(pairs):
        conv, diff, _ = ctx.run(p, t, s)
        assert res[i].status == 0, i
        assert _same(outs[i][0], conv) and _same(outs[i][1], diff), i
        assert res[i].kernSum == ctx.trace().kernSum, i
    print(f"64 pairs, sides {min(sides)}..{max(sides)}, clip-refit passes "
          f"{sum(r.kscIterations for r in res)}")


@pytest.mark.skipif(not screen.MODEL_PATH.exists(), reason="no trained screen model")
def test_screening_workload_3072_crowded(ctx):
    p = xbach.Params.default()
    cfg = xbach.synth_cfg("crowd3072")
    (t, s), = list(screen.contaminated_pairs(ctx, p, [cfg]))
    ctx.set_form(xbach.FORM_PLAIN)
    _, dp, _ = ctx.run(p, t, s)
    it_plain = ctx.trace().kscIterations
    ctx.screen_load(np.load(screen.MODEL_PATH))
    ctx.set_form(xbach.FORM_MIXED | xbach.FORM_SCREEN)
    try:
        _, ds, _ = ctx.run(p, t, s)
        it_scr = ctx.trace().kscIterations
        n_scr = ctx.screen_samples()[3]
    finally:
        ctx.set_form(xbach.FORM_PLAIN)
    _, d64, _, _, _ = _oracle.run(p, t, s, mode=_oracle.MODE_FP64)
    r64, rp, rs = (_compare.residual_rms(d) for d in (d64, dp, ds))
    print(f"crowd3072: passes plain {it_plain} mixed+screened {it_scr} (screened {n_scr}); "
          f"residual RMS fp64 {r64:.5f} plain {rp:.5f} mixed+screened {rs:.5f}")
    assert abs(rp - r64) <= 1e-6 * r64
    assert abs(rs - r64) <= 0.01 * r64
