"""Launch shapes (xb_set_shapes, xb_batch_autotune): the tiling of the
convolution (tiles per workgroup), the input-mask pass (rows per tile), the
L D L^T solve (trailing tiles per load group), whether the substamp search
stages each stamp in LDS, and the stamp statistics' threads per stamp are free.

Bar: every shape gives the default shape's convolved image, difference image
and mask bit for bit, in the plain form and in the fast form (mixed cmv,
L D L^T solves with refinement), on a square pair and on a ragged rectangle
whose last tiles are partial; invalid shapes are refused; the batch autotuner
returns valid shapes per size bucket and its tuned runs equal single runs.
"""
import numpy as np
import pytest

import xbach

pytestmark = pytest.mark.gpu

FAST = xbach.FORM_MIXED | xbach.FORM_LDLT | xbach.FORM_REFINE


def _same(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.fixture(scope="module")
def ctx():
    c = xbach.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module", params=["square", "ragged"])
def pair(request):
    if request.param == "square":
        return xbach.synth_pair(xbach.synth_cfg("pair512"))
    # sides that are not multiples of the 21-pixel tiles nor of the mask tiles
    return xbach.synth_pair(xbach.synth_cfg("pair512", w=923, h=611, nstars=1200, seed=913))


def _run(ctx, form, shapes, t, s):
    p = xbach.Params.default()
    ctx.set_form(form)
    ctx.set_shapes(shapes)
    try:
        conv, diff, _ = ctx.run(p, t, s)
        mask = ctx.mask()
    finally:
        ctx.set_shapes(None)
        ctx.set_form(xbach.FORM_PLAIN)
    return conv, diff, mask


@pytest.mark.parametrize("form", [xbach.FORM_PLAIN, FAST], ids=["plain", "fast"])
def test_every_shape_gives_the_same_bits(ctx, pair, form):
    t, s = pair
    ref = _run(ctx, form, None, t, s)
    variants = [xbach.Shapes.make(conv_tiles=ct) for ct in (1, 2, 8)]
    variants += [xbach.Shapes.make(mask_rows=mr) for mr in (16, 32)]
    variants.append(xbach.Shapes.make(find_staged=0))
    variants.append(xbach.Shapes.make(stats_threads=256))
    variants.append(xbach.Shapes.make(stats_threads=512))
    if form == FAST:
        variants.append(xbach.Shapes.make(ldlt_group=2))
    variants.append(xbach.Shapes.make(conv_tiles=1, mask_rows=16, ldlt_group=2))
    for v in variants:
        conv, diff, mask = _run(ctx, form, v, t, s)
        assert _same(conv, ref[0]), v.as_dict()
        assert _same(diff, ref[1]), v.as_dict()
        assert np.array_equal(mask, ref[2]), v.as_dict()


def test_invalid_shapes_are_refused(ctx):
    for bad in (xbach.Shapes.make(conv_tiles=3), xbach.Shapes.make(mask_rows=48),
                xbach.Shapes.make(ldlt_group=8), xbach.Shapes.make(conv_tiles=0),
                xbach.Shapes.make(find_staged=2), xbach.Shapes.make(stats_threads=768)):
        with pytest.raises(xbach.XbachError):
            ctx.set_shapes(bad)
    assert ctx.shapes().as_dict() == xbach.Shapes.make().as_dict()


def test_autotuned_shapes_are_valid_and_runs_equal_single_runs():
    p = xbach.Params.default()
    pairs = [xbach.synth_pair(xbach.synth_cfg("pair512", seed=600 + i)) for i in range(4)]
    singles = []
    with xbach.Context(0) as c:
        c.set_form(FAST)
        for t, s in pairs:
            conv, diff, _ = c.run(p, t, s)
            singles.append((conv, diff))
    with xbach.Context(0) as c, xbach.Batch(0, workers=4) as b:
        bufs = []
        for t, s in pairs:
            q = [c.buf_alloc(t.nbytes) for _ in range(4)]
            c.buf_upload(q[0], t)
            c.buf_upload(q[1], s)
            bufs.append(q)
        cols = [[q[k] for q in bufs] for k in range(4)]
        ws = [t.shape[1] for t, _ in pairs]
        hs = [t.shape[0] for t, _ in pairs]
        b.set_form(FAST)
        table = b.autotune(p, ws, hs, *cols, reps=1)
        assert table[0] >= 1 and table[1:] == [0, 0, 0]
        sh = b.shapes(0).as_dict()
        print("autotuned bucket 0:", table[0], "in flight,", sh)
        assert sh["conv_tiles"] in (1, 2, 4, 8)
        assert sh["mask_rows"] in (16, 32, 64)
        assert sh["ldlt_group"] in (2, 4)
        assert sh["find_staged"] in (0, 1)
        assert sh["stats_threads"] in (256, 512, 1024)
        # buckets the tuning did not touch keep a batch's defaults (the
        # statistics at 512 threads per stamp)
        for k in (1, 2, 3):
            assert b.shapes(k).as_dict() == xbach.Shapes.make(stats_threads=512).as_dict()
        res = b.run_device(p, ws, hs, *cols)
        assert all(r.status == 0 for r in res)
        for i, (conv, diff) in enumerate(singles):
            oc = np.empty_like(conv)
            od = np.empty_like(diff)
            c.buf_download(oc, bufs[i][2])
            c.buf_download(od, bufs[i][3])
            assert _same(oc, conv), i
            assert _same(od, diff), i
        with pytest.raises(xbach.XbachError):
            b.set_shapes(xbach.Shapes.make(conv_tiles=5))
