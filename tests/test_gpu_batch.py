"""Batches of independent pairs (xb_batch_*): many pairs of uneven sizes in
flight at once, each bit-identical to the single-pair pipeline.

Bar: every pair's convolved and difference images, kernel sum, direction and
clip-refit count equal a lone xb_pipeline_run of the same pair bit for bit
(the single-pair path is itself bit-identical to the oracle, and one pair per
batch is checked against the oracle here too); a pair that fails reports its
own status without disturbing the others; host- and device-resident batches
agree; results do not depend on the number of pairs in flight.
"""
import numpy as np
import pytest

import _compare
import _oracle
import xbach

pytestmark = pytest.mark.gpu

NO_SUBSTAMPS = 3


def _same(a, b):
    nan = np.isnan(b)
    return np.array_equal(np.isnan(a), nan) and np.array_equal(a[~nan].view(np.uint32),
                                                               b[~nan].view(np.uint32))


@pytest.fixture(scope="module")
def pairs():
    # uneven square sides (the batching law, sides cut to 512..1280 to keep the
    # test short) plus a ragged rectangle and a pair with no usable stamps
    cfgs = xbach.batch_cfgs(9, seed=77, side=(512, 1280))
    ps = [xbach.synth_pair(c) for c in cfgs]
    ps.append(xbach.synth_pair(xbach.synth_cfg("pair512", w=900, h=700, nstars=1500, seed=31)))
    z = np.zeros((300, 300), np.float32)
    ps.append((z, z))
    return ps


@pytest.fixture(scope="module")
def singles(pairs):
    p = xbach.Params.default()
    out = []
    with xbach.Context(0) as c:
        for t, s in pairs[:-1]:
            conv, diff, _ = c.run(p, t, s)
            tr = c.trace()
            out.append((conv, diff, tr.kernSum, tr.convTemplate, tr.kscIterations))
    return out


def test_batch_host_equals_single_runs(pairs, singles):
    p = xbach.Params.default()
    with xbach.Batch(0, workers=4) as b:
        outs, res = b.run(p, pairs, check=False)
        msg = b._L.xb_batch_last_error(b._h)
    for i, (conv, diff, ks, ct, it) in enumerate(singles):
        r = res[i]
        assert r.status == 0, (i, r.status, msg)
        assert _same(outs[i][0], conv) and _same(outs[i][1], diff), i
        assert (r.kernSum, r.convTemplate, r.kscIterations) == (ks, ct, it), i
    assert res[-1].status == NO_SUBSTAMPS
    assert all(0 <= r.worker < 4 for r in res)
    # worker contexts never capture graphs, so none can fail
    assert all(r.captureFailures == 0 for r in res)


def test_batch_pair_against_oracle(pairs):
    # the smallest pair of the batch, straight against the CPU restatement
    p = xbach.Params.default()
    k = min(range(len(pairs) - 1), key=lambda i: pairs[i][0].size)
    with xbach.Batch(0, workers=2) as b:
        outs, res = b.run(p, [pairs[k], pairs[0]])
    cr, dr, _, trr, _ = _oracle.run(p, *pairs[k])
    assert _same(outs[0][0], cr) and _same(outs[0][1], dr)
    assert res[0].kernSum == trr.kernSum


@pytest.mark.parametrize("inflight", [1, 3])
def test_batch_device_resident(pairs, singles, inflight):
    p = xbach.Params.default()
    use = pairs[:-1]
    with xbach.Context(0) as c, xbach.Batch(0, workers=3) as b:
        bufs = []
        try:
            for t, s in use:
                dt, ds, dc, dd = (c.buf_alloc(t.nbytes) for _ in range(4))
                c.buf_upload(dt, t)
                c.buf_upload(ds, s)
                bufs.append((dt, ds, dc, dd))
            b.set_inflight(inflight)
            res = b.run_device(p, [t.shape[1] for t, _ in use], [t.shape[0] for t, _ in use],
                               [q[0] for q in bufs], [q[1] for q in bufs],
                               [q[2] for q in bufs], [q[3] for q in bufs])
            assert {r.worker for r in res} <= set(range(inflight))
            for i, (t, _) in enumerate(use):
                conv = np.empty_like(t)
                diff = np.empty_like(t)
                c.buf_download(conv, bufs[i][2])
                c.buf_download(diff, bufs[i][3])
                assert _same(conv, singles[i][0]) and _same(diff, singles[i][1]), i
                assert res[i].kernSum == singles[i][2]
        finally:
            for q in bufs:
                for d in q:
                    c.buf_free(d)


def test_batch_argument_errors():
    p = xbach.Params.default()
    with pytest.raises(xbach.XbachError, match="ERR_CAPACITY"):
        xbach.Batch(0, workers=33)  # above XB_BATCH_MAX_WORKERS
    with xbach.Batch(0, workers=2) as b:
        with pytest.raises(xbach.XbachError, match="ERR_CAPACITY"):
            b.set_inflight(3)  # more pairs in flight than workers
        t = np.ones((64, 64), np.float32)
        with pytest.raises(xbach.XbachError, match="ERR_ARG"):
            b.run(xbach.Params.default(fKernelWidth=19), [(t, t)])
        outs, res = b.run(p, [])
        assert outs == [] and res == []
    # device pairs are read and written in place: an output that overlaps an
    # input is refused for that pair (XB_ERR_ARG), the other pair still runs
    t, s = xbach.synth_pair(xbach.synth_cfg("pair512"))
    with xbach.Context(0) as c, xbach.Batch(0, workers=2) as b:
        q = [c.buf_alloc(t.nbytes) for _ in range(6)]
        try:
            for k in (0, 3):
                c.buf_upload(q[k], t)
                c.buf_upload(q[k + 1], s)
            w, h = [t.shape[1]] * 2, [t.shape[0]] * 2
            res = b.run_device(p, w, h, [q[0], q[3]], [q[1], q[4]], [q[2], q[5]],
                               [q[1], None], check=False)  # pair 0's difference over its science
            assert res[0].status == 1, res[0].status
            assert res[1].status == 0, res[1].status
        finally:
            for d in q:
                c.buf_free(d)


def test_batch_autotune(pairs, singles):
    # the autotuner picks a count per size bucket; runs afterwards go bucket by
    # bucket and stay bit-identical to lone runs
    p = xbach.Params.default()
    use = pairs[:-1]
    w = [t.shape[1] for t, _ in use]
    h = [t.shape[0] for t, _ in use]
    with xbach.Context(0) as c, xbach.Batch(0, workers=4) as b:
        bufs = []
        try:
            for t, s in use:
                dt, ds, dc, dd = (c.buf_alloc(t.nbytes) for _ in range(4))
                c.buf_upload(dt, t)
                c.buf_upload(ds, s)
                bufs.append((dt, ds, dc, dd))
            args = ([q[0] for q in bufs], [q[1] for q in bufs], [q[2] for q in bufs],
                    [q[3] for q in bufs])
            table = b.autotune(p, w, h, *args, reps=1)
            present = {0 if max(a, c_) <= 1024 else 1 for a, c_ in zip(w, h)}
            for k in range(4):
                assert (table[k] > 0) == (k in present), table
                assert table[k] <= 4
            res = b.run_device(p, w, h, *args)
            for i, (t, _) in enumerate(use):
                conv = np.empty_like(t)
                diff = np.empty_like(t)
                c.buf_download(conv, bufs[i][2])
                c.buf_download(diff, bufs[i][3])
                assert res[i].status == 0
                assert _same(conv, singles[i][0]) and _same(diff, singles[i][1]), i
                assert res[i].kernSum == singles[i][2]
        finally:
            for q in bufs:
                for d in q:
                    c.buf_free(d)


@pytest.mark.parametrize("inflight", [3, 16])
def test_lockstep_batch_equals_single_runs(pairs, singles, inflight):
    """XB_BATCH_LOCKSTEP: groups of pairs run stage by stage on one host thread,
    every global solve of a group as one batched solve; each pair bit-identical
    to its lone run, the failing pair reports its own status."""
    p = xbach.Params.default()
    with xbach.Batch(0, workers=16, flags=xbach.BATCH_LOCKSTEP) as b:
        b.set_inflight(inflight)
        outs, res = b.run(p, pairs, check=False)
        msg = b._L.xb_batch_last_error(b._h)
    for i, (conv, diff, ks, ct, it) in enumerate(singles):
        r = res[i]
        assert r.status == 0, (i, r.status, msg)
        assert _same(outs[i][0], conv) and _same(outs[i][1], diff), i
        assert (r.kernSum, r.convTemplate, r.kscIterations) == (ks, ct, it), i
    assert res[-1].status == NO_SUBSTAMPS
    assert all(0 <= r.worker < inflight for r in res)


def _nan_pairs():
    # a few NaN pixels in each image and some at the same pixel of both, on a
    # square and a ragged pair (kernels stay finite, but the NaN pixels change
    # the stamp statistics and the substamps chosen): batch workers run both
    # images' statistics passes as one dispatch and write the NaN marks
    # afterwards, which must leave every pass reading the mask as the lone
    # run's T-then-S order does
    out = []
    for seed, w, h, n in ((41, 512, 512, 16), (43, 700, 530, 12)):
        t, s = xbach.synth_pair(xbach.synth_cfg("pair512", w=w, h=h, seed=seed))
        rng = np.random.default_rng(100 * n)
        for arr in (t, s):
            pts = rng.integers(20, min(w, h) - 20, (n, 2))
            arr[pts[:, 0], pts[:, 1]] = np.nan
        pts = rng.integers(20, min(w, h) - 20, (n // 2, 2))
        t[pts[:, 0], pts[:, 1]] = np.nan
        s[pts[:, 0], pts[:, 1]] = np.nan
        out.append((t, s))
    return out


@pytest.mark.parametrize("flags", [0, xbach.BATCH_LOCKSTEP], ids=["workers", "lockstep"])
def test_batch_nan_pixels_equal_single_runs(flags):
    p = xbach.Params.default()
    ps = _nan_pairs()
    ref = []
    with xbach.Context(0) as c:
        for t, s in ps:
            conv, diff, _ = c.run(p, t, s)
            tr = c.trace()
            ref.append((conv, diff, tr.kernSum, tr.convTemplate, tr.kscIterations))
    with xbach.Batch(0, workers=2, flags=flags) as b:
        outs, res = b.run(p, ps, check=False)
        msg = b._L.xb_batch_last_error(b._h)
    for i, (conv, diff, ks, ct, it) in enumerate(ref):
        r = res[i]
        assert r.status == 0, (i, r.status, msg)
        assert _same(outs[i][0], conv) and _same(outs[i][1], diff), i
        assert (r.kernSum, r.convTemplate, r.kscIterations) == (ks, ct, it), i


@pytest.mark.parametrize("form", [xbach.FORM_MIXED | xbach.FORM_REFINE,
                                  xbach.FORM_MIXED | xbach.FORM_LDLT,
                                  xbach.FORM_MIXED | xbach.FORM_LDLT | xbach.FORM_REFINE])
def test_lockstep_refined_equals_single_refined(pairs, form):
    # the batched solves (with each system's iterative refinement, or one
    # L D L^T launch for all systems) give the single-context form's results
    # bit for bit
    p = xbach.Params.default()
    use = pairs[:4]
    ref = []
    with xbach.Context(0) as c:
        c.set_form(form)
        for t, s in use:
            conv, diff, _ = c.run(p, t, s)
            ref.append((conv, diff, c.trace().kernSum))
    with xbach.Batch(0, workers=4, flags=xbach.BATCH_LOCKSTEP) as b:
        b.set_form(form)
        outs, res = b.run(p, use)
    for i, (conv, diff, ks) in enumerate(ref):
        assert res[i].status == 0
        assert _same(outs[i][0], conv) and _same(outs[i][1], diff), i
        assert res[i].kernSum == ks, i


@pytest.mark.parametrize("form", [xbach.FORM_PLAIN, xbach.FORM_MIXED | xbach.FORM_REFINE,
                                  xbach.FORM_MIXED | xbach.FORM_LDLT | xbach.FORM_REFINE])
def test_lockstep_graph_replays_equal_single_runs(pairs, form):
    # each batched solve is captured into a hipGraph at its second occurrence
    # and replayed after that: four runs of the same batch (the later ones
    # replay every solve), with a lone pipeline run on another context between
    # them, each bit-identical to the single runs, and no capture failed
    p = xbach.Params.default()
    use = pairs[:4]
    ref = []
    with xbach.Context(0) as c:
        c.set_form(form)
        for t, s in use:
            conv, diff, _ = c.run(p, t, s)
            ref.append((conv, diff, c.trace().kernSum))
        with xbach.Batch(0, workers=4, flags=xbach.BATCH_LOCKSTEP) as b:
            b.set_form(form)
            for rep in range(4):
                outs, res = b.run(p, use)
                for i, (conv, diff, ks) in enumerate(ref):
                    assert res[i].status == 0, (rep, i)
                    assert res[i].captureFailures == 0, (rep, i)
                    assert _same(outs[i][0], conv) and _same(outs[i][1], diff), (rep, i)
                    assert res[i].kernSum == ks, (rep, i)
                c.run(p, *use[rep % len(use)])  # eager work between the replays


def test_lockstep_autotune(pairs, singles):
    p = xbach.Params.default()
    use = pairs[:-1]
    w = [t.shape[1] for t, _ in use]
    h = [t.shape[0] for t, _ in use]
    with xbach.Context(0) as c, xbach.Batch(0, workers=8, flags=xbach.BATCH_LOCKSTEP) as b:
        bufs = []
        try:
            for t, s in use:
                dt, ds, dc, dd = (c.buf_alloc(t.nbytes) for _ in range(4))
                c.buf_upload(dt, t)
                c.buf_upload(ds, s)
                bufs.append((dt, ds, dc, dd))
            args = ([q[0] for q in bufs], [q[1] for q in bufs], [q[2] for q in bufs],
                    [q[3] for q in bufs])
            table = b.autotune(p, w, h, *args, reps=1)
            assert all(0 <= t <= 8 for t in table) and max(table) > 0, table
            res = b.run_device(p, w, h, *args)
            for i, (t, _) in enumerate(use):
                conv = np.empty_like(t)
                diff = np.empty_like(t)
                c.buf_download(conv, bufs[i][2])
                c.buf_download(diff, bufs[i][3])
                assert res[i].status == 0
                assert _same(conv, singles[i][0]) and _same(diff, singles[i][1]), i
        finally:
            for q in bufs:
                for d in q:
                    c.buf_free(d)
