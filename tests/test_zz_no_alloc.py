"""The allocation-free gradient entry point (include/xbach.h
xb_mlp_grad_noalloc) against xb_mlp_grad, through the C ABI.

Both run the same gradient kernels over the same block order, so the results
must agree bit for bit.  Two calls in a row share the library's static
per-device scratch: the first at the documented maximum of samples, the second
with fewer, so partial sums the first call left in the scratch must not reach
the second result.  A size above the maximum is refused before any launch.
"""
import numpy as np
import pytest

import xbach

from test_mlp import data, random_model


def _grad_noalloc(ctx, model, X, y, wt):
    n = len(X)
    bufs = []
    try:
        def up(a):
            a = np.ascontiguousarray(a, np.float32)
            d = ctx.buf_alloc(a.nbytes)
            bufs.append(d)
            ctx.buf_upload(d, a)
            return d
        dm, dX, dy, dw = up(model), up(X), up(y), up(wt)
        out = np.full(xbach.MLP_NGRAD + 1, np.nan, np.float32)
        dg = up(out)
        rc = ctx.mlp_grad_noalloc(dm, dX, dy, dw, n, dg)
        assert rc == 0, (rc, ctx._L.xb_last_error(ctx._h))
        ctx.sync()
        ctx.buf_download(out, dg)
        return out
    finally:
        for d in bufs:
            ctx.buf_free(d)


@pytest.mark.gpu
def test_noalloc_gradient_equals_allocating_entry_point():
    rng = np.random.default_rng(11)
    model = random_model(rng)
    with xbach.Context(0) as ctx:
        for n in (xbach.MLP_NOALLOC_MAX, 701):
            X, y, wt = data(rng, n)
            print(f"xb_mlp_grad_noalloc n={n}", flush=True)
            got = _grad_noalloc(ctx, model, X, y, wt)
            g, loss = ctx.mlp_grad(model, X, y, wt)
            ref = np.append(g, np.float32(loss))
            assert np.all(np.isfinite(got))
            assert np.array_equal(got, ref), (n, np.abs(got - ref).max())


@pytest.mark.gpu
def test_noalloc_gradient_refuses_sizes_above_the_maximum():
    with xbach.Context(0) as ctx:
        n = xbach.MLP_NOALLOC_MAX + 1
        # refused on the host before anything is launched: the pointers are
        # never dereferenced
        rc = ctx.mlp_grad_noalloc(1, 1, 1, 1, n, 1)
        assert rc == 8  # XB_ERR_CAPACITY
        assert b"exceeds the maximum" in ctx._L.xb_last_error(ctx._h)
