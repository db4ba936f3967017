"""The GPU global solve (xb_lu_solve) against the oracle's Numerical Recipes
ludcmp + lubksb (src/bachUtil.cpp:319-411) on edge cases: bitwise equal
results (NaN positions equal, NaN payloads free)."""
import ctypes as C

import numpy as np
import pytest

import _oracle
import xbach

pytestmark = pytest.mark.gpu
N = 292


@pytest.fixture(scope="module")
def ctx():
    c = xbach.Context(0)
    yield c
    c.close()


def oracle_solve(a, b):
    A = np.ascontiguousarray(a, np.float64).copy()
    B = np.zeros(N + 1)
    B[:] = b[: N + 1]
    idx = np.zeros(N + 1, np.int32)
    _oracle.lib().xo_lu_solve(A.ctypes.data, N, B.ctypes.data, idx.ctypes.data)
    return B


def same(x, y):
    x = np.asarray(x[1:N + 1])
    y = np.asarray(y[1:N + 1])
    nx, ny = np.isnan(x), np.isnan(y)
    return bool(np.array_equal(nx, ny) and np.array_equal(x[~nx].view(np.uint64),
                                                          y[~ny].view(np.uint64)))


def padded(m):
    a = np.zeros((N + 1, N + 1))
    a[1:, 1:] = m
    return a


def rhs(rng):
    b = np.zeros(298)
    b[1:N + 1] = rng.normal(size=N)
    return b


def test_normal_matrix_like(ctx):
    rng = np.random.default_rng(1)
    x = rng.normal(size=(400, N)) * np.logspace(0, 5, N)
    a = padded(x.T @ x)
    b = rhs(rng)
    assert same(ctx.lu_solve(a, b), oracle_solve(a, b))


def test_general_matrix_with_pivoting(ctx):
    rng = np.random.default_rng(2)
    m = rng.normal(size=(N, N))
    m[np.arange(0, N, 3), np.arange(0, N, 3)] = 0.0
    a = padded(m)
    b = rhs(rng)
    assert same(ctx.lu_solve(a, b), oracle_solve(a, b))


def test_pivot_ties_pick_last_row(ctx):
    rng = np.random.default_rng(3)
    perm = rng.permutation(N)
    m = np.zeros((N, N))
    m[np.arange(N), perm] = 2.0 ** rng.integers(-4, 4, N)
    m[:, 7] += 1.0  # ties between rows in several columns
    a = padded(m)
    b = rhs(rng)
    assert same(ctx.lu_solve(a, b), oracle_solve(a, b))


def test_zero_row_returns_before_factorising(ctx):
    rng = np.random.default_rng(4)
    m = rng.normal(size=(N, N))
    m[100, :] = 0.0
    a = padded(m)
    b = rhs(rng)
    assert same(ctx.lu_solve(a, b), oracle_solve(a, b))


@pytest.mark.parametrize("col", [1, 2, 40, 200, 292])
def test_nan_column_replays_reference_pivoting(ctx, col):
    rng = np.random.default_rng(col)
    m = rng.normal(size=(N, N)) + 3 * np.eye(N)
    m[:, col - 1] = np.nan
    a = padded(m)
    b = rhs(rng)
    assert same(ctx.lu_solve(a, b), oracle_solve(a, b))


def test_rhs_leading_zeros(ctx):
    # lubksb's ii rule: substitution starts at the first nonzero entry
    rng = np.random.default_rng(6)
    x = rng.normal(size=(400, N))
    a = padded(x.T @ x)
    b = rhs(rng)
    b[1:60] = 0.0
    b[5] = -0.0
    assert same(ctx.lu_solve(a, b), oracle_solve(a, b))


def _true_solution(a, b):
    """The solution to ~1e-19 relative: float64 solves corrected with residuals
    formed in extended precision (np.longdouble) until they stop changing."""
    A, B = a[1:, 1:], b[1:N + 1]
    AL, BL = A.astype(np.longdouble), B.astype(np.longdouble)
    x = np.linalg.solve(A, B).astype(np.longdouble)
    for _ in range(8):
        r = BL - AL @ x
        x = x + np.linalg.solve(A, r.astype(np.float64)).astype(np.longdouble)
    return x


@pytest.mark.parametrize("cond", [1e3, 1e13])
def test_iterative_refinement(ctx, cond):
    """XB_FORM_REFINE: one step of FP64 iterative refinement after the solve
    (residual with compensated FP64 products and sums, the same LU factors).
    Against the extended-precision solution, the refined error is never worse
    than the plain solve's and, for an ill-conditioned system (cond 1e13, where
    the plain solve keeps ~cond * eps of error), smaller by over 10x."""
    rng = np.random.default_rng(int(np.log10(cond)))
    u, _ = np.linalg.qr(rng.normal(size=(N, N)))
    v, _ = np.linalg.qr(rng.normal(size=(N, N)))
    sv = np.logspace(0, -np.log10(cond), N)
    a = padded((u * sv) @ v.T)
    b = rhs(rng)
    xt = _true_solution(a, b)
    errs = {}
    for name, form in (("plain", xbach.FORM_PLAIN), ("refined", xbach.FORM_REFINE)):
        ctx.set_form(form)
        try:
            x = ctx.lu_solve(a, b)
        finally:
            ctx.set_form(xbach.FORM_PLAIN)
        if name == "plain":
            assert same(x, oracle_solve(a, b))
        d = x[1:N + 1].astype(np.longdouble) - xt
        errs[name] = float(np.sqrt(np.sum(d * d)) / np.sqrt(np.sum(xt * xt)))
    print(f"cond {cond:.0e}: normwise error plain {errs['plain']:.2e} refined "
          f"{errs['refined']:.2e}")
    assert errs["refined"] <= max(errs["plain"], 1e-15)
    if cond >= 1e10:
        assert errs["refined"] < 0.1 * errs["plain"]


# ---------------------------------------------------------------------------
# XB_FORM_LDLT: L D L^T of the symmetric positive definite normal matrix, no
# pivoting (not the reference's order, so not bit-identical to it); bars
# against the extended-precision solution.
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def fit_systems():
    """The global normal systems the fit builds on the seeded 512x512 pair (the
    oracle's cd direction trials and clip-refit passes): exactly symmetric,
    positive definite, condition ~1e10 after diagonal scaling."""
    O = _oracle.lib()
    t, s = xbach.synth_pair(xbach.synth_cfg("pair512"))
    O.xo_capture(1)
    try:
        _oracle.run(xbach.Params.default(), t, s)
        out = []
        for k in range(O.xo_captured_count()):
            m = np.zeros(293 * 293)
            r = np.zeros(298)
            O.xo_captured_get(k, m.ctypes.data, r.ctypes.data)
            out.append((m.reshape(293, 293), r))
    finally:
        O.xo_capture(0)
    assert len(out) >= 4
    return out


def _ldlt(ctx, a, b, refine):
    ctx.set_form(xbach.FORM_LDLT | (xbach.FORM_REFINE if refine else 0))
    try:
        return ctx.lu_solve(a, b)
    finally:
        ctx.set_form(xbach.FORM_PLAIN)


def _err(x, xt):
    d = x[1:N + 1].astype(np.longdouble) - xt
    return float(np.sqrt(np.sum(d * d)) / np.sqrt(np.sum(xt * xt)))


def test_ldlt_on_the_fits_systems(ctx, fit_systems):
    """Unrefined L D L^T is as accurate as ludcmp on the fit's own systems (its
    normwise error within 5x of the reference's, measured 0.3-1x); one
    refinement step takes it to ~1e-11 (bar 1e-9, and 100x below ludcmp's).
    The solution slots outside 1..292 are the right-hand side's, as lubksb
    leaves them."""
    for k, (a, b) in enumerate(fit_systems):
        xt = _true_solution(a, b)
        e_ref = _err(oracle_solve(a, b), xt)
        x0 = _ldlt(ctx, a, b, False)
        x1 = _ldlt(ctx, a, b, True)
        e0, e1 = _err(x0, xt), _err(x1, xt)
        print(f"system {k}: normwise error ludcmp {e_ref:.2e} ldlt {e0:.2e} ldlt+refine {e1:.2e}")
        assert e0 <= max(5 * e_ref, 1e-12)
        assert e1 <= 1e-9 and e1 <= 0.01 * e_ref
        for x in (x0, x1):
            assert x[0] == b[0] and np.array_equal(x[N + 1:], b[N + 1:len(x)])


def test_ldlt_normal_matrix_like(ctx):
    rng = np.random.default_rng(5)
    x = rng.normal(size=(400, N)) * np.logspace(0, 5, N)
    a = padded(x.T @ x)
    b = rhs(rng)
    xt = _true_solution(a, b)
    e_ref = _err(oracle_solve(a, b), xt)
    e0 = _err(_ldlt(ctx, a, b, False), xt)
    e1 = _err(_ldlt(ctx, a, b, True), xt)
    print(f"normwise error ludcmp {e_ref:.2e} ldlt {e0:.2e} ldlt+refine {e1:.2e}")
    assert e0 <= max(5 * e_ref, 1e-12)
    assert e1 <= max(e0, 1e-15)


def test_ldlt_zero_row_takes_the_reference_path(ctx, fit_systems):
    """A zero row: ludcmp returns before factorising and lubksb runs on the
    unfactorised matrix; XB_FORM_LDLT does the same, bit for bit."""
    a, b = fit_systems[0]
    a = a.copy()
    a[58, :] = 0.0
    a[:, 58] = 0.0
    for refine in (False, True):
        assert same(_ldlt(ctx, a, b, refine), oracle_solve(a, b))
