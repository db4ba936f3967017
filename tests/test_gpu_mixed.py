"""The mixed-precision form (XB_FORM_MIXED): cmv's basis convolution on the
matrix cores in BF16 split precision (three BF16 parts, six part products,
FP32 accumulation) or FP16 split precision (XB_FORM_FP16) and its
normal-matrix products on the FP64 matrix cores; everything else FP64.

Bars:
  * the stamp fit data read back per stamp (xb_get_stamp_fit) of the plain
    form equal the oracle's fillStamps/createQ/createB (cl_kern/cmv.cl:70-248)
    bit for bit — this pins how the read-back maps onto the reference's W, Q, B;
  * the mixed form's W, Q, B against the oracle in full FP64: per W row,
    max |dW| <= 5e-6 max |W|; |dQ_ij| <= 5e-6 sqrt(Q_ii Q_jj);
    |dB_i| <= 5e-6 sqrt(Q_ii * sum(target^2)) (the split is exact in FP32,
    the dropped part products are below 2^-24, and FP32 accumulation over
    504 taps leaves ~1e-6 in W, which Q and B inherit; measured worst 1.7e-6
    on pair512);
  * the same with W in FP16 split precision (XB_FORM_FP16: two FP16 parts,
    22 bits, range-scaled): within 2e-5;
  * end to end on the 2048x2048 pair, for the BF16 leg, the FP16 leg, the
    BF16 leg with FP64 iterative refinement of every global solve, and the BF16
    leg with the global solves by L D L^T (XB_FORM_LDLT), unrefined and refined: the
    substamps selected (sss) are the plain form's, bit for bit; the difference
    image's residual RMS over good interior pixels is within 1 % of the FP64
    oracle's, its pixels within 1e-3 sigma_pix = sqrt(|T| + |S|) of the FP64
    oracle's, and the kernel coefficients normwise within 1e-2 (reported per
    component and normwise: the ill-conditioned system leaves some directions
    loosely determined, so per-component agreement at 1e-6 is out of reach for
    any W that differs from the FP64 one, refined or not).
"""
import numpy as np
import pytest

import _compare
import _oracle
import xbach

pytestmark = pytest.mark.gpu


def even_rows():
    """Even-even basis rows n > 0 (src/bach.cpp:181-188 order: Gaussian g,
    then x degree, then y degree)."""
    out = []
    for dg in (6, 4, 2):
        for ix in range(dg + 1):
            for iy in range(dg - ix + 1):
                out.append(len(out) > 0 and ix % 2 == 0 and iy % 2 == 0)
    return np.array(out + [False] * 3)


EVEN = even_rows()


def gpu_fit(ctx, im, k):
    f = ctx.stamp_fit(im, k)
    W = f["W"].copy()
    W[EVEN] -= W[0]
    G = np.tril(f["G"])
    G = G + np.tril(G, -1).T
    Q = np.zeros((51, 51))
    Q[1:, 1:] = G[:50, :50]
    B = np.zeros(51)
    B[1:] = f["Bc"][:50]
    return W, Q, B, f


def stamps_after_cmv(ctx, params, t, s, form):
    ctx.set_form(form)
    ctx.init(params, t, s)
    ctx.sss()
    ctx.cmv()
    return ctx.trace()


@pytest.fixture(scope="module")
def ctx():
    c = xbach.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def pair512():
    return xbach.synth_pair(xbach.synth_cfg("pair512"))


def test_plain_stamp_fit_readback_is_the_reference_fill(ctx, pair512):
    p = xbach.Params.default()
    t, s = pair512
    tr = stamps_after_cmv(ctx, p, t, s, xbach.FORM_PLAIN)
    for im in (0, 1):
        for k in (0, 7, tr.stampCount[im] - 1):
            W, Q, B, f = gpu_fit(ctx, im, k)
            img, img2 = (t, s) if im == 0 else (s, t)
            Wr, Qr, Br = _oracle.fill_one(p, img, img2, f["ssx"], f["ssy"])
            assert np.array_equal(W, Wr), (im, k)
            assert np.array_equal(np.tril(Q), np.tril(Qr)), (im, k)
            assert np.array_equal(B, Br), (im, k)


LEGS = {"bf16": xbach.FORM_MIXED, "fp16": xbach.FORM_MIXED | xbach.FORM_FP16}
FILL_BAR = {"bf16": 5e-6, "fp16": 2e-5}


@pytest.mark.parametrize("leg", ["bf16", "fp16"])
def test_mixed_stamp_fit_against_fp64(ctx, pair512, leg):
    p = xbach.Params.default()
    t, s = pair512
    tr = stamps_after_cmv(ctx, p, t, s, LEGS[leg])
    worst = {"W": 0.0, "Q": 0.0, "B": 0.0}
    for im in (0, 1):
        for k in range(0, tr.stampCount[im], 5):
            W, Q, B, f = gpu_fit(ctx, im, k)
            img, img2 = (t, s) if im == 0 else (s, t)
            W64, Q64, B64 = _oracle.fill_one(p, img, img2, f["ssx"], f["ssy"],
                                             mode=_oracle.MODE_FP64)
            dw = np.max(np.abs(W - W64), axis=1) / np.maximum(np.max(np.abs(W64), axis=1), 1e-300)
            d = np.sqrt(np.abs(np.diag(Q64)))[1:]
            dq = np.abs(np.tril(Q - Q64))[1:, 1:] / np.maximum(np.outer(d, d), 1e-300)
            x0, y0 = f["ssx"] - 15, f["ssy"] - 15
            tgt = img2[y0:y0 + 31, x0:x0 + 31].astype(np.float64)
            db = np.abs(B - B64)[1:] / np.maximum(d * np.sqrt(np.sum(tgt * tgt)), 1e-300)
            worst["W"] = max(worst["W"], float(dw.max()))
            worst["Q"] = max(worst["Q"], float(dq.max()))
            worst["B"] = max(worst["B"], float(db.max()))
    ctx.set_form(xbach.FORM_PLAIN)
    print(f"mixed fill ({leg}) errors vs FP64:", worst)
    bar = FILL_BAR[leg]
    assert worst["W"] <= bar and worst["Q"] <= bar and worst["B"] <= bar, worst


def coefficient_errors(x, x64):
    """Kernel-solution errors against the FP64 oracle's: largest per-component
    error relative to that component (over components above 1e-6 of the
    largest), the largest error relative to the largest component, and the
    normwise error ||x - x64|| / ||x64||."""
    x, x64 = np.asarray(x, np.float64), np.asarray(x64, np.float64)
    d = np.abs(x - x64)
    big = np.abs(x64) > 1e-6 * np.abs(x64).max()
    return {"per_component": float(np.max(d[big] / np.abs(x64[big]))),
            "of_largest": float(d.max() / np.abs(x64).max()),
            "normwise": float(np.linalg.norm(x - x64) / np.linalg.norm(x64))}


@pytest.fixture(scope="module")
def pair2048_fp64():
    p = xbach.Params.default()
    t, s = xbach.synth_pair(xbach.synth_cfg("pair2048"))
    _, d64, _, tr64, _ = _oracle.run(p, t, s, mode=_oracle.MODE_FP64)
    return t, s, d64, tr64.as_dict()


@pytest.mark.parametrize("leg", ["bf16", "fp16", "bf16+refine", "bf16+ldlt", "bf16+ldlt+refine"])
def test_mixed_pipeline_2048_against_fp64_oracle(ctx, pair2048_fp64, leg):
    p = xbach.Params.default()
    t, s, d64, tr64 = pair2048_fp64
    form = (LEGS[leg.split("+")[0]] | (xbach.FORM_REFINE if "refine" in leg else 0)
            | (xbach.FORM_LDLT if "ldlt" in leg else 0))
    ctx.set_form(xbach.FORM_PLAIN)
    _, dplain, _ = ctx.run(p, t, s)
    trp = ctx.trace().as_dict()
    ctx.set_form(form)
    try:
        _, dmix, _ = ctx.run(p, t, s)
        trm = ctx.trace().as_dict()
    finally:
        ctx.set_form(xbach.FORM_PLAIN)
    for k in ("sky0", "sky1", "fwhm0", "fwhm1", "ssX0", "ssX1", "ssY0", "ssY1", "ssVal0",
              "ssVal1", "stampIndex0", "stampIndex1"):
        assert _compare._eq(trm[k], trp[k]), k
    r64 = _compare.residual_rms(d64)
    rm = _compare.residual_rms(dmix)
    rp = _compare.residual_rms(dplain)
    sp = _compare.sigma_pix(t, s)
    good = (d64 != np.float32(1e-30)) & (dmix != np.float32(1e-30))
    rel = np.abs(dmix[good].astype(np.float64) - d64[good]) / sp[good]
    ce = coefficient_errors(trm["solution"], tr64["solution"])
    cp = coefficient_errors(trp["solution"], tr64["solution"])
    print(f"{leg}: residual RMS fp64 {r64:.6f} plain {rp:.6f} mixed {rm:.6f} "
          f"({(rm - r64) / r64:+.3%}); |d diff|/sigma_pix median {np.median(rel):.2e} "
          f"p99 {np.percentile(rel, 99):.2e} max {rel.max():.2e}; coefficient errors "
          f"{ce} (plain form {cp}); ksc iterations fp64 {tr64['kscIterations']} "
          f"plain {trp['kscIterations']} mixed {trm['kscIterations']}")
    assert abs(rm - r64) <= 0.01 * r64
    assert rel.max() <= 1e-3
    assert ce["normwise"] <= 1e-2
