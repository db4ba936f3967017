"""Comparison helpers shared by the parity tests.

Bars (stated where used):
  * every decision of the fit — stamp statistics, substamp coordinates and
    values, kept stamps, test stamps, merits, direction, solutions, refit
    iterations, final substamp indices, the pixel mask — bit-exact;
  * convolved image: |d| <= 1e-5 |ref| + 1e-3 sigma_pix on interior pixels,
    where sigma_pix = sqrt(|T| + |S|) is the per-pixel noise of the
    reference's chi2 (cl_kern/cd.cl:578);
  * difference image: |d| <= 1e-3 sigma_pix, identical 1e-30 (masked) pixels.
"""
from __future__ import annotations

import numpy as np

TRACE_EXACT_SCALARS = ["w", "h", "stampsx", "stampsy", "fStampWidth", "nGrid", "retried",
                       "convTemplate", "kscIterations", "stampCount", "nTest"]
TRACE_EXACT_ARRAYS = ["sky0", "sky1", "fwhm0", "fwhm1", "ssCount0", "ssCount1", "ssX0", "ssX1",
                      "ssY0", "ssY1", "ssVal0", "ssVal1", "stampIndex0", "stampIndex1",
                      "finalCurrent"]


def _eq(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape:
        return False
    if a.dtype.kind == "f":
        return bool(np.array_equal(a.view(np.uint64) if a.dtype == np.float64 else a,
                                   b.view(np.uint64) if b.dtype == np.float64 else b)
                    or np.array_equal(a, b, equal_nan=True))
    return bool(np.array_equal(a, b))


def trace_mismatches(got: dict, ref: dict, fit_exact: bool = True) -> list[str]:
    """Names of trace fields that are not bit-identical."""
    bad = []
    for k in TRACE_EXACT_SCALARS:
        if got[k] != ref[k]:
            bad.append(f"{k}: {got[k]} != {ref[k]}")
    for k in TRACE_EXACT_ARRAYS:
        if not _eq(got[k], ref[k]):
            g, r = np.asarray(got[k]), np.asarray(ref[k])
            if g.shape == r.shape:
                idx = np.argwhere(~((g == r) | (np.isnan(g) & np.isnan(r))
                                    if g.dtype.kind == "f" else (g == r)))
                first = tuple(idx[0]) if len(idx) else None
                bad.append(f"{k}: {len(idx)} differ, first at {first}: "
                           f"{g[first] if first is not None else ''} vs "
                           f"{r[first] if first is not None else ''}")
            else:
                bad.append(f"{k}: shape {g.shape} vs {r.shape}")
    if fit_exact:
        for k in ["merit", "kernelSumCd", "kernSum"]:
            if not _eq(np.asarray(got[k], np.float64), np.asarray(ref[k], np.float64)):
                bad.append(f"{k}: {got[k]} != {ref[k]}")
        for k in ["solution", "cdSolution0", "cdSolution1"]:
            if not _eq(got[k], ref[k]):
                d = np.abs(np.asarray(got[k]) - np.asarray(ref[k]))
                bad.append(f"{k}: max |d| {d.max():.3e}")
    return bad


def sigma_pix(tmpl, sci):
    return np.sqrt(np.abs(tmpl.astype(np.float64)) + np.abs(sci.astype(np.float64)))


def image_errors(conv, diff, conv_ref, diff_ref, tmpl, sci):
    """Worst normalised errors of conv and diff (<= 1 passes)."""
    sp = sigma_pix(tmpl, sci)
    c = conv.astype(np.float64)
    cr = conv_ref.astype(np.float64)
    d = diff.astype(np.float64)
    dr = diff_ref.astype(np.float64)
    ok = np.isfinite(cr) & np.isfinite(dr)
    tol_c = 1e-5 * np.abs(cr) + 1e-3 * sp
    conv_err = np.max(np.where(ok, np.abs(c - cr) / tol_c, 0.0))
    tol_d = 1e-3 * sp
    diff_err = np.max(np.where(ok, np.abs(d - dr) / tol_d, 0.0))
    masked_ref = dr == np.float32(1e-30)
    masked_got = d == np.float32(1e-30)
    return {
        "conv_err": float(conv_err),
        "diff_err": float(diff_err),
        "masked_mismatch": int(np.count_nonzero(masked_ref != masked_got)),
        "conv_max_abs": float(np.max(np.abs(c - cr))),
        "diff_max_abs": float(np.max(np.abs(d - dr))),
    }


def residual_rms(diff, mask_pixels=None):
    d = diff.astype(np.float64)
    good = d != np.float32(1e-30)
    if mask_pixels is not None:
        good &= mask_pixels
    return float(np.sqrt(np.mean(d[good] ** 2)))
