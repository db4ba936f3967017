"""ctypes binding of the CPU oracle (oracle/liboracle.so) — test infrastructure.

Only tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this;
the product path (libxbach) never does.
"""
from __future__ import annotations

import ctypes as C
import pathlib
import sys

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "thesis-emx-bach_amd"))
import xbach  # noqa: E402

ORACLE_PATH = ROOT / "oracle" / "liboracle.so"
MODE_REF, MODE_FP64 = 0, 1

_O = None


def lib() -> C.CDLL:
    global _O
    if _O is not None:
        return _O
    if not ORACLE_PATH.exists():
        raise RuntimeError(f"{ORACLE_PATH} missing: run `make -C oracle`")
    O = C.CDLL(str(ORACLE_PATH))
    P = C.POINTER
    vp = C.c_void_p
    O.xo_run.argtypes = [P(xbach.Params), C.c_int, C.c_int, vp, vp, C.c_int, C.c_int, vp, vp,
                         vp, P(xbach.Trace), vp]
    O.xo_ran1_seq.argtypes = [C.c_int, vp]
    O.xo_sigma_clip.argtypes = [vp, C.c_int, C.c_int, C.c_double, P(C.c_double), P(C.c_double)]
    O.xo_lu_solve.argtypes = [vp, C.c_int, vp, vp]
    O.xo_lu_solve_stamp.argtypes = [vp, C.c_int, vp]
    O.xo_basis.argtypes = [P(xbach.Params), vp, vp, vp]
    O.xo_bitonic.argtypes = [vp, C.c_int]
    O.xo_histogram.argtypes = [vp, C.c_int, vp, C.c_int, C.c_double, C.c_double, C.c_double,
                               C.c_double, P(C.c_double), P(C.c_double)]
    O.xo_fill_one.argtypes = [P(xbach.Params), C.c_int, C.c_int, C.c_int, vp, vp, C.c_int,
                              C.c_int, vp, vp, vp]
    O.xo_capture.argtypes = [C.c_int]
    O.xo_captured_count.restype = C.c_int
    O.xo_captured_get.argtypes = [C.c_int, vp, vp]
    _O = O
    return O


def run(params, tmpl, sci, mode=MODE_REF, threads=0):
    """Whole path on the CPU: (conv, diff, mask, trace, stage_ms)."""
    t = np.ascontiguousarray(tmpl, np.float32)
    s = np.ascontiguousarray(sci, np.float32)
    h, w = t.shape
    conv = np.empty_like(t)
    diff = np.empty_like(t)
    mask = np.empty(t.shape, np.uint16)
    tr = xbach.Trace()
    ms = np.zeros(9)
    rc = lib().xo_run(C.byref(params), w, h, t.ctypes.data, s.ctypes.data, mode, threads,
                      conv.ctypes.data, diff.ctypes.data, mask.ctypes.data, C.byref(tr),
                      ms.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"oracle failed: {xbach.STATUS_NAMES.get(rc, rc)}")
    return conv, diff, mask, tr, ms


def ran1(n: int) -> np.ndarray:
    out = np.empty(n)
    lib().xo_ran1_seq(n, out.ctypes.data)
    return out


def sigma_clip(data, max_iter=3, alpha=3.0):
    d = np.ascontiguousarray(data, np.float64)
    m, s = C.c_double(), C.c_double()
    lib().xo_sigma_clip(d.ctypes.data, len(d), max_iter, alpha, C.byref(m), C.byref(s))
    return m.value, s.value


def lu_solve(a, b):
    """NR ludcmp + lubksb, 1-based: a is n x n, b length n (0-based inputs)."""
    n = a.shape[0]
    A = np.zeros((n + 1, n + 1))
    A[1:, 1:] = a
    B = np.zeros(n + 1)
    B[1:] = b
    idx = np.zeros(n + 1, np.int32)
    rc = lib().xo_lu_solve(A.ctypes.data, n, B.ctypes.data, idx.ctypes.data)
    return rc, B[1:].copy()


def basis(params):
    nps = 49
    kw = params.fKernelWidth
    fx = np.zeros(nps * kw)
    fy = np.zeros(nps * kw)
    vec = np.zeros(nps * kw * kw)
    n = lib().xo_basis(C.byref(params), fx.ctypes.data, fy.ctypes.data, vec.ctypes.data)
    return n, fx.reshape(nps, kw), fy.reshape(nps, kw), vec.reshape(nps, kw, kw)


def bitonic(v):
    a = np.ascontiguousarray(v, np.float64).copy()
    lib().xo_bitonic(a.ctypes.data, len(a))
    return a


def histogram(pix, sorted_samples, sample_count, mean, inv_sd, iq=1.35, alpha=3.0):
    p = np.ascontiguousarray(pix, np.float64)
    s = np.ascontiguousarray(sorted_samples, np.float64)
    sky, fw = C.c_double(), C.c_double()
    lib().xo_histogram(p.ctypes.data, len(p), s.ctypes.data, sample_count, mean, inv_sd, iq,
                       alpha, C.byref(sky), C.byref(fw))
    return sky.value, fw.value


def fill_one(params, img, img2, ssx, ssy, mode=MODE_REF):
    im = np.ascontiguousarray(img, np.float64)
    im2 = np.ascontiguousarray(img2, np.float64)
    h, w = im.shape
    W = np.zeros((52, 961))
    Q = np.zeros((51, 51))
    B = np.zeros(51)
    lib().xo_fill_one(C.byref(params), mode, w, h, im.ctypes.data, im2.ctypes.data, ssx, ssy,
                      W.ctypes.data, Q.ctypes.data, B.ctypes.data)
    return W, Q, B


def captured_matrices():
    O = lib()
    out = []
    for i in range(O.xo_captured_count()):
        M = np.zeros((293, 293))
        r = np.zeros(298)
        O.xo_captured_get(i, M.ctypes.data, r.ctypes.data)
        out.append((M, r))
    return out
