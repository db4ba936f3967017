"""Python host binding of libxbach, the MI355X-native X-BACH subtraction path.

The product is the C-ABI library ``libxbach.so`` (``include/xbach.h``) and the
C++ CLI ``xbach`` built next to it.  This module mirrors the reference's stage
interface (include/bach.h:66-87: init, sss, cmv, cd, ksc, conv, sub) over that
C ABI with ctypes, so Python callers (tests, the benchmark) drive exactly the
code the CLI runs.  Loading fails loudly when the library has not been built:
there is no CPU fallback.

FITS helpers read/write single-HDU BITPIX -32/-64 images (the reference reads
with CCfits, src/fitsUtil.cpp:8-38, and writes FLOAT_IMG, src/fitsUtil.cpp:40-64).
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib

import numpy as np

PKG_DIR = pathlib.Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libxbach.so"

MAX_STAMPS = 1024
MAX_SS = 16
MAX_SOL = 512

# status codes (include/xbach.h)
XB_OK = 0
STATUS_NAMES = {
    0: "XB_OK", 1: "XB_ERR_ARG", 2: "XB_ERR_SIZE_MISMATCH", 3: "XB_ERR_NO_SUBSTAMPS",
    4: "XB_ERR_STAMP_PIXELS", 5: "XB_ERR_HIP", 6: "XB_ERR_NO_DEVICE",
    7: "XB_ERR_NOT_BUILT", 8: "XB_ERR_CAPACITY", 9: "XB_ERR_IO",
}

# mask bits (include/datatypeUtil.h:172-188)
MASK_BAD_PIX_VAL = 1 << 0
MASK_SAT_PIXEL = 1 << 1
MASK_LOW_PIXEL = 1 << 2
MASK_NAN_PIXEL = 1 << 3
MASK_BAD_CONV = 1 << 4
MASK_OK_CONV = 1 << 6
MASK_BAD_INPUT = 1 << 7
MASK_BAD_PIXEL_T = 1 << 8
MASK_SKIP_T = 1 << 9
MASK_BAD_PIXEL_S = 1 << 10
MASK_SKIP_S = 1 << 11
MASK_BAD_OUTPUT = 1 << 12


class Params(C.Structure):
    """The reference's ``Arguments`` (include/argsUtil.h:7-57) minus I/O names."""

    _fields_ = [
        ("stampsx", C.c_int32), ("stampsy", C.c_int32),
        ("threshLow", C.c_double), ("threshHigh", C.c_double),
        ("threshKernFit", C.c_double), ("sigKernFit", C.c_double),
        ("sigClipAlpha", C.c_double), ("iqRange", C.c_double),
        ("maxKSStamps", C.c_int32),
        ("hSStampWidth", C.c_int32), ("fSStampWidth", C.c_int32),
        ("hKernelWidth", C.c_int32), ("fKernelWidth", C.c_int32),
        ("inSpreadMaskFactor", C.c_double),
        ("normalizeTemplate", C.c_int32),
        ("backgroundOrder", C.c_int32), ("kernelOrder", C.c_int32),
        ("ngauss", C.c_int32), ("dg", C.c_int32 * 3), ("bg", C.c_float * 3),
        ("verbose", C.c_int32),
    ]

    @classmethod
    def default(cls, **kw) -> "Params":
        p = cls()
        p.stampsx = p.stampsy = 10
        p.threshLow, p.threshHigh = 0.0, 25000.0
        p.threshKernFit, p.sigKernFit = 20.0, 2.0
        p.sigClipAlpha, p.iqRange = 3.0, 1.35
        p.maxKSStamps = 3
        p.hSStampWidth, p.fSStampWidth = 15, 31
        p.hKernelWidth, p.fKernelWidth = 10, 21
        p.inSpreadMaskFactor = 1.0
        p.normalizeTemplate = 1
        p.backgroundOrder, p.kernelOrder = 1, 2
        p.ngauss = 3
        p.dg[:] = [6, 4, 2]
        # 1/(2 sigma^2) for sigma = 0.7, 1.5, 3.0, stored as float (argsUtil.h:49-53)
        p.bg[:] = [1.0 / (2.0 * 0.7 * 0.7), 1.0 / (2.0 * 1.5 * 1.5), 1.0 / (2.0 * 3.0 * 3.0)]
        p.verbose = 0
        for k, v in kw.items():
            setattr(p, k, v)
        return p


class Trace(C.Structure):
    """Decision record of one run (include/xbach.h ``xb_trace``)."""

    _fields_ = [
        ("w", C.c_int32), ("h", C.c_int32),
        ("stampsx", C.c_int32), ("stampsy", C.c_int32),
        ("fStampWidth", C.c_int32), ("nGrid", C.c_int32),
        ("retried", C.c_int32),
        ("sky", (C.c_double * MAX_STAMPS) * 2),
        ("fwhm", (C.c_double * MAX_STAMPS) * 2),
        ("ssCount", (C.c_int32 * MAX_STAMPS) * 2),
        ("ssX", (C.c_int32 * (MAX_STAMPS * MAX_SS)) * 2),
        ("ssY", (C.c_int32 * (MAX_STAMPS * MAX_SS)) * 2),
        ("ssVal", (C.c_double * (MAX_STAMPS * MAX_SS)) * 2),
        ("stampCount", C.c_int32 * 2),
        ("stampIndex", (C.c_int32 * MAX_STAMPS) * 2),
        ("nTest", C.c_int32 * 2),
        ("kernelSumCd", C.c_double * 2),
        ("merit", C.c_double * 2),
        ("cdSolution", (C.c_double * MAX_SOL) * 2),
        ("convTemplate", C.c_int32),
        ("kscIterations", C.c_int32),
        ("nSol", C.c_int32),
        ("solution", C.c_double * MAX_SOL),
        ("finalCurrent", C.c_int32 * MAX_STAMPS),
        ("kernSum", C.c_double),
        ("captureFailures", C.c_int32),
        ("screened", C.c_int32), ("screenRestored", C.c_int32),
        ("screenConfirmed", C.c_int32),
    ]

    def as_dict(self) -> dict:
        """Numpy views of the populated part of the record."""
        n = self.nGrid
        d = {
            "w": self.w, "h": self.h, "stampsx": self.stampsx, "stampsy": self.stampsy,
            "fStampWidth": self.fStampWidth, "nGrid": n, "retried": self.retried,
            "convTemplate": self.convTemplate, "kscIterations": self.kscIterations,
            "nSol": self.nSol, "kernSum": self.kernSum,
            "stampCount": tuple(self.stampCount), "nTest": tuple(self.nTest),
            "merit": tuple(self.merit), "kernelSumCd": tuple(self.kernelSumCd),
        }
        for im in range(2):
            d[f"sky{im}"] = np.array(self.sky[im][:n])
            d[f"fwhm{im}"] = np.array(self.fwhm[im][:n])
            d[f"ssCount{im}"] = np.array(self.ssCount[im][:n])
            d[f"ssX{im}"] = np.array(self.ssX[im][: n * MAX_SS]).reshape(n, MAX_SS)
            d[f"ssY{im}"] = np.array(self.ssY[im][: n * MAX_SS]).reshape(n, MAX_SS)
            d[f"ssVal{im}"] = np.array(self.ssVal[im][: n * MAX_SS]).reshape(n, MAX_SS)
            d[f"stampIndex{im}"] = np.array(self.stampIndex[im][: self.stampCount[im]])
            d[f"cdSolution{im}"] = np.array(self.cdSolution[im][: self.nSol])
        d["solution"] = np.array(self.solution[: self.nSol])
        d["finalCurrent"] = np.array(self.finalCurrent[: self.stampCount[1 - self.convTemplate]]
                                     if self.convTemplate == 0 else
                                     self.finalCurrent[: self.stampCount[0]])
        return d


class Diag(C.Structure):
    """xb_diag: HOTPANTS-style quality figures of a result."""
    _fields_ = [("kernSum", C.c_double), ("sssig", C.c_double), ("ssscat", C.c_double),
                ("fsig", C.c_double), ("fscat", C.c_double), ("x2nrm", C.c_double),
                ("nx2nrm", C.c_int64), ("dmean", C.c_double), ("dsige", C.c_double),
                ("dsig", C.c_double), ("nok", C.c_int64), ("dmeano", C.c_double),
                ("dsigeo", C.c_double), ("dsigo", C.c_double), ("convTemplate", C.c_int32),
                ("normalizeTemplate", C.c_int32)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class Shapes(C.Structure):
    """xb_shapes: launch shapes of the kernels whose tiling is free (every
    shape gives the same bits)."""
    _fields_ = [("conv_tiles", C.c_int32), ("mask_rows", C.c_int32),
                ("ldlt_group", C.c_int32), ("find_staged", C.c_int32),
                ("stats_threads", C.c_int32), ("pad", C.c_int32)]

    @classmethod
    def make(cls, conv_tiles: int = 4, mask_rows: int = 64, ldlt_group: int = 4,
             find_staged: int = 1, stats_threads: int = 1024) -> "Shapes":
        return cls(conv_tiles, mask_rows, ldlt_group, find_staged, stats_threads, 0)

    def as_dict(self) -> dict:
        return {"conv_tiles": self.conv_tiles, "mask_rows": self.mask_rows,
                "ldlt_group": self.ldlt_group, "find_staged": self.find_staged,
                "stats_threads": self.stats_threads}


class SynthCfg(C.Structure):
    _fields_ = [
        ("w", C.c_int32), ("h", C.c_int32), ("nstars", C.c_int32),
        ("fluxMin", C.c_double), ("fluxMax", C.c_double), ("fluxSlope", C.c_double),
        ("fwhmT", C.c_double), ("fwhmS0", C.c_double), ("dfwhmX", C.c_double),
        ("dfwhmY", C.c_double), ("sky", C.c_double), ("skyGx", C.c_double),
        ("skyGy", C.c_double), ("sciFluxScale", C.c_double), ("seed", C.c_uint64),
    ]


# Seeded synthetic workloads (names give what they are).
SYNTH_CONFIGS = {
    # 512x512, 300 stars, FWHM 2.0 (template) / 3.0 (science), sky 100
    "pair512": dict(w=512, h=512, nstars=300, fluxMin=1e2, fluxMax=1e5, fluxSlope=-1.5,
                    fwhmT=2.0, fwhmS0=3.0, dfwhmX=0.0, dfwhmY=0.0, sky=100.0, skyGx=0.0,
                    skyGy=0.0, sciFluxScale=1.0, seed=512001),
    # 2048x2048, 6000 stars, science FWHM 2.5 -> 3.5 ramp, template 1.8, sky 200 + gradient
    "pair2048": dict(w=2048, h=2048, nstars=6000, fluxMin=1e2, fluxMax=1e5, fluxSlope=-1.5,
                     fwhmT=1.8, fwhmS0=2.5, dfwhmX=0.5, dfwhmY=0.5, sky=200.0, skyGx=0.02,
                     skyGy=0.01, sciFluxScale=1.0, seed=2048001),
    # 4096x4096, 20000 stars, the 2048 pair's PSFs and sky (fit on a 20x20 stamp grid)
    "pair4096": dict(w=4096, h=4096, nstars=20000, fluxMin=1e2, fluxMax=1e5, fluxSlope=-1.5,
                     fwhmT=1.8, fwhmS0=2.5, dfwhmX=0.5, dfwhmY=0.5, sky=200.0, skyGx=0.02,
                     skyGy=0.01, sciFluxScale=1.0, seed=4096001),
    # 3072x3072 crowded field, 30000 stars, FWHM 2.0 / 3.0, sky 100
    "crowd3072": dict(w=3072, h=3072, nstars=30000, fluxMin=1e2, fluxMax=1e5, fluxSlope=-1.5,
                      fwhmT=2.0, fwhmS0=3.0, dfwhmX=0.0, dfwhmY=0.0, sky=100.0, skyGx=0.0,
                      skyGy=0.0, sciFluxScale=1.0, seed=3072001),
}


def batch_cfgs(n: int, seed: int = 64064, side=(512, 3072)) -> list[SynthCfg]:
    """Seeded batch of independent pairs: square side ~ U[side], star density
    ~ U[1e-3, 5e-3] per px, template and science FWHM each ~ U[1.5, 4.0] px,
    fluxes and sky as pair512.  The draws come from numpy's PCG64 seeded with
    `seed`; pair i's pixels are seeded with seed * 1000 + i."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        s = int(rng.integers(side[0], side[1] + 1))
        dens = float(rng.uniform(1e-3, 5e-3))
        ft, fs = (float(v) for v in rng.uniform(1.5, 4.0, 2))
        out.append(synth_cfg("pair512", w=s, h=s, nstars=int(round(dens * s * s)), fwhmT=ft,
                             fwhmS0=fs, seed=seed * 1000 + i))
    return out


def synth_cfg(name: str | None = None, **kw) -> SynthCfg:
    base = dict(SYNTH_CONFIGS[name]) if name else {}
    base.update(kw)
    c = SynthCfg()
    for k, v in base.items():
        setattr(c, k, v)
    return c


# ----------------------------------------------------------------------------
# FITS (single primary HDU, BITPIX -32 / -64)
# ----------------------------------------------------------------------------
def _fits_open(path):
    path = str(path)
    if path.endswith(".gz"):
        import gzip
        with gzip.open(path, "rb") as f:
            return f.read()
    with open(path, "rb") as f:
        return f.read()


def read_fits(path) -> tuple[dict, np.ndarray]:
    """Header cards and FP64 data (h, w) of the primary HDU."""
    raw = _fits_open(path)
    hdr, off = {}, 0
    while True:
        blk = raw[off:off + 2880]
        off += 2880
        done = False
        for i in range(36):
            card = blk[i * 80:(i + 1) * 80].decode("ascii", "replace")
            key = card[:8].strip()
            if key == "END":
                done = True
                break
            if card[8:10] == "= ":
                val = card[10:].split("/")[0].strip().strip("'").strip()
                hdr[key] = val
        if done:
            break
    bitpix = int(hdr["BITPIX"])
    if bitpix not in (-32, -64):
        raise ValueError(f"fits image of type{bitpix} is not supported.")
    n1, n2 = int(hdr["NAXIS1"]), int(hdr["NAXIS2"])
    dt = ">f4" if bitpix == -32 else ">f8"
    data = np.frombuffer(raw[off:off + n1 * n2 * abs(bitpix) // 8], dtype=dt)
    data = data.reshape(n2, n1).astype(np.float64)
    bscale = float(hdr.get("BSCALE", 1) or 1)
    bzero = float(hdr.get("BZERO", 0) or 0)
    if bscale != 1.0 or bzero != 0.0:
        data = data * bscale + bzero
    return hdr, data


def write_fits(path, data: np.ndarray) -> None:
    """FLOAT_IMG primary HDU, overwriting (the reference's "!" prefix)."""
    data = np.asarray(data, dtype=">f4")
    h, w = data.shape
    cards = [
        "SIMPLE  =                    T",
        "BITPIX  =                  -32",
        "NAXIS   =                    2",
        f"NAXIS1  = {w:20d}",
        f"NAXIS2  = {h:20d}",
        "EXTEND  =                    T",
        "END",
    ]
    hdr = "".join(c.ljust(80) for c in cards).encode("ascii")
    hdr += b" " * ((2880 - len(hdr) % 2880) % 2880)
    body = data.tobytes()
    body += b"\0" * ((2880 - len(body) % 2880) % 2880)
    with open(path, "wb") as f:
        f.write(hdr + body)


# ----------------------------------------------------------------------------
# libxbach
# ----------------------------------------------------------------------------
_lib = None


class XbachError(RuntimeError):
    pass


def lib() -> C.CDLL:
    """Load libxbach.so.  Raises when it has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `make -C {PKG_DIR}` "
            "(or __graft_entry__.build()); the GPU path has no fallback")
    L = C.CDLL(str(LIB_PATH))
    P = C.POINTER
    vp = C.c_void_p
    L.xb_params_default.argtypes = [P(Params)]
    L.xb_ctx_create.argtypes = [C.c_int, C.c_int, P(vp)]
    L.xb_ctx_destroy.argtypes = [vp]
    L.xb_last_error.argtypes = [vp]
    L.xb_last_error.restype = C.c_char_p
    L.xb_device_name.argtypes = [vp, C.c_char_p, C.c_size_t]
    L.xb_init.argtypes = [vp, P(Params), C.c_int, C.c_int, vp, vp]
    L.xb_init_device.argtypes = [vp, P(Params), C.c_int, C.c_int, vp, vp]
    L.xb_sss.argtypes = [vp]
    L.xb_cmv.argtypes = [vp]
    L.xb_cd.argtypes = [vp, P(C.c_int)]
    L.xb_ksc.argtypes = [vp]
    L.xb_conv.argtypes = [vp, P(C.c_double)]
    L.xb_sub.argtypes = [vp]
    L.xb_get_conv.argtypes = [vp, vp]
    L.xb_get_diff.argtypes = [vp, vp]
    L.xb_get_mask.argtypes = [vp, vp]
    L.xb_get_trace.argtypes = [vp, P(Trace)]
    L.xb_device_outputs.argtypes = [vp, P(vp), P(vp)]
    L.xb_pipeline_run.argtypes = [vp, P(Params), C.c_int, C.c_int, vp, vp, vp, vp, vp]
    L.xb_pipeline_run_device.argtypes = [vp, P(Params), C.c_int, C.c_int, vp, vp]
    L.xb_synth_pair.argtypes = [P(SynthCfg), vp, vp]
    L.xb_lu_solve.argtypes = [vp, vp, vp, vp]
    L.xb_buf_alloc.argtypes = [vp, C.c_size_t, P(vp)]
    L.xb_buf_free.argtypes = [vp, vp]
    L.xb_buf_upload.argtypes = [vp, vp, vp, C.c_size_t]
    L.xb_buf_download.argtypes = [vp, vp, vp, C.c_size_t]
    L.xb_profile_enable.argtypes = [vp, C.c_int]
    L.xb_profile_get.argtypes = [vp, vp, vp, C.c_int]
    L.xb_fits_read.argtypes = [C.c_char_p, P(C.c_int), P(C.c_int), vp, C.c_size_t, C.c_char_p,
                               C.c_size_t]
    L.xb_fits_write.argtypes = [C.c_char_p, C.c_int, C.c_int, vp, C.c_char_p, C.c_size_t]
    L.xb_sync.argtypes = [vp]
    L.xb_set_form.argtypes = [vp, C.c_int]
    L.xb_set_shapes.argtypes = [vp, P(Shapes)]
    L.xb_get_shapes.argtypes = [vp, P(Shapes)]
    L.xb_batch_get_shapes.argtypes = [vp, C.c_int, P(Shapes)]
    L.xb_batch_set_shapes.argtypes = [vp, P(Shapes)]
    L.xb_get_diagnostics.argtypes = [vp, P(Diag)]
    L.xb_screen_load.argtypes = [vp, vp]
    L.xb_batch_screen_load.argtypes = [vp, vp]
    L.xb_screen_samples.argtypes = [vp, vp, vp, vp, C.c_int, P(C.c_int), P(C.c_int)]
    L.xb_mlp_forward.argtypes = [vp, vp, vp, C.c_int, vp]
    L.xb_mlp_grad.argtypes = [vp, vp, vp, vp, vp, C.c_int, vp]
    L.xb_mlp_grad_noalloc.argtypes = [vp, vp, vp, vp, vp, C.c_int, vp]
    L.xb_mlp_train.argtypes = [vp, vp, vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_float,
                               C.c_uint64, vp]
    L.xb_batch_set_form.argtypes = [vp, C.c_int]
    L.xb_get_stamp_fit.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp, vp, vp]
    L.xb_batch_create.argtypes = [C.c_int, C.c_int, C.c_int, P(vp)]
    L.xb_batch_destroy.argtypes = [vp]
    L.xb_batch_last_error.argtypes = [vp]
    L.xb_batch_last_error.restype = C.c_char_p
    L.xb_batch_set_inflight.argtypes = [vp, C.c_int]
    L.xb_batch_autotune.argtypes = [vp, P(Params), C.c_int, vp, vp, vp, vp, vp, vp, C.c_int, vp]
    L.xb_batch_run.argtypes = [vp, P(Params), C.c_int, vp, vp, vp, vp, vp, vp, vp]
    L.xb_batch_run_device.argtypes = [vp, P(Params), C.c_int, vp, vp, vp, vp, vp, vp, vp]
    _lib = L
    return L


def fits_read_native(path) -> np.ndarray:
    """FP32 image (h, w) read by the library's FITS reader (xb_fits_read)."""
    L = lib()
    w, h = C.c_int(), C.c_int()
    msg = C.create_string_buffer(512)
    p = str(path).encode()
    rc = L.xb_fits_read(p, C.byref(w), C.byref(h), None, 0, msg, 512)
    if rc != 0:
        raise XbachError(msg.value.decode())
    out = np.empty((h.value, w.value), np.float32)
    rc = L.xb_fits_read(p, C.byref(w), C.byref(h), out.ctypes.data, out.size, msg, 512)
    if rc != 0:
        raise XbachError(msg.value.decode())
    return out


def fits_write_native(path, data: np.ndarray) -> None:
    """FLOAT_IMG primary HDU written by the library (xb_fits_write)."""
    d = np.ascontiguousarray(data, dtype=np.float32)
    msg = C.create_string_buffer(512)
    rc = lib().xb_fits_write(str(path).encode(), d.shape[1], d.shape[0], d.ctypes.data, msg, 512)
    if rc != 0:
        raise XbachError(msg.value.decode())


# forms of the path (include/xbach.h XB_FORM_*)
# xb_ctx_create flags (include/xbach.h)
CTX_ONE_STREAM = 1
CTX_EAGER = 2  # global solves launched kernel by kernel instead of replayed hipGraphs
BATCH_LOCKSTEP = 4  # xb_batch_create: groups of pairs in lockstep, one batched solve per solve point
FORM_PLAIN, FORM_MIXED, FORM_SCREEN, FORM_RECORD = 0, 1, 2, 4
FORM_FP16, FORM_REFINE = 8, 16  # with FORM_MIXED: FP16 split W; FP64 iterative refinement
FORM_LDLT = 32  # global solves by L D L^T + one FP64 refinement step (not the reference's order)
# the screen's network (include/xbach.h XB_MLP_*)
MLP_NFEAT, MLP_NPARAM, MLP_NGRAD = 32, 2242, 2177
MLP_NOALLOC_MAX = 16384  # XB_MLP_NOALLOC_MAX

# per-kernel timing slots (include/xbach.h XB_PROF_*)
PROF_SLOTS = ["conv", "sub", "solve", "fill", "sss", "mask"]


def synth_pair(cfg: SynthCfg) -> tuple[np.ndarray, np.ndarray]:
    """Seeded synthetic (template, science) FP32 pair of shape (h, w)."""
    t = np.empty((cfg.h, cfg.w), np.float32)
    s = np.empty((cfg.h, cfg.w), np.float32)
    rc = lib().xb_synth_pair(C.byref(cfg), t.ctypes.data, s.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"xb_synth_pair failed: {STATUS_NAMES.get(rc, rc)}")
    return t, s


def synth_pairs(cfgs, threads: int = 8) -> list[tuple[np.ndarray, np.ndarray]]:
    """synth_pair over many configurations on host threads (same results)."""
    from concurrent.futures import ThreadPoolExecutor

    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        return list(ex.map(synth_pair, cfgs))


class Context:
    """One device context (xb_ctx): stream, resident buffers, stage API."""

    def __init__(self, device: int = 0, flags: int = 0):
        self._L = lib()
        h = C.c_void_p()
        rc = self._L.xb_ctx_create(device, flags, C.byref(h))
        if rc != 0:
            raise XbachError(f"xb_ctx_create failed: {STATUS_NAMES.get(rc, rc)}")
        self._h = h
        self.w = self.h = 0

    def close(self):
        if self._h:
            self._L.xb_ctx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            msg = self._L.xb_last_error(self._h)
            raise XbachError(f"{what}: {STATUS_NAMES.get(rc, rc)}: "
                             f"{msg.decode() if msg else ''}")

    @property
    def device_name(self) -> str:
        buf = C.create_string_buffer(256)
        self._L.xb_device_name(self._h, buf, 256)
        return buf.value.decode()

    # ---- stages (include/bach.h:66-87) -------------------------------------
    def init(self, params: Params, tmpl: np.ndarray, sci: np.ndarray):
        t = np.ascontiguousarray(tmpl, np.float32)
        s = np.ascontiguousarray(sci, np.float32)
        self.h, self.w = t.shape
        self._check(self._L.xb_init(self._h, C.byref(params), self.w, self.h,
                                    t.ctypes.data, s.ctypes.data), "xb_init")

    def init_device(self, params: Params, w: int, h: int, d_tmpl: int, d_sci: int):
        self.w, self.h = w, h
        self._check(self._L.xb_init_device(self._h, C.byref(params), w, h, d_tmpl, d_sci),
                    "xb_init_device")

    def sss(self):
        self._check(self._L.xb_sss(self._h), "xb_sss")

    def cmv(self):
        self._check(self._L.xb_cmv(self._h), "xb_cmv")

    def cd(self) -> bool:
        ct = C.c_int()
        self._check(self._L.xb_cd(self._h, C.byref(ct)), "xb_cd")
        return bool(ct.value)

    def ksc(self):
        self._check(self._L.xb_ksc(self._h), "xb_ksc")

    def conv(self) -> float:
        ks = C.c_double()
        self._check(self._L.xb_conv(self._h, C.byref(ks)), "xb_conv")
        return ks.value

    def sub(self):
        self._check(self._L.xb_sub(self._h), "xb_sub")

    # ---- results ------------------------------------------------------------
    def conv_image(self) -> np.ndarray:
        out = np.empty((self.h, self.w), np.float32)
        self._check(self._L.xb_get_conv(self._h, out.ctypes.data), "xb_get_conv")
        return out

    def diff_image(self) -> np.ndarray:
        out = np.empty((self.h, self.w), np.float32)
        self._check(self._L.xb_get_diff(self._h, out.ctypes.data), "xb_get_diff")
        return out

    def mask(self) -> np.ndarray:
        out = np.empty((self.h, self.w), np.uint16)
        self._check(self._L.xb_get_mask(self._h, out.ctypes.data), "xb_get_mask")
        return out

    def trace(self) -> Trace:
        tr = Trace()
        self._check(self._L.xb_get_trace(self._h, C.byref(tr)), "xb_get_trace")
        return tr

    def run(self, params: Params, tmpl: np.ndarray, sci: np.ndarray):
        """Whole path; returns (conv, diff, stage_ms[9])."""
        t = np.ascontiguousarray(tmpl, np.float32)
        s = np.ascontiguousarray(sci, np.float32)
        self.h, self.w = t.shape
        conv = np.empty_like(t)
        diff = np.empty_like(t)
        ms = np.zeros(9, np.float64)
        self._check(self._L.xb_pipeline_run(self._h, C.byref(params), self.w, self.h,
                                             t.ctypes.data, s.ctypes.data, conv.ctypes.data,
                                             diff.ctypes.data, ms.ctypes.data),
                    "xb_pipeline_run")
        return conv, diff, ms

    def run_device(self, params: Params, w: int, h: int, d_tmpl: int, d_sci: int):
        self.w, self.h = w, h
        self._check(self._L.xb_pipeline_run_device(self._h, C.byref(params), w, h,
                                                   d_tmpl, d_sci), "xb_pipeline_run_device")

    def lu_solve(self, a: np.ndarray, b: np.ndarray) -> np.ndarray:
        """Global solve (ludcmp + lubksb) of a 293x293 1-based system."""
        A = np.ascontiguousarray(a, np.float64)
        B = np.zeros(298)
        B[: len(b)] = b
        if A.shape != (293, 293):
            raise ValueError("expects the 293x293 (1-based) global system")
        x = np.empty(298)
        self._check(self._L.xb_lu_solve(self._h, A.ctypes.data, B.ctypes.data, x.ctypes.data),
                    "xb_lu_solve")
        return x

    # ---- device buffers / profiling -----------------------------------------
    def buf_alloc(self, nbytes: int) -> int:
        d = C.c_void_p()
        self._check(self._L.xb_buf_alloc(self._h, nbytes, C.byref(d)), "xb_buf_alloc")
        return d.value

    def buf_free(self, d: int):
        self._check(self._L.xb_buf_free(self._h, d), "xb_buf_free")

    def buf_upload(self, d: int, arr: np.ndarray):
        a = np.ascontiguousarray(arr)
        self._check(self._L.xb_buf_upload(self._h, d, a.ctypes.data, a.nbytes), "xb_buf_upload")

    def buf_download(self, arr: np.ndarray, d: int):
        self._check(self._L.xb_buf_download(self._h, arr.ctypes.data, d, arr.nbytes),
                    "xb_buf_download")

    def profile(self, enable: bool = True):
        self._check(self._L.xb_profile_enable(self._h, 1 if enable else 0), "xb_profile_enable")

    def profile_get(self) -> dict:
        ms = np.zeros(len(PROF_SLOTS))
        cnt = np.zeros(len(PROF_SLOTS), np.int32)
        self._check(self._L.xb_profile_get(self._h, ms.ctypes.data, cnt.ctypes.data,
                                           len(PROF_SLOTS)), "xb_profile_get")
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(PROF_SLOTS)}

    def sync(self):
        """Wait for the work queued on the context (xb_sync)."""
        self._check(self._L.xb_sync(self._h), "xb_sync")

    def set_form(self, form: int):
        """FORM_PLAIN (the reference's arithmetic) or FORM_MIXED (cmv on MFMA in
        BF16 split precision); applies from the next init."""
        self._check(self._L.xb_set_form(self._h, form), "xb_set_form")

    def set_shapes(self, shapes: "Shapes | None"):
        """Launch shapes (None: defaults)."""
        self._check(self._L.xb_set_shapes(self._h, C.byref(shapes) if shapes else None),
                    "xb_set_shapes")

    def shapes(self) -> "Shapes":
        s = Shapes()
        self._check(self._L.xb_get_shapes(self._h, C.byref(s)), "xb_get_shapes")
        return s

    def screen_load(self, model: np.ndarray):
        m = np.ascontiguousarray(model, np.float32)
        if m.size != MLP_NPARAM:
            raise ValueError(f"screen model has {MLP_NPARAM} floats")
        self._check(self._L.xb_screen_load(self._h, m.ctypes.data), "xb_screen_load")

    def screen_samples(self):
        """(features[n][32], labels[n], prob[n], n_screened) of the last run."""
        n, scr = C.c_int(), C.c_int()
        self._check(self._L.xb_screen_samples(self._h, None, None, None, 0, C.byref(n),
                                              C.byref(scr)), "xb_screen_samples")
        f = np.empty((n.value, MLP_NFEAT), np.float32)
        lab = np.empty(n.value, np.int32)
        pr = np.empty(n.value, np.float32)
        if n.value:
            self._check(self._L.xb_screen_samples(self._h, f.ctypes.data, lab.ctypes.data,
                                                  pr.ctypes.data, n.value, C.byref(n),
                                                  C.byref(scr)), "xb_screen_samples")
        return f, lab, pr, scr.value

    def mlp_forward(self, model, X) -> np.ndarray:
        m = np.ascontiguousarray(model, np.float32)
        X = np.ascontiguousarray(X, np.float32)
        p = np.empty(len(X), np.float32)
        self._check(self._L.xb_mlp_forward(self._h, m.ctypes.data, X.ctypes.data, len(X),
                                           p.ctypes.data), "xb_mlp_forward")
        return p

    def mlp_grad(self, model, X, y, wt) -> tuple[np.ndarray, float]:
        m = np.ascontiguousarray(model, np.float32)
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        wt = np.ascontiguousarray(wt, np.float32)
        g = np.empty(MLP_NGRAD + 1, np.float32)
        self._check(self._L.xb_mlp_grad(self._h, m.ctypes.data, X.ctypes.data, y.ctypes.data,
                                        wt.ctypes.data, len(X), g.ctypes.data), "xb_mlp_grad")
        return g[:MLP_NGRAD], float(g[MLP_NGRAD])

    def mlp_grad_noalloc(self, d_model: int, d_X: int, d_y: int, d_wt: int, n: int,
                         d_grad: int) -> int:
        """xb_mlp_grad_noalloc on device pointers: enqueues the gradient and
        returns its status code (0 = queued; wait with sync())."""
        return int(self._L.xb_mlp_grad_noalloc(self._h, d_model, d_X, d_y, d_wt, n, d_grad))

    def mlp_train(self, model, X, y, wt, epochs=40, batch=1024, lr=3e-3, seed=1):
        m = np.array(model, np.float32)
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        wt = np.ascontiguousarray(wt, np.float32)
        loss = np.empty(epochs, np.float32)
        self._check(self._L.xb_mlp_train(self._h, m.ctypes.data, X.ctypes.data, y.ctypes.data,
                                         wt.ctypes.data, len(X), epochs, batch, lr, seed,
                                         loss.ctypes.data), "xb_mlp_train")
        return m, loss

    def diagnostics(self) -> dict:
        d = Diag()
        self._check(self._L.xb_get_diagnostics(self._h, C.byref(d)), "xb_get_diagnostics")
        return d.as_dict()

    def stamp_fit(self, im: int, k: int) -> dict:
        """W (52x961 raw rows), G (52x52 lower), Bc, Br and the substamp centre
        of kept stamp k of image im, as the last fill left them."""
        W = np.empty((52, 961))
        G = np.empty((52, 52))
        Bc = np.empty(52)
        Br = np.empty(3)
        xy = np.empty(2, np.int32)
        self._check(self._L.xb_get_stamp_fit(self._h, im, k, W.ctypes.data, G.ctypes.data,
                                             Bc.ctypes.data, Br.ctypes.data, xy.ctypes.data),
                    "xb_get_stamp_fit")
        return {"W": W, "G": G, "Bc": Bc, "Br": Br, "ssx": int(xy[0]), "ssy": int(xy[1])}

    def device_outputs(self) -> tuple[int, int]:
        c, d = C.c_void_p(), C.c_void_p()
        self._check(self._L.xb_device_outputs(self._h, C.byref(c), C.byref(d)),
                    "xb_device_outputs")
        return c.value, d.value


class PairResult(C.Structure):
    """xb_pair_result: per-pair outcome of a batch run."""
    _fields_ = [("status", C.c_int32), ("convTemplate", C.c_int32),
                ("kscIterations", C.c_int32), ("worker", C.c_int32),
                ("kernSum", C.c_double), ("ms", C.c_double),
                ("captureFailures", C.c_int32), ("screened", C.c_int32),
                ("screenRestored", C.c_int32), ("pad", C.c_int32)]


def _ptr_array(ptrs) -> C.Array:
    return (C.c_void_p * max(len(ptrs), 1))(*[p if p else None for p in ptrs])


class Batch:
    """Many independent pairs through the pipeline at once (xb_batch): worker
    threads, each with its own context, pull pairs largest first."""

    def __init__(self, device: int = 0, workers: int = 4, flags: int = 0):
        self._L = lib()
        h = C.c_void_p()
        self._h = None
        rc = self._L.xb_batch_create(device, workers, flags, C.byref(h))
        if rc != 0:
            raise XbachError(f"xb_batch_create failed: {STATUS_NAMES.get(rc, rc)}")
        self._h = h
        self.workers = workers

    def close(self):
        if self._h:
            self._L.xb_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            msg = self._L.xb_batch_last_error(self._h)
            raise XbachError(f"{what}: {STATUS_NAMES.get(rc, rc)}: "
                             f"{msg.decode() if msg else ''}")

    def set_form(self, form: int):
        self._check(self._L.xb_batch_set_form(self._h, form), "xb_batch_set_form")

    def screen_load(self, model: np.ndarray):
        m = np.ascontiguousarray(model, np.float32)
        self._check(self._L.xb_batch_screen_load(self._h, m.ctypes.data),
                    "xb_batch_screen_load")

    def autotune(self, params: Params, w, h, d_tmpl, d_sci, d_conv=None, d_diff=None,
                 reps: int = 2) -> list[int]:
        """Pairs in flight per size bucket (<=1024, <=2048, <=3072, larger)."""
        n = len(w)
        wa = np.ascontiguousarray(w, np.int32)
        ha = np.ascontiguousarray(h, np.int32)
        table = np.zeros(4, np.int32)
        self._check(self._L.xb_batch_autotune(
            self._h, C.byref(params), n, wa.ctypes.data, ha.ctypes.data, _ptr_array(d_tmpl),
            _ptr_array(d_sci), _ptr_array(d_conv) if d_conv is not None else None,
            _ptr_array(d_diff) if d_diff is not None else None, reps, table.ctypes.data),
            "xb_batch_autotune")
        return [int(v) for v in table]

    def set_inflight(self, n: int):
        self._check(self._L.xb_batch_set_inflight(self._h, n), "xb_batch_set_inflight")

    def shapes(self, bucket: int) -> "Shapes":
        """Launch shapes of a size bucket (the autotuner's, or the defaults)."""
        s = Shapes()
        self._check(self._L.xb_batch_get_shapes(self._h, bucket, C.byref(s)),
                    "xb_batch_get_shapes")
        return s

    def set_shapes(self, shapes: "Shapes | None"):
        self._check(self._L.xb_batch_set_shapes(self._h, C.byref(shapes) if shapes else None),
                    "xb_batch_set_shapes")

    def run(self, params: Params, pairs, check: bool = True):
        """pairs: list of (tmpl, sci) host arrays.  Returns ([(conv, diff)], results)."""
        ts = [np.ascontiguousarray(t, np.float32) for t, _ in pairs]
        ss = [np.ascontiguousarray(s, np.float32) for _, s in pairs]
        n = len(ts)
        w = np.array([t.shape[1] for t in ts], np.int32)
        h = np.array([t.shape[0] for t in ts], np.int32)
        convs = [np.empty_like(t) for t in ts]
        diffs = [np.empty_like(t) for t in ts]
        res = (PairResult * max(n, 1))()
        rc = self._L.xb_batch_run(self._h, C.byref(params), n, w.ctypes.data, h.ctypes.data,
                                  _ptr_array([t.ctypes.data for t in ts]),
                                  _ptr_array([s.ctypes.data for s in ss]),
                                  _ptr_array([c.ctypes.data for c in convs]),
                                  _ptr_array([d.ctypes.data for d in diffs]), res)
        if check:
            self._check(rc, "xb_batch_run")
        return list(zip(convs, diffs)), [res[i] for i in range(n)]

    def run_device(self, params: Params, w, h, d_tmpl, d_sci, d_conv=None, d_diff=None,
                   check: bool = True):
        """Device-resident pairs (lists of device pointers); returns the results."""
        n = len(w)
        wa = np.ascontiguousarray(w, np.int32)
        ha = np.ascontiguousarray(h, np.int32)
        res = (PairResult * max(n, 1))()
        rc = self._L.xb_batch_run_device(
            self._h, C.byref(params), n, wa.ctypes.data, ha.ctypes.data, _ptr_array(d_tmpl),
            _ptr_array(d_sci), _ptr_array(d_conv) if d_conv is not None else None,
            _ptr_array(d_diff) if d_diff is not None else None, res)
        if check:
            self._check(rc, "xb_batch_run_device")
        return [res[i] for i in range(n)]


def diff_fits_stats(golden: np.ndarray, ours: np.ndarray) -> dict:
    """The comparison of tools/run_test.py:34-93 (diff_fits): max/mean absolute
    error, relative error only where the golden pixel is > 0, NaN mismatches."""
    g = np.asarray(golden, np.float64)
    b = np.asarray(ours, np.float64)
    nan_g, nan_b = np.isnan(g), np.isnan(b)
    wrong_nans = int(np.count_nonzero(nan_g != nan_b))
    ok = ~(nan_g | nan_b)
    err = np.abs(g - b)[ok]
    pos = (g > 0) & ok
    rel = np.abs(g - b)[pos] / g[pos]
    return {
        "max_abs": float(err.max()) if err.size else 0.0,
        "mean_abs": float(err.mean()) if err.size else 0.0,
        "max_rel": float(rel.max()) if rel.size else 0.0,
        "mean_rel": float(rel.sum() / max(int(ok.sum()), 1)),
        "wrong_nans": wrong_nans,
    }
