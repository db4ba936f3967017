"""CPU-side checks of the product library (no GPU needed): it loads, exports
every entry point include/xbach.h declares, its defaults match the reference's
Arguments, the synthetic generator is seeded and Poisson, FITS round-trips."""
import ctypes as C
import pathlib
import re

import numpy as np
import pytest

import xbach

ROOT = pathlib.Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "xbach.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(xb_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = xbach.lib()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_params_default_matches_reference_arguments():
    p = xbach.Params()
    xbach.lib().xb_params_default(C.byref(p))
    q = xbach.Params.default()
    for name, _ in xbach.Params._fields_:
        a, b = getattr(p, name), getattr(q, name)
        if hasattr(a, "__len__"):
            assert list(a) == list(b), name
        else:
            assert a == b, name
    # include/argsUtil.h:49-53: bg stored as float
    assert p.bg[0] == np.float32(1.0 / (2.0 * 0.7 * 0.7))


def test_no_gpu_is_an_error_not_a_crash():
    h = C.c_void_p()
    rc = xbach.lib().xb_ctx_create(0, 0, C.byref(h))
    # in a container without a GPU this is XB_ERR_NO_DEVICE; on a GPU box it
    # succeeds and is destroyed again
    if rc == 0:
        xbach.lib().xb_ctx_destroy(h)
    else:
        assert rc == 6
    b = C.c_void_p()
    assert xbach.lib().xb_batch_create(0, 0, 0, C.byref(b)) == 1  # no workers: XB_ERR_ARG
    rc = xbach.lib().xb_batch_create(0, 2, 0, C.byref(b))
    if rc == 0:
        xbach.lib().xb_batch_destroy(b)
    else:
        assert rc == 6 and not b.value


def test_batch_cfgs_follow_the_batching_law():
    cfgs = xbach.batch_cfgs(64)
    assert [c.seed for c in cfgs] == [64064 * 1000 + i for i in range(64)]
    sides = np.array([c.w for c in cfgs])
    assert np.array_equal(sides, [c.h for c in cfgs])
    assert sides.min() >= 512 and sides.max() <= 3072 and np.ptp(sides) > 1500
    dens = np.array([c.nstars / c.w ** 2 for c in cfgs])
    assert dens.min() > 0.99e-3 and dens.max() < 5.01e-3
    f = np.array([[c.fwhmT, c.fwhmS0] for c in cfgs])
    assert f.min() >= 1.5 and f.max() <= 4.0
    assert [c.seed for c in xbach.batch_cfgs(3)] == [c.seed for c in cfgs[:3]]


def test_synth_is_seeded_and_poisson():
    cfg = xbach.synth_cfg("pair512")
    t1, s1 = xbach.synth_pair(cfg)
    t2, s2 = xbach.synth_pair(cfg)
    assert np.array_equal(t1, t2) and np.array_equal(s1, s2)
    cfg2 = xbach.synth_cfg("pair512", seed=cfg.seed + 1)
    t3, _ = xbach.synth_pair(cfg2)
    assert not np.array_equal(t1, t3)
    # a star-free pair is pure Poisson sky
    blank = xbach.synth_cfg("pair512", nstars=0, sky=150.0)
    t, s = xbach.synth_pair(blank)
    assert np.all(t == np.round(t)) and t.dtype == np.float32
    assert abs(t.mean() - 150.0) < 0.2 and abs(t.var() / 150.0 - 1.0) < 0.03
    low = xbach.synth_cfg("pair512", nstars=0, sky=4.0)  # Knuth branch
    t, _ = xbach.synth_pair(low)
    assert abs(t.mean() - 4.0) < 0.05 and abs(t.var() / 4.0 - 1.0) < 0.05


def test_synth_flux_and_psf():
    cfg = xbach.synth_cfg("pair512", nstars=1, sky=0.0, fluxMin=1e5, fluxMax=1e5, seed=3)
    # Poisson of a pure star: total counts ~ flux, template narrower than science
    t, s = xbach.synth_pair(cfg)
    assert abs(t.sum() - 1e5) < 5 * np.sqrt(1e5)
    assert t.max() > s.max()


def test_fits_round_trip(tmp_path):
    a = (np.arange(35 * 20, dtype=np.float32).reshape(20, 35) - 100.5) * 3.25
    p = tmp_path / "x.fits"
    xbach.write_fits(p, a)
    hdr, b = xbach.read_fits(p)
    assert hdr["BITPIX"] == "-32" and int(hdr["NAXIS1"]) == 35 and int(hdr["NAXIS2"]) == 20
    assert np.array_equal(b.astype(np.float32), a)
    assert p.stat().st_size % 2880 == 0


def _device_code_objects(lib_path, workdir):
    """The gfx950 code objects of the built library: one offload bundle per
    translation unit in its .hip_fatbin section, unbundled with the ROCm tools."""
    import shutil
    import subprocess

    bundler = pathlib.Path("/opt/rocm/llvm/bin/clang-offload-bundler")
    if not bundler.exists() or shutil.which("objcopy") is None:
        pytest.skip("objcopy or the ROCm offload bundler is missing")
    fb = workdir / "fatbin"
    # objcopy rewrites the file it is given: it works on a copy, never on the
    # library this process has loaded
    copy = workdir / "lib.so"
    shutil.copyfile(lib_path, copy)
    subprocess.run(["objcopy", "--dump-section", f".hip_fatbin={fb}", str(copy)], check=True)
    data = fb.read_bytes()
    magic = b"__CLANG_OFFLOAD_BUNDLE__"
    starts = [m.start() for m in re.finditer(re.escape(magic), data)]
    out = []
    for i, a in enumerate(starts):
        b = starts[i + 1] if i + 1 < len(starts) else len(data)
        part = workdir / f"bundle{i}"
        part.write_bytes(data[a:b])
        co = workdir / f"code{i}.co"
        subprocess.run([str(bundler), "--unbundle", "--type=o",
                        "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", f"--input={part}",
                        f"--output={co}"], check=True)
        out.append(co)
    return out


def test_no_kernel_uses_private_scratch(tmp_path):
    """Every kernel of libxbach has a zero private segment: per-queue scratch
    for a kernel with a large private segment, multiplied over many hardware
    queues, once made the runtime abort with OUT_OF_RESOURCES."""
    import subprocess

    readelf = pathlib.Path("/opt/rocm/llvm/bin/llvm-readelf")
    if not readelf.exists():
        pytest.skip("llvm-readelf is missing")
    kernels = {}
    for co in _device_code_objects(xbach.LIB_PATH, tmp_path):
        notes = subprocess.run([str(readelf), "--notes", str(co)], check=True,
                               capture_output=True, text=True).stdout
        # one YAML map per kernel under amdhsa.kernels ("  - " items); its own
        # keys sit at four spaces (argument maps are deeper)
        for block in re.split(r"\n  - ", notes)[1:]:
            name = re.search(r"^    \.name:\s+(\S+)", block, re.M)
            size = re.search(r"^    \.private_segment_fixed_size:\s+(\d+)", block, re.M)
            if name and size:
                kernels[name.group(1)] = int(size.group(1))
    assert len(kernels) >= 30, sorted(kernels)
    assert {k: v for k, v in kernels.items() if v} == {}
