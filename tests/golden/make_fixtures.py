"""Writes the committed golden fixtures from the reference's own data files.

Reads (as data/text, nothing imported or executed):
  res/test0.fits, res/test1.fits           test 1 inputs (science, template)
  res/test0_bg.fits, res/test2.fits        the reference's other held images
                                           (330x330, no goldens: oracle parity)
  tests/test01_conv.fits, test01_sub.fits  HOTPANTS 5.1.11 goldens of test 1
                                           (tools/run_test.py:13 tolerances)
  tools/cond_num.py:3-297                  the dumped 293x293 global normal
                                           matrix `test01_mtx` (6 significant
                                           digits), parsed as text
Writes tests/golden/test01_*.fits.gz and tests/golden/test01_mtx.npz.
Run here (the reference is not on the GPU box): python tests/golden/make_fixtures.py
"""
import gzip
import pathlib
import re
import shutil

import numpy as np

REF = pathlib.Path("/root/reference")
OUT = pathlib.Path(__file__).resolve().parent


def copy_gz(src, dst):
    # a fixture already written is kept byte for byte (gzip headers carry a time)
    if dst.exists():
        return
    with open(src, "rb") as f, open(dst, "wb") as raw, \
            gzip.GzipFile(filename="", mode="wb", compresslevel=9, fileobj=raw, mtime=0) as g:
        shutil.copyfileobj(f, g)


def parse_matrix(text, name):
    start = text.index(f"{name} = np.array([")
    body = text[start:]
    end = body.index("])\n")
    rows = re.findall(r"\[([^\[\]]*)\]", body[: end + 1])
    mat = np.array([[float(v) for v in r.split(",")] for r in rows], dtype=np.float64)
    return mat


def main():
    copy_gz(REF / "res/test0.fits", OUT / "test01_science.fits.gz")
    copy_gz(REF / "res/test1.fits", OUT / "test01_template.fits.gz")
    copy_gz(REF / "tests/test01_conv.fits", OUT / "test01_conv_hotpants.fits.gz")
    copy_gz(REF / "tests/test01_sub.fits", OUT / "test01_sub_hotpants.fits.gz")
    copy_gz(REF / "res/test0_bg.fits", OUT / "test0_bg.fits.gz")
    copy_gz(REF / "res/test2.fits", OUT / "test2.fits.gz")
    text = (REF / "tools/cond_num.py").read_text()
    m = parse_matrix(text, "test01_mtx")
    assert m.shape == (293, 293), m.shape
    np.savez_compressed(OUT / "test01_mtx.npz", mtx=m)
    print("wrote fixtures to", OUT)


if __name__ == "__main__":
    main()
