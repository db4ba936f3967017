import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "thesis-emx-bach_amd"))
sys.path.insert(0, str(ROOT / "tests"))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")


@pytest.fixture(scope="session")
def test01_pair():
    """Test 1 of the reference (tools/run_test.py:13): science res/test0.fits,
    template res/test1.fits, with the HOTPANTS 5.1.11 goldens."""
    import numpy as np
    import xbach

    _, sci = xbach.read_fits(GOLDEN / "test01_science.fits.gz")
    _, tmpl = xbach.read_fits(GOLDEN / "test01_template.fits.gz")
    _, gconv = xbach.read_fits(GOLDEN / "test01_conv_hotpants.fits.gz")
    _, gsub = xbach.read_fits(GOLDEN / "test01_sub_hotpants.fits.gz")
    return tmpl.astype(np.float32), sci.astype(np.float32), gconv, gsub
