"""Host check of the projector data the kernels stage: the row segments (in-kernel projector
values) cover exactly the sphere points of the item tables, and the kernel's FP32 formula
(displacement from the integer part of the image centre, r^2 < 1e-14 -> Y_1m = 0) reproduces
the tables' FP64 projector values to FP32 rounding, nuclei on grid points included.  Builds a
small harness from tests/harness/tile_segments.cpp and the library's qumasun_amd/csrc/qg_setup.cpp
with hipcc (host code only)."""
import os
import re
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def harness():
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    d = tempfile.mkdtemp()
    exe = os.path.join(d, "tile_segments")
    cmd = [HIPCC, "-O1", "-std=c++17", "-ffp-contract=off", "-I" + os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "harness", "tile_segments.cpp"),
           os.path.join(ROOT, "qumasun_amd", "csrc", "qg_setup.cpp"), "-o", exe]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:]
    yield exe
    shutil.rmtree(d, ignore_errors=True)


@pytest.mark.parametrize("n,ty,tz", [(40, 16, 2), (40, 4, 2), (20, 20, 2), (40, 6, 10), (24, 5, 3)])
def test_segments_match_tables(harness, n, ty, tz):
    r = subprocess.run([harness, str(n), str(ty), str(tz)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout
    m = re.search(r"items (\d+) segs (\d+) max\|diff\| (\S+) missing (\d+) extra (\d+)", r.stdout)
    assert m, r.stdout
    items, segs, diff, missing, extra = int(m[1]), int(m[2]), float(m[3]), int(m[4]), int(m[5])
    assert items == segs and missing == 0 and extra == 0, r.stdout
    assert diff < 5e-7, r.stdout   # FP32 evaluation against FP64 tables of values up to ~1.6
