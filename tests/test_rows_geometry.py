"""Host check of the row kernels' projector geometry (qg_setup.cpp build_rows_geometry): the
per-(tile, plane) items hold exactly the sphere points of the per-image intervals, grouped by
image as the offsets say, and a projection replayed with the projection kernel's bookkeeping
(thread -> image, rank and stride over the image's items, partial slots summed per atom)
equals the direct dot product of the FP64 projector tables with a random psi.  Builds a small
harness from tests/harness/rows_geometry.cpp and qumasun_amd/csrc/qg_setup.cpp with hipcc
(host code only)."""
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
    exe = os.path.join(d, "rows_geometry")
    cmd = [HIPCC, "-O1", "-std=c++17", "-ffp-contract=off", "-I" + os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "harness", "rows_geometry.cpp"),
           os.path.join(ROOT, "qumasun_amd", "csrc", "qg_setup.cpp"), "-o", exe]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:]
    yield exe
    shutil.rmtree(d, ignore_errors=True)


# (nx, ny, nz, tile rows, projection slab, second species with separate s/p cutoffs)
CASES = [(40, 40, 40, 8, 16, 0), (40, 40, 40, 16, 40, 0), (24, 28, 20, 6, 7, 0), (32, 24, 36, 8, 5, 1),
         (40, 40, 40, 4, 3, 1), (16, 16, 16, 16, 16, 0)]


@pytest.mark.parametrize("order", ["image", "group"])
@pytest.mark.parametrize("case", CASES)
def test_rows_items_and_projection_bookkeeping(harness, case, order):
    """Projection order (by image, with the thread map and partial slots) and the apply
    kernel's order (items cut at 4-point groups, sorted by group, with the group runs)."""
    r = subprocess.run([harness] + [str(v) for v in case] + [str(int(order == "group"))], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout
    m = re.search(r"items (\d+) refpts (\d+) missing (\d+) extra (\d+) dup (\d+) badgroup (\d+) badmap (\d+) "
                  r"maxrel (\S+) slots (\d+) maxcol (\d+) sp4 (\d) njs (\d)", r.stdout)
    assert m, r.stdout
    items, refpts, missing, extra, dup, badgroup, badmap = (int(m[i]) for i in range(1, 8))
    maxrel = float(m[8])
    assert items == refpts and missing == 0 and extra == 0 and dup == 0, r.stdout
    assert badgroup == 0 and badmap == 0, r.stdout
    assert maxrel < 1e-13, r.stdout
    assert int(m[11]) == (0 if case[5] else 1), r.stdout   # s+p fast form only for the one-species cell
