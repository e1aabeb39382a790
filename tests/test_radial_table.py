"""Pin of the product's tabulated radial projectors (QG_RADIAL_TABLE) on the host, independent
of the oracle's restatement: the library's geometry build (qg_setup.cpp) tabulates one atom's
projector from a table sampling a polynomial in xi = log r, and the radial values recovered
from its per-point tables must reproduce that polynomial wherever the interpolation branch of
RadialGrid2::GetValueByXi (RadialGrid2.h:81-136) is exact for its degree:
  linear (first and last interval): degree <= 1; cubic Hermite next to them: degree <= 2;
  quintic Hermite with 4th-order derivative estimates inside: degree <= 4;
  r < r_min gives table[0] (PseudoPotOperator_mpi_v2.h:154-158); beyond the last point, 0.
The inexact branches must be inexact (so a swapped branch cannot pass).  s and p channels
(the p values are divided by their real Y_1m, RealSphericalHarmonics.h:149-161)."""
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
    exe = os.path.join(d, "radial_table")
    cmd = [HIPCC, "-O1", "-std=c++17", "-ffp-contract=off", "-I" + os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "harness", "radial_table.cpp"),
           os.path.join(ROOT, "qumasun_amd", "csrc", "qg_setup.cpp"), "-o", exe]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:]
    yield exe
    shutil.rmtree(d, ignore_errors=True)


EXACT_UP_TO = {"lin": 1, "cub": 2, "quint": 4, "rmin": 4, "zero": 4}


@pytest.mark.parametrize("l", [0, 1])
@pytest.mark.parametrize("deg", [0, 1, 2, 3, 4])
def test_table_branches_polynomial_exactness(harness, deg, l):
    r = subprocess.run([harness, str(deg), str(l)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout
    m = re.search(r"rmin (\d+) (\S+) lin (\d+) (\S+) cub (\d+) (\S+) quint (\d+) (\S+) zero (\d+) (\S+) outside (\d+)",
                  r.stdout)
    assert m, r.stdout
    res = {k: (int(m[2 * i + 1]), float(m[2 * i + 2])) for i, k in enumerate(["rmin", "lin", "cub", "quint", "zero"])}
    assert int(m[11]) == 0, r.stdout            # no table point outside the cutoff
    for k, (cnt, err) in res.items():
        assert cnt > 10, (k, r.stdout)          # every branch is sampled
        if deg <= EXACT_UP_TO[k]:
            assert err < 1e-13, (k, r.stdout)
        else:
            assert err > 1e-6, (k, r.stdout)    # the branch is not of higher order than it should be
