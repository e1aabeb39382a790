"""Host check of the split-BF16 Gram kernel's tiling (qg_gram_map.hpp, the maps the kernel
uses): every staged global value lies inside its operand and K chunk, coverage is exact (each
value of a tile's rows and chunk staged once), LDS slots stay inside the tile, and the partial
tiles stay inside the workspace, each written once.  Shapes: the GPU tests' small products
(unaligned K and ragged tiles) and the solver's Rayleigh-Ritz products, whose N x N, N x 2N and
N x 3N calls share one workspace sized for N x 3N."""
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
    exe = os.path.join(d, "gram_bounds")
    r = subprocess.run([HIPCC, "-O2", "-std=c++17", os.path.join(ROOT, "tests", "harness", "gram_bounds.cpp"), "-o",
                        exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:]
    yield exe
    shutil.rmtree(d, ignore_errors=True)


def _ws(K, M, N):
    # the workspace a caller allocates for one shape: its unbounded chunk count x M x N
    tiles = -(-M // 256) * -(-N // 256)
    nch = max(1, min(max(1, 512 // tiles), max(1, K // 4096)))
    return nch * M * N


def run(harness, K, M, N, lda, ldb, cap=0, cover=True):
    r = subprocess.run([harness, str(K), str(M), str(N), str(lda), str(ldb), str(cap), str(int(cover))],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
    assert r.returncode == 0, r.stdout
    m = re.search(r"chunks (\d+) kchunk (\d+) blocks (\d+) reads (\d+) bad_read (\d+) bad_lds (\d+) bad_cover (\d+) "
                  r"bad_write (\d+) dup_write (\d+) unwritten (\d+)", r.stdout)
    assert m, r.stdout
    vals = [int(m[i]) for i in range(5, 11)]
    assert vals == [0] * 6, r.stdout
    return int(m[1])


@pytest.mark.parametrize("K,M,N", [(10007, 40, 72), (64000, 16, 16), (4096, 128, 130), (70001, 300, 520),
                                   (4097, 1, 1)])
def test_small_products(harness, K, M, N):
    run(harness, K, M, N, K, K)
    run(harness, K, M, N, K + 3, K + 5)   # leading dimensions larger than K


@pytest.mark.parametrize("K,N", [(120 ** 3, 512), (40 ** 3, 16), (80 ** 3, 128)])
def test_solver_products_share_one_workspace(harness, K, N):
    cap = _ws(K, N, 3 * N)
    for nb in (N, 2 * N, 3 * N):
        nch = run(harness, K, N, nb, K, K, cap=cap, cover=K < 10 ** 6)
        assert nch * N * nb <= cap
