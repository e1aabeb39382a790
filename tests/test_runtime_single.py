"""The library and torch share one HIP runtime in a process (CPU check, no device calls).

torch's wheel bundles libamdhip64.so / libhsa-runtime64.so under the same sonames as /opt/rocm's
copies.  qumasun_amd's loader imports torch before it maps the library, so the library's
libamdhip64.so.7 and libhsa-runtime64.so.1 resolve to torch's copies; loading the library
first maps a second HIP and HSA runtime when torch is imported afterwards, and torch's first
device call then fails.  Each case runs in a fresh interpreter.
"""
import subprocess
import sys
import textwrap

import pytest

import qumasun_amd._native as N

_PROBE = textwrap.dedent("""
    import sys
    sys.path.insert(0, {root!r})
    {prelude}
    import torch  # noqa: F401
    maps = open("/proc/self/maps").read()
    for lib in ("libamdhip64", "libhsa-runtime64"):
        print(len({{l.split()[-1] for l in maps.splitlines() if lib in l}}))
""")


def _mapped(prelude):
    import os
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    out = subprocess.run([sys.executable, "-c", _PROBE.format(root=root, prelude=prelude)],
                         stdout=subprocess.PIPE, text=True, check=True).stdout.split()
    return [int(x) for x in out]


@pytest.mark.skipif(not __import__("os").path.exists(N.LIB_PATH), reason="library not built")
def test_loader_maps_one_runtime():
    assert _mapped("import qumasun_amd._native as N; N.lib()") == [1, 1]
