"""Build recipe of the MI355X engine: compiles qumasun_amd/csrc for gfx950 with hipcc and
links qumasun_amd/lib/libqumasun_gpu.so (in-tree, so it travels with the repository)."""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
BUILD = os.path.join(HERE, "_build")
LIBNAME = "libqumasun_gpu.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-variable", "-Wno-unused-function",
          "-Wno-unused-but-set-variable"]
# no hand-written or compiler-generated packed FP32 math (v_pk_*_f32)
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _sources():
    hip = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    cpp = sorted(f for f in os.listdir(CSRC) if f.endswith(".cpp"))
    return hip, cpp


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + out)
    return out


def build(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(BUILD, exist_ok=True)
    hip, cpp = _sources()
    jobs = []
    objs = []
    for f in hip:
        o = os.path.join(BUILD, f + ".o")
        objs.append(o)
        jobs.append([HIPCC, *COMMON, *DEVICE, "-c", os.path.join(CSRC, f), "-o", o])
    for f in cpp:
        o = os.path.join(BUILD, f + ".o")
        objs.append(o)
        # host-only setup code: no contraction, so sphere membership uses exact products
        jobs.append([HIPCC, *COMMON, "-ffp-contract=off", "-c", os.path.join(CSRC, f), "-o", o])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    lib = os.path.join(LIBDIR, LIBNAME)
    tmp = lib + ".tmp"
    _run([HIPCC, *DEVICE, "-shared", "-fPIC", "-o", tmp, *objs,
          "-L/opt/rocm/lib", "-lrocblas", "-lrocsolver", "-lrocfft", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, lib)
    return lib


HARNESS = os.path.join(os.path.dirname(HERE), "tests", "harness")
COMPAT_DRIVER = os.path.join(LIBDIR, "compat_drive")


def compat_driver(rebuild=False):
    """Build (when missing or asked) the test driver of the reference-shaped C++ surface,
    tests/harness/compat_drive.cpp, linked against the library; returns its path."""
    if rebuild or not os.path.exists(COMPAT_DRIVER):
        inc = os.path.join(os.path.dirname(HERE), "include")
        tmp = COMPAT_DRIVER + ".tmp"
        _run([HIPCC, "-O2", "-std=c++17", "-I" + inc, os.path.join(HARNESS, "compat_drive.cpp"), "-o", tmp,
              "-L" + LIBDIR, "-l:" + LIBNAME, "-Wl,-rpath," + LIBDIR])
        os.replace(tmp, COMPAT_DRIVER)
    return COMPAT_DRIVER


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
