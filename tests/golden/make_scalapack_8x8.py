"""Extracts the embedded 8x8 symmetric (A, B) pair of the reference's test_scalapack.cpp
(source/test_scalapack.cpp:37-55, a dumped LOBPCG Sa/Sb pair; the test holds no expected
output) into scalapack_8x8.json.  Input data only."""
import json
import os
import re

SRC = "/root/reference/source/test_scalapack.cpp"
text = open(SRC).read()
out = {}
for name in ("A", "B"):
    m = re.search(r"double\s+%s\[\]\s*=\s*\{(.*?)\};" % name, text, re.S)
    vals = [float(v) for v in re.findall(r"-?\d+\.\d+", m.group(1))]
    assert len(vals) == 64, (name, len(vals))
    out[name] = vals
out["source"] = "source/test_scalapack.cpp:37-55 (row-major 8x8, symmetric)"
json.dump(out, open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "scalapack_8x8.json"), "w"), indent=1)
print("ok")
