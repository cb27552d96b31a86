"""Writes tsqbx_bessel.json: the spherical Bessel / Hankel golden vectors held
in the reference's test_spherical_bessel_functions and
test_spherical_hankel_functions (test/test_target_specific_qbx.py:47-128),
read as text from the reference tree (nothing of it is imported or run).

    python tests/golden/make_tsqbx_bessel.py [/root/reference]
"""
import json
import os
import re
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
src = open(os.path.join(ref, "test", "test_target_specific_qbx.py")).read()

num = r"[+-]?\d\.\d+e[+-]\d+"
cnum = re.compile(rf"({num})\s*\+\s*({num})j")


def arrays(name):
    m = re.search(rf"{name}\s*=\s*np\.array\(\[(.*?)\]\)", src, re.S)
    return [[float(a), float(b)] for a, b in cnum.findall(m.group(1))]


out = {
    "source": "test/test_target_specific_qbx.py:47-128 (values from scipy.special)",
    "jn": {"nterms": 9, "z": [0.0, 3.0], "j": arrays("j_expected"),
           "jder": arrays("jder_expected")},
    "hn": {"nterms": 9, "z": [2.0, 3.0], "h": arrays("h_expected"),
           "hder": arrays("hder_expected")},
}
for k in ("j", "jder"):
    assert len(out["jn"][k]) == 10
for k in ("h", "hder"):
    assert len(out["hn"][k]) == 10
with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "tsqbx_bessel.json"),
          "w") as f:
    json.dump(out, f, indent=1)
