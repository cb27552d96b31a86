"""The autotuner's search and its knob table (CPU).

* coordinate descent finds the minimum of a separable bowl over the knobs'
  candidate values, measuring each distinct setting once, and stops when a
  pass improves nothing;
* the knob identifiers match the QBX_TUNE_* enumeration of the C header and
  every candidate value lies in the range the library accepts (a power of
  two within the bounds the header states).
"""
import os
import re

import pytest

from pytential_amd import autotune

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_coordinate_descent_finds_bowl_minimum():
    knobs = {"a": (1, 2, 4, 8, 16), "b": (32, 64, 128, 256), "c": (1, 2)}
    target = {"a": 4, "b": 256, "c": 1}
    calls = []

    def measure(s):
        calls.append(tuple(sorted(s.items())))
        return sum((knobs[n].index(s[n]) - knobs[n].index(target[n])) ** 2 for n in s) + 1.0

    best, t, trials = autotune.coordinate_descent(knobs, {"a": 1, "b": 32, "c": 2}, measure,
                                                  passes=3)
    assert best == target and t == 1.0
    assert len(calls) == len(set(calls)) == len(trials)


def test_knob_table_matches_header():
    hdr = open(os.path.join(ROOT, "include", "qbx_mi355x.h")).read()
    enum = dict((name.lower(), int(v)) for name, v in
                re.findall(r"QBX_TUNE_([A-Z0-9_]+) = (\d+)", hdr) if name != "COUNT")
    names = {n: pid for n, (pid, _) in autotune.KNOBS.items()}
    names.update(autotune.SETTINGS)
    assert names == enum
    bounds = dict((m[0].lower(), (int(m[1]), int(m[2]))) for m in re.findall(
        r"QBX_TUNE_([A-Z0-9_]+) = \d+,\s*/\* [^(]*\((\d+)\.\.(\d+)\)", hdr))
    for n, (_, cands) in autotune.KNOBS.items():
        for v in cands:
            assert v & (v - 1) == 0, (n, v)
            if n in bounds:
                lo, hi = bounds[n]
                assert lo <= v <= hi, (n, v)


@pytest.mark.parametrize("bad", [0, 3, 2048])
def test_coordinate_descent_never_measures_twice(bad):
    seen = []

    def measure(s):
        seen.append(tuple(sorted(s.items())))
        return float(s["x"] == bad)

    autotune.coordinate_descent({"x": (1, 2, 4)}, {"x": 2}, measure, passes=4)
    assert len(seen) == len(set(seen))
