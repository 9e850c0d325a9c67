"""Stage 4's tone curve, clamp(pow(max(0, v), 0.454545f), 0, 1)
(opencl/bmfr.cl:851-856), as the library computes it
(bmfr_amd/csrc/tone_curve.h, a table-driven FP64 evaluation of our
own).

The header is compiled for the host with g++ (contraction off); it uses only
IEEE-exact operations, so the host build computes the device build's bits.
tests/native/tone_curve_check.cpp runs it on every float in [0, 1] (all
inputs above 1 clamp to 1) and on the special values, against
  - pow in double precision rounded once to float (correctly rounded up to
    that double rounding): at most 1 ulp, and 1 ulp only where the exact
    power sits within a few 1e-16 of a rounding boundary (1 of the 1.07e9
    inputs);
  - the C library's powf (the oracle's tone map): at most 1 ulp.
A dropped polynomial term, a wrong coefficient, a wrong table entry or a
wrong reduction moves hundreds to millions of results, so the mismatch bound
pins the method.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def check_output(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("tone") / "tone_curve_check")
    subprocess.run([gxx, "-O2", "-std=c++17", "-ffp-contract=off", "-fopenmp",
                    "-I", os.path.join(ROOT, "bmfr_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "tone_curve_check.cpp"), "-o", exe],
                   check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True, timeout=600).stdout
    return {k: int(v) for k, v in re.findall(r"(\w+) (-?\d+)", out)}


def test_every_float_in_unit_interval(check_output):
    r = check_output
    assert r["checked"] == 0x3F800000 + 1
    assert r["max_ulp_vs_double_pow"] <= 1
    assert r["mismatches_vs_double_pow"] <= 100
    assert r["max_ulp_vs_powf"] <= 1


def test_special_values(check_output):
    assert check_output["special_bad"] == 0


def test_tables_are_the_generator_output():
    """tone_tables.h is exactly what gen_tone_tables.py writes (every entry
    the correctly rounded double of its definition, computed at 60 digits)."""
    import sys
    csrc = os.path.join(ROOT, "bmfr_amd", "csrc")
    out = subprocess.run([sys.executable, os.path.join(csrc, "gen_tone_tables.py")], check=True,
                         capture_output=True, text=True, timeout=120).stdout
    assert out == open(os.path.join(csrc, "tone_tables.h")).read()
