"""The shared-divisor quotient of stages 4 and 5 (bmfr_amd/csrc/device_common.h:
quotient_reciprocal, quotient): a / d as (float)((double)a * y) with y an
FP64 reciprocal refined by two Newton steps.  The header argues that this is
the IEEE quotient's bits for every pair of floats; tests/native/quotient_check.c
restates the arithmetic on the host, with the device's first approximation
replaced by 1/d perturbed by up to 2^-14 (far worse than v_rcp_f64), and
compares it with IEEE float division on 4e8 pairs: random bit patterns,
numerators chosen so that the quotient sits next to a float rounding boundary
(where a wrong rounding would show first), integer divisors (1/spp) and
special values (zeros, infinities, NaN, subnormals).  The GPU parity tests
compare the stages' outputs with the oracle's IEEE divisions bit for bit."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_shared_divisor_quotient_is_ieee(tmp_path):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    exe = str(tmp_path / "quotient_check")
    subprocess.run([gcc, "-O2", "-ffp-contract=off", os.path.join(ROOT, "tests", "native", "quotient_check.c"),
                    "-o", exe, "-lm"], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=600).stdout
    assert "bad 0" in out, out
    assert int(out.split("checked ")[1].split()[0]) > 390_000_000, out
