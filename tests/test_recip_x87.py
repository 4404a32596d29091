"""The device's copy of niftilib's long-double reciprocal (niftyreg_amd/csrc/recip_x87.hpp).

niftilib inverts matrices with `deti = 1.0l / deti` (x87 long double, then double); the device
reproduces that double rounding in double arithmetic.  The header is compiled for the host with
g++ and compared against the real long-double quotient on random doubles of many binades: it
must agree everywhere, while the plain double reciprocal differs in about one case in 4000.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "niftyreg_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ absent")
def test_recip_x87_matches_long_double(tmp_path):
    exe = str(tmp_path / "recip_check")
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-I" + CSRC, "-o", exe,
                    os.path.join(HERE, "native", "recip_x87_check.cpp")], check=True)
    out = subprocess.run([exe, "2000000"], check=True, capture_output=True, text=True).stdout.split()
    checked, bad, plain = (int(v) for v in out)
    assert checked == 2000000 and bad == 0
    assert plain > 100  # the double rounding matters: the check is not vacuous
