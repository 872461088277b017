"""The count draw's division-free step quotient equals the IEEE quotient, exhaustively.

The GPU's inversion loops (`step_quot` in csrc/lp_kernels.hip) replace the correctly
rounded t / s of the oracle's Poisson and negative-binomial inversions
(oracle/lp_oracle.cpp, count draws after model.py:1327-1363) by t * RN(1/s) refined with
its fma residual.  Bit-exact parity of the count draws rests on that identity for
s = 1 .. 64; it is checked here over every float significand, and two planted mistakes
(no residual correction; the reciprocal of the wrong step) must be caught.
"""

import os
import shutil
import subprocess

import pytest

SRC = os.path.join(os.path.dirname(__file__), "step_quotient_check.c")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not found")
    exe = str(tmp_path_factory.mktemp("quot") / "step_quotient_check")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, SRC, "-lm"], check=True)
    return exe


def test_step_quotient_equals_ieee_division_for_every_significand(checker):
    out = subprocess.run([checker, "64"], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout
    assert out.stdout.strip() == "mismatches 0"


@pytest.mark.parametrize("mistake", ["nocorr", "lane"])
def test_step_quotient_check_catches_planted_mistakes(checker, mistake):
    out = subprocess.run([checker, "8", mistake], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout
    assert int(out.stdout.split()[-1]) > 0
