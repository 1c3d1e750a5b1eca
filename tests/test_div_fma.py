"""The FMA-corrected division that the thermal kernels use where a divisor repeats
(div_fma: Markstein's theorem) gives the correctly rounded quotient: checked against
IEEE division on 10 million random operand pairs (CPU, hardware FMA)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_fma_division_is_correctly_rounded(tmp_path):
    exe = str(tmp_path / "div_fma_check")
    try:
        subprocess.run(["gcc", "-O2", "-ffp-contract=off", os.path.join(HERE, "csrc", "div_fma_check.c"),
                        "-o", exe, "-lm"], check=True, capture_output=True)
    except (OSError, subprocess.CalledProcessError) as e:  # no C compiler
        pytest.skip("gcc unavailable: %s" % e)
    out = subprocess.run([exe, "10000000"], check=True, capture_output=True, text=True).stdout
    assert int(out.strip()) == 0
