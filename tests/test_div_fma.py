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


def test_min_of_reciprocals_is_reciprocal_of_max():
    """dt_mass's side measure: the device reduces the largest product y = area x side
    and takes one reciprocal, where the reference takes the minimum of 1/y over every
    product.  RN(1/y) does not increase with y > 0, so the two agree bit for bit:
    checked on groups of 12 products (an element) and on whole arrays spanning many
    binades, including near-ties (neighbouring doubles)."""
    import numpy as np
    rng = np.random.default_rng(7)
    y = np.exp(rng.uniform(-200.0, 200.0, size=(200000, 12)))
    y[::3, 5] = np.nextafter(y[::3, 4], np.inf)  # neighbouring doubles
    y[1::3, 7] = y[1::3, 2]                       # equal products
    one = np.float64(1.0)
    assert np.array_equal(np.min(one / y, axis=1), one / np.max(y, axis=1))
    assert np.min(one / y) == one / np.max(y)
    assert np.array_equal(np.minimum(1e28, np.min(one / y, axis=1)), np.minimum(1e28, one / np.max(y, axis=1)))
