"""Correctly rounded sin / tan / exp / pow of the FP64 path (geoflac_amd/csrc/gf_ddmath.h).

The host build of the same code (in oracle/liboracle_cr.so) is checked against
60-digit decimal evaluations; the device build is checked bit for bit against the
host build, and against the host libm that the reference calls.
"""
import math
from decimal import Decimal, getcontext

import numpy as np
import pytest

from tests import oracle_lib as ol

getcontext().prec = 60


def _dsin(x):
    x = Decimal(x)
    s, t, n = Decimal(0), x, 1
    while abs(t) > Decimal(10) ** -75:
        s += t
        n += 2
        t = -t * x * x / ((n - 1) * n)
    return s


def _dcos(x):
    x = Decimal(x)
    s, t, n = Decimal(0), Decimal(1), 0
    while abs(t) > Decimal(10) ** -75:
        s += t
        n += 2
        t = -t * x * x / ((n - 1) * n)
    return s


def _nearest(v):
    f = float(v)
    cands = [f, math.nextafter(f, math.inf), math.nextafter(f, -math.inf)]
    return min(cands, key=lambda c: abs(Decimal(c) - v))


def _args(seed, n):
    rng = np.random.default_rng(seed)
    return {
        0: (rng.uniform(0, math.pi / 2, n), None),
        1: (rng.uniform(0, math.pi / 2 - 1e-3, n), None),
        2: (rng.uniform(-40, 40, n), None),
        3: (10 ** rng.uniform(-18, -10, n), rng.uniform(-1, 1, n)),
    }


def _exact(fn, x, y):
    if fn == 0:
        return _nearest(_dsin(x))
    if fn == 1:
        return _nearest(_dsin(x) / _dcos(x))
    if fn == 2:
        return _nearest(Decimal(x).exp())
    return _nearest((Decimal(y) * Decimal(x).ln()).exp())


@pytest.mark.parametrize("fn", [0, 1, 2, 3])
def test_host_build_is_correctly_rounded(fn):
    L = ol.lib(cr=True)
    x, y = _args(7, 400)[fn]
    for k in range(len(x)):
        yy = 0.0 if y is None else float(y[k])
        assert L.gfo_math(fn, float(x[k]), yy) == _exact(fn, float(x[k]), yy), (fn, x[k], yy)


@pytest.mark.gpu
@pytest.mark.parametrize("fn", [0, 1, 2, 3])
def test_device_equals_host_build_and_mostly_libm(fn):
    from geoflac_amd import gpu
    print("case: device correctly rounded math, fn %d" % fn, flush=True)
    x, y = _args(11, 200000)[fn]
    dev = gpu.math_check(fn, x, y)
    L = ol.lib(cr=True)
    Lg = ol.lib(cr=False)
    host = np.array([L.gfo_math(fn, float(a), 0.0 if y is None else float(b))
                     for a, b in zip(x, x if y is None else y)])
    libm = np.array([Lg.gfo_math(fn, float(a), 0.0 if y is None else float(b))
                     for a, b in zip(x, x if y is None else y)])
    assert np.array_equal(dev, host)
    # glibc is not correctly rounded for a small fraction of arguments
    assert np.mean(dev != libm) < 5e-3
