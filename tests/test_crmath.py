"""Correctly rounded sin / tan / exp / pow of the FP64 path (geoflac_amd/csrc/gf_ddmath.h).

The host build of the same code (in oracle/liboracle_cr.so) is checked against
60-digit decimal evaluations, its fast paths (table-driven, with a rounding test)
against its slow path over millions of arguments; the device build (fast paths) is
checked bit for bit against the host's slow path, and against the host libm that the
reference calls.
"""
import ctypes as C
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
    fast = np.empty(len(x))
    ns = C.c_long(0)
    L.gfo_math_fast(fn, np.ascontiguousarray(x).ctypes.data, None if y is None else np.ascontiguousarray(y).ctypes.data,
                    fast.ctypes.data, len(x), C.byref(ns))
    for k in range(len(x)):
        yy = 0.0 if y is None else float(y[k])
        ex = _exact(fn, float(x[k]), yy)
        assert L.gfo_math(fn, float(x[k]), yy) == ex, (fn, x[k], yy)
        assert fast[k] == ex, (fn, x[k], yy)


def _fast_vs_slow(fn, x, y=None):
    L = ol.lib(cr=True)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = None if y is None else np.ascontiguousarray(y, dtype=np.float64)
    out = np.empty_like(x)
    nslow = C.c_long(0)
    bad = L.gfo_math_fast(fn, x.ctypes.data, None if y is None else y.ctypes.data, out.ctypes.data, len(x),
                          C.byref(nslow))
    return bad, nslow.value


@pytest.mark.parametrize("fn,lo,hi", [(0, 0.0, 1.2), (0, 1.2, 1.6), (1, 0.0, 1.2), (1, 1.2, 1.5707963267948966),
                                       (2, -700.0, 700.0), (2, -1.0, 1.0), (3, None, None)])
def test_fast_paths_equal_slow_path(fn, lo, hi):
    """Whatever the fast path returns is the slow path's value, bit for bit, and the
    rounding test refuses few arguments (where the fast value lies too close to a
    rounding boundary for its error bound; exp near the ends of its range, where the
    result leaves [2^-1000, 2^1000], always goes to the slow path)."""
    rng = np.random.default_rng(3 + fn)
    n = 1_000_000
    if fn == 3:
        x = 10 ** rng.uniform(-30, 30, n)
        y = rng.uniform(-3, 3, n)
    else:
        x, y = rng.uniform(lo, hi, n), None
    bad, nslow = _fast_vs_slow(fn, x, y)
    assert bad == 0
    inside = fn == 3 or (fn in (0, 1) and hi <= 1.2) or (fn == 2 and hi < 600)
    if inside:  # arguments the fast path covers
        assert nslow < (1e-3 if fn == 3 else 1e-4) * n, nslow
    edges = np.array([0.0, -0.0, 1e-300, 5e-324, 1.2, 1.2000000000000002, 1.5707963267948966, 2.0, -1.0,
                      699.9, -699.9, 700.0, np.nan, np.inf, -np.inf])
    assert _fast_vs_slow(fn, edges, None if fn != 3 else np.full(len(edges), 0.7))[0] == 0


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
