"""Mixed precision.

GF_PREC_MIXED_STORAGE: velocity, force and the stress deviation stored in FP32, all
arithmetic FP64.  Bar: rel-L2 of velocity and stress against the FP64 path (itself
bitwise the correctly rounded oracle) <= 1e-5 after 1000 steps of the reference's
60x20 box.

GF_PREC_MIXED: the same storage with FP32 strain-rate, stress and force arithmetic.
The plastic/viscous branch choice of every element (src/fl_rheol.f90:266-288) is a
discontinuous function of the stress, so the run is chaotic: FP32 arithmetic flips
near-tie elements and the fields drift apart (about 1e-2 in velocity after 100
steps).  Its test holds the fields to bounds that catch a broken kernel (an unstable
or wrong update) and prints the per-field errors; temperature and coordinates, which
stay FP64, are held tight.
"""
import sys

import numpy as np
import pytest

from geoflac_amd import abi, state
from tests import oracle_lib as ol

CASE = "box60x20"
FIELDS = ["vel", "stress0", "temp", "cord", "aps", "strain", "visn", "e2sr"]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _run(precision, nsteps):
    from geoflac_amd import gpu
    p = ol.golden_params(CASE)
    st = ol.golden_state(CASE, 0)
    with gpu.Model(p, st, precision=precision) as m:
        m.step(nsteps)
        m.download()
    return st


def test_mixed_precision_is_accepted_by_the_abi_header():
    hdr = open(__file__.replace("tests/test_gpu_mixed.py", "include/gf_abi.h")).read()
    assert "GF_PREC_MIXED 1" in hdr


def _errors(nsteps, precision):
    ref = _run(abi.PREC_FP64, nsteps)
    mix = _run(precision, nsteps)
    err = {}
    for f in FIELDS:
        a, b = getattr(mix, f), getattr(ref, f)
        err[f] = (state.rel_l2(a, b), float(np.max(np.abs(a - b))))
    _say("  " + ", ".join("%s rel-L2 %.2e max %.2e" % (f, e[0], e[1]) for f, e in err.items()))
    assert mix.clock.nloop == ref.clock.nloop
    return err, mix, ref


@pytest.mark.gpu
@pytest.mark.parametrize("nsteps", [100, 1000])
def test_mixed_storage_matches_fp64(nsteps):
    _say("case: %d steps of %s, FP32 storage / FP64 arithmetic vs FP64" % (nsteps, CASE))
    err, mix, ref = _errors(nsteps, abi.PREC_MIXED_STORAGE)
    assert err["vel"][0] <= 1e-5, err["vel"]
    assert err["stress0"][0] <= 1e-5, err["stress0"]
    assert err["temp"][0] <= 1e-9, err["temp"]
    assert abs(mix.clock.dt - ref.clock.dt) <= 1e-6 * ref.clock.dt


@pytest.mark.gpu
@pytest.mark.parametrize("nsteps", [100, 1000])
def test_mixed_fp32_arithmetic_stays_bounded(nsteps):
    _say("case: %d steps of %s, FP32 arithmetic vs FP64" % (nsteps, CASE))
    err, mix, ref = _errors(nsteps, abi.PREC_MIXED)
    # measured: vel 7e-3 / 2.2e-2, stress 7.7e-5 / 3.3e-4, temp 4.7e-10 / 2.6e-8,
    # cord 1.3e-7 / 4.0e-6 after 100 / 1000 steps
    assert err["vel"][0] <= 0.1, err["vel"]
    assert err["stress0"][0] <= 3e-3, err["stress0"]
    assert err["temp"][0] <= 1e-6, err["temp"]
    assert err["cord"][0] <= 1e-4, err["cord"]
    assert np.isfinite(mix.vel).all() and np.isfinite(mix.stress0).all()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [abi.PREC_MIXED, abi.PREC_MIXED_STORAGE])
def test_mixed_round_trip_of_fp32_fields(precision):
    """Upload/download of the FP32 planes: a state downloaded right after upload holds
    the FP32 roundings of the host values and nothing else changes."""
    from geoflac_amd import gpu
    _say("case: mixed-precision upload/download round trip of %s" % CASE)
    p = ol.golden_params(CASE)
    st = ol.golden_state(CASE, 100)
    orig = st.copy()
    with gpu.Model(p, st, precision=precision) as m:
        m.download()
    for f in state.State.FIELDS:
        a, b = getattr(st, f), getattr(orig, f)
        if f in ("vel", "force"):
            assert np.array_equal(a, b.astype(np.float32).astype(np.float64)), f
        elif f == "stress0":
            # kept as an FP32 deviation from the element's mean normal stress (FP64):
            # the round trip is exact to half an FP32 ulp of that deviation
            pref = -(b[:, :, 0, :] + b[:, :, 1, :] + b[:, :, 3, :]).sum(axis=2) / 12.0
            dev = b.copy()
            for c in (0, 1, 3):
                dev[:, :, c, :] += pref[:, :, None]
            bound = np.abs(dev) * 2.0 ** -24 + 1e-30
            assert np.all(np.abs(a - b) <= bound * 1.0001 + np.abs(pref)[:, :, None, None] * 2.0 ** -52), f
        else:
            assert np.array_equal(a, b), f
