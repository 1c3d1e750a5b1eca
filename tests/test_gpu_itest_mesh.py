"""The remeshing test itest_mesh (src/rem_test.f90:9-139) on the device
(gf_itest_mesh) against the oracle's restatement: the decision and the reason's value
exactly, the smallest triangle angle to 1e-13 relative (the device's acos against the
host's; the oracle is pinned by closed forms, tests/test_oracle.py, and the drop-in's
remeshing count equals the reference's, tests/test_fortran_dropin.py).  Cases: the
deformed remesh60x20 states after 100 and 300 reference steps, the undeformed box with
angle_rem on either side of its smallest angle, a shortened box (mode_rem 11) and a
surface change beyond a third of the top element height."""
import math
import sys

import pytest

from geoflac_amd import state
from tests import oracle_lib as ol


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _both(p, st, *args):
    from geoflac_amd import gpu
    with gpu.Model(p, st) as m:
        dev = m.itest_mesh(*args)
    ref = ol.itest_mesh(p, st, *args)
    _say("  args %s: device %s, oracle %s" % (args, dev, ref))
    return dev, ref


def _agree(dev, ref):
    assert dev[0] == ref[0] and dev[2] == ref[2]
    assert abs(dev[1] - ref[1]) <= 1e-13 * abs(ref[1])


@pytest.mark.gpu
@pytest.mark.parametrize("step", [100, 300])
def test_itest_mesh_on_deformed_states_equals_oracle(step):
    _say("case: remesh60x20 after %d reference steps" % step)
    p = ol.golden_params("remesh60x20")
    st = ol.golden_full("remesh60x20", step)
    dev, ref = _both(p, st, 1, 0.5, 0.0)
    _agree(dev, ref)
    amin = ref[1]
    for ar in (amin * (1 - 1e-9), amin * (1 + 1e-9)):  # either side of the smallest angle
        _agree(*_both(p, st, 1, 0.5, ar))


@pytest.mark.gpu
def test_itest_mesh_reasons_equal_oracle():
    p = ol.golden_params("box60x20")
    st = ol.golden_full("box60x20", 0)
    nz, nx = p.nz, p.nx
    dx = abs(st.cord[0, 1, 0] - st.cord[0, 0, 0])
    dz = abs(st.cord[0, 0, 1] - st.cord[1, 0, 1])
    want = math.degrees(math.atan(min(dx, dz) / max(dx, dz)))
    _say("case: undeformed box60x20, smallest angle %.15g" % want)
    dev, ref = _both(p, st, 1, 0.5, want * (1 + 1e-9))
    _agree(dev, ref)
    assert dev[0] == 1
    sh = st.copy()
    sh.cord[:, nx - 1, 0] -= 3 * dx
    _say("case: right wall moved in by 3 element widths, mode_rem 11")
    dev, ref = _both(p, sh, 11, 2.0, 0.0)
    _agree(dev, ref)
    assert dev[0] == 2
    p2 = ol.golden_params("topo60x20")
    st2 = ol.golden_full("topo60x20", 0, marker_room=True)
    dz2 = abs(st2.cord[0, 0, 1] - st2.cord[1, 0, 1])
    st2.dhacc[:] = 0.0
    st2.dhacc[3] = 0.34 * dz2
    _say("case: topo60x20 with a surface change of 0.34 of the top element height")
    dev, ref = _both(p2, st2, 1, 0.5, 0.0)
    _agree(dev, ref)
    assert dev[0] == 3
