"""Surface processes on the device (diff_topo with resurface, src/fl_move.f90:138-472):
topo60x20, topography diffusion over a weakened block, where resurface appends
sediment markers and removes eroded ones from step ~5000 on.  The device runs are
compared bit for bit with the correctly rounded oracle from the same reference state
(the oracle itself reproduces the reference's states bit for bit,
tests/test_oracle.py::test_surface_processes_bitwise_equal_to_reference)."""
import sys

import numpy as np
import pytest

from geoflac_amd import state
from tests import oracle_lib as ol

CASE = "topo60x20"
MARKER_ARRAYS = ("mark_a1", "mark_a2", "mark_ntriag", "mark_phase", "mark_dead", "mark_age")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _diff(st, ref):
    """Fields (and live marker arrays) that differ, with their rel-L2."""
    out = {}
    for f in state.State.FIELDS + list(state.TOPO) + ["mark_id_elem", "nmark_elem"]:
        a, b = getattr(st, f), getattr(ref, f)
        if not np.array_equal(a, b):
            out[f] = state.rel_l2(a, b)
    if st.nmarkers != ref.nmarkers:
        out["nmarkers"] = (st.nmarkers, ref.nmarkers)
    else:
        for f in MARKER_ARRAYS:
            a, b = st.live_markers(f), ref.live_markers(f)
            if not np.array_equal(a, b):
                out[f] = int((a != b).sum())
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("start,nsteps", [(0, 100), (5000, 3000), (8000, 4000)])
def test_surface_processes_bitwise_equal_to_oracle(start, nsteps):
    """nsteps device steps from the reference's state after `start` steps equal the
    correctly rounded oracle's, every field, the topography arrays and the markers
    (count, lists, barycentric coordinates, triangles, phases, dead flags, ages)."""
    from geoflac_amd import gpu
    _say("case: %s, %d steps from the %d-step reference state" % (CASE, nsteps, start))
    p = ol.golden_params(CASE)
    st = ol.golden_full(CASE, start, marker_room=True)
    ref = st.copy()
    n0, dead0 = st.nmarkers, int((st.live_markers("mark_dead") == 0).sum())
    with gpu.Model(p, st) as m:
        m.step(nsteps)
        m.download()
    assert ol.step(p, ref, nsteps, cr=True) == 0
    bad = _diff(st, ref)
    _say("  markers %d -> %d (dead %d -> %d); differing %s"
         % (n0, st.nmarkers, dead0, int((st.live_markers("mark_dead") == 0).sum()), bad))
    assert not bad, bad
    if start == 5000:
        assert st.nmarkers > n0  # resurface appended sediment markers on the way


@pytest.mark.gpu
def test_fl_move_with_resurface_bitwise_equal_to_oracle():
    """One fl_move (gf_fl_move) on a resurface step (nloop a multiple of ifreq_avgsr)
    from the 8000-step reference state: coordinates, topography, markers and the
    recounted phase ratios equal the oracle's fl_move."""
    from geoflac_amd import gpu
    p = ol.golden_params(CASE)
    st = ol.golden_full(CASE, 8000, marker_room=True)
    assert st.clock.nloop % p.ifreq_avgsr == 0
    ref = st.copy()
    _say("case: %s, fl_move at nloop %d" % (CASE, st.clock.nloop))
    with gpu.Model(p, st) as m:
        m.call("fl_move")
        m.download()
    assert ol.call("gfo_fl_move", p, ref, cr=True) == 0
    bad = _diff(st, ref)
    _say("  differing %s" % bad)
    assert not bad, bad
