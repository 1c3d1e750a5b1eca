"""Host-side new-start set-up (geoflac_amd/model.py) against the reference's setflac.

The host set-up (mesh, markers and the phases they give, geotherm, lithostatic stress,
boundary conditions) followed by the oracle's init_areas / rmasses / init_visc /
dt_mass must reproduce the reference's step-0 state bit for bit, markers included.
"""
import ctypes as C
import os

import numpy as np
import pytest

from geoflac_amd import inp, model
from tests import oracle_lib as ol

CASES = ["box60x20", "layers400x100", "hetero80x40", "phases60x30", "bcs60x20", "idt1_60x20", "topo60x20"]
MARKERS = ["mark_a1", "mark_a2", "mark_ntriag", "mark_phase", "mark_dead", "mark_id_elem", "nmark_elem"]


@pytest.mark.parametrize("case", CASES)
def test_host_setup_matches_reference_new_start(case):
    p = inp.read_inp(os.path.join(ol.GOLDEN, "inputs", case + ".inp"))
    gp, st = model.build_box(p)
    ref = ol.golden_state(case, 0)
    for f in MARKERS + ["phase_ratio", "iphase", "nopbou", "bcstress"]:
        a = st.live_markers(f) if f in MARKERS[:5] else getattr(st, f)
        assert np.array_equal(a, getattr(ref, f)), f
    refp = ol.golden_params(case)
    assert gp.vbc == refp.vbc
    assert gp.ivis_present == refp.ivis_present
    assert gp.nelem_serp == refp.nelem_serp
    assert gp.nopbmax == refp.nopbmax
    assert gp.pisos == refp.pisos
    assert gp.topo_kappa == refp.topo_kappa
    assert list(gp.erodibility)[:gp.nphase] == list(refp.erodibility)[:refp.nphase]
    h = st.host_arrays()
    L = ol.lib()
    assert L.gfo_init_areas(C.byref(gp), C.byref(h)) == 0
    assert L.gfo_rmasses(C.byref(gp), C.byref(h), C.byref(st.clock)) == 0
    assert L.gfo_init_visc(C.byref(gp), C.byref(h)) == 0
    assert L.gfo_dt_mass(C.byref(gp), C.byref(h), C.byref(st.clock)) == 0
    for f in st.FIELDS:
        if f == "force":
            continue
        a, b = getattr(st, f), getattr(ref, f)
        assert np.array_equal(a, b), f
    assert st.clock.dt == ref.clock.dt
    assert st.clock.dt_elastic == ref.clock.dt_elastic


def test_layer_phases_regular_markers():
    """Without markers the phases come from the 9 regular points of each element; on
    box60x20 the 35 km interface lies on an element boundary, so the random offsets
    of the markers cannot change the count."""
    case = "box60x20"
    p = inp.read_inp(os.path.join(ol.GOLDEN, "inputs", case + ".inp"))
    gp, st = model.build_box(p, markers=False)
    ref = ol.golden_state(case, 0)
    assert np.array_equal(st.phase_ratio, ref.phase_ratio)
    assert np.array_equal(st.iphase, ref.iphase)


def test_cohesion_noise_is_seeded_and_bounded():
    p = inp.read_inp(os.path.join(ol.GOLDEN, "inputs", "box60x20.inp"))
    _, a = model.build_box(p, coh_noise=0.01, seed=7)
    _, b = model.build_box(p, coh_noise=0.01, seed=7)
    assert np.array_equal(a.coh_scale, b.coh_scale)
    assert a.coh_scale.dtype == np.float32 and a.coh_scale.shape == (20, 60)
    assert np.all(np.abs(a.coh_scale - 1.0) <= 0.0100001) and a.coh_scale.std() > 0.004
