"""Host-side new-start set-up (geoflac_amd/model.py) against the reference's setflac.

Given the reference's own marker-derived phase ratios, the host set-up followed by
the oracle's init_areas / rmasses / init_visc / dt_mass must reproduce the
reference's step-0 state bit for bit.
"""
import ctypes as C
import os

import numpy as np

from geoflac_amd import inp, model
from tests import oracle_lib as ol

CASE = "box60x20"


def test_host_setup_matches_reference_new_start():
    p = inp.read_inp(os.path.join(ol.GOLDEN, "inputs", CASE + ".inp"))
    gp, st = model.build_box(p)
    ref = ol.golden_state(CASE, 0)
    # phases: markers are jittered randomly in the reference; take its ratios
    st.phase_ratio[...] = ref.phase_ratio
    st.iphase[...] = ref.iphase
    pisos = model.init_stress(p, st)
    refp = ol.golden_params(CASE)
    assert gp.vbc == refp.vbc
    assert gp.ivis_present == refp.ivis_present
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
    p = inp.read_inp(os.path.join(ol.GOLDEN, "inputs", CASE + ".inp"))
    gp, st = model.build_box(p)
    ref = ol.golden_state(CASE, 0)
    # the 35 km interface lies on an element boundary: jitter cannot change the count
    assert np.array_equal(st.phase_ratio, ref.phase_ratio)
    assert np.array_equal(st.iphase, ref.iphase)
