"""Error paths of the C-ABI that a drop-in caller relies on (run on the device, every
kernel exiting normally): an unproven FP64 thermal step, a state whose phases have
change_phase rules but no markers, and more distinct phases than the device keeps
ratio planes for."""
import os
import sys

import numpy as np
import pytest

from geoflac_amd import abi, inp, sweep
from tests import oracle_lib as ol

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.gpu
def test_unproven_thermal_step_returns_its_status():
    """A 200x40 variant of box60x20 has cells taller than wide: its columns couple
    strongly enough that the tiled fixed point needs a second pass to be proven.  With
    the pass limit set to 1 from the host a step must return GF_ERR_THERM_SWEEPS
    instead of going on silently; with the default limit the same steps are proven."""
    from geoflac_amd import gpu
    _say("case: box60x20 variant 200x40, thermal pass limit 1, up to 10 steps")
    p = sweep.variant(inp.read_inp(os.path.join(ROOT, "tests/golden/inputs/box60x20.inp")), 200, 40)
    gp, st, m = sweep.make_model(p)
    with m:
        m.set_therm_passes(1)
        code = None
        for _ in range(10):
            try:
                m.step(1)
            except gpu.GfError as e:
                code = e.code
                break
        assert code == abi.GF_ERR_THERM_SWEEPS, code
        m.therm_stats()   # (clears the counters, which include the unproven step)
        m.set_therm_passes(4)
        m.step(10)
        unconv, extra, _, _ = m.therm_stats()
        _say("  with 4 passes: unproven %d, extra passes %d" % (unconv, extra))
        assert unconv == 0 and extra > 0


@pytest.mark.gpu
def test_step_refuses_rule_phases_without_markers():
    """box60x20 holds phase 4 (kmant1), which has change_phase rules: without markers
    the step cannot run change_phase and refuses (GF_ERR_STATE) rather than skip it."""
    from geoflac_amd import gpu
    _say("case: box60x20 step 0 without its markers")
    p = ol.golden_params("box60x20")
    st = ol.golden_state("box60x20", 0)
    st.mark_a1 = None  # drop the markers
    with gpu.Model(p, st) as m:
        with pytest.raises(gpu.GfError) as exc:
            m.step(1)
        assert exc.value.code == abi.GF_ERR_STATE
        assert "change_phase" in str(exc.value)


@pytest.mark.gpu
def test_too_many_phases_for_the_ratio_planes():
    """phases60x30 with five more phases present (markers of kcont1, kcont2, kweak,
    karc1 and kweakmc in a corner element each, ratios recounted): eight planes are in
    use before the first change_phase, whose rules add three phases.  The step returns
    GF_ERR_PHASES."""
    from geoflac_amd import gpu
    _say("case: phases60x30 with eight phases present, one step")
    p = ol.golden_params("phases60x30")
    st = ol.golden_state("phases60x30", 0)
    ez = st.nz - 1
    for n, ph in enumerate((2, 6, 12, 14, 15)):
        j, i = 0, 1 + n   # top-row elements
        e = j + ez * i
        ids = st.mark_id_elem[:st.nmark_elem[j, i], j, i] - 1
        st.mark_phase[ids] = ph
        st.phase_ratio[:, j, i] = 0.0
        st.phase_ratio[ph - 1, j, i] = 1.0
        st.iphase[j, i] = ph
    with gpu.Model(p, st) as m:
        with pytest.raises(gpu.GfError) as exc:
            m.step(1)
        assert exc.value.code == abi.GF_ERR_PHASES


@pytest.mark.gpu
def test_maxwell_stop_22_on_the_device():
    """src/rh_maxwell.f90:24-30: an element whose Maxwell number theta = rm dt / (2 visn)
    exceeds 0.5 stops the run (STOP 22).  Two elements of box60x20's 10-step state get
    a viscosity with theta = 0.6, and the loop counter is set so that the step updates
    no viscosity (nloop 11, not a multiple of ifreq_visc): the device
    returns GF_STOP_MAXWELL naming the first of them in the reference's loop order
    (column by column), and so does the oracle."""
    from geoflac_amd import gpu
    from tests.test_gpu_parity import golden_full
    _say("case: box60x20 10-step state, theta 0.6 at elements (5,7) and (12,3), one step")
    p = ol.golden_params("box60x20")
    st = golden_full(10, "box60x20")
    st.clock.nloop = 11
    for j, i in ((5, 7), (12, 3)):
        rm = p.rm[int(st.iphase[j - 1, i - 1]) - 1]
        st.visn[j - 1, i - 1] = rm * st.clock.dt / (2 * 0.6)
    ref = st.copy()
    assert ol.step(p, ref, 1, cr=True) == abi.GF_STOP_MAXWELL
    with gpu.Model(p, st) as m:
        with pytest.raises(gpu.GfError) as exc:
            m.step(1)
    _say("  %s" % exc.value)
    assert exc.value.code == abi.GF_STOP_MAXWELL
    assert "(12,3)" in str(exc.value)
