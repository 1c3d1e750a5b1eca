"""rmasses and dt_mass at different frequencies (src/flac.f90:68-72): the step's move
leaves its areas pending, and the mass kernels form them from the moved coordinates.
With ifreq_rmasses 3 and ifreq_imasses 5, steps that run only one of the two launch
k_rmass or k_amass alone (after k_dens, or after k_dt_reduce and k_dt_scalars), and
steps that run both form the two masses in one pass (k_masses).  The device run is
compared bit for bit with the correctly rounded oracle from the same state."""
import os
import sys

import numpy as np
import pytest

from geoflac_amd import abi, inp, state, sweep

BASE = "tests/golden/inputs/box60x20.inp"


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.gpu
def test_masses_at_different_frequencies_bitwise_equal_to_oracle():
    from tests import oracle_lib as ol
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    q = inp.read_inp(os.path.join(root, BASE))
    q["ifreq_rmasses"] = 3
    q["ifreq_imasses"] = 5
    q["ntest_rem"] = 30  # (the reference's input rule: a multiple of every frequency)
    n = 31
    _say("case: box60x20 with ifreq_rmasses 3, ifreq_imasses 5, %d steps, device vs oracle" % n)
    gp, st, m = sweep.make_model(q, abi.PREC_FP64)
    m.download()
    ref = st.copy()
    m.step(n)
    m.download()
    m.close()
    assert ol.step(gp, ref, n, cr=True) == 0
    diff = {f: state.rel_l2(getattr(st, f), getattr(ref, f)) for f in state.State.FIELDS
            if not np.array_equal(getattr(st, f), getattr(ref, f))}
    _say("  differing fields %s" % diff)
    assert not diff
    assert st.clock.nloop == ref.clock.nloop and st.clock.time == ref.clock.time and st.clock.dt == ref.clock.dt
