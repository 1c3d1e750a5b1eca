"""Batches of independent models (gf_batch_*): every model of a batch ends bit for bit
where it ends when run alone (the batched launches run the single-model kernel bodies
on each model's own parameters, clock and arrays), with and without graph replay and
after the workgroup-size autotuner; plus a seeded sweep's variants on the CPU."""
import sys

import numpy as np
import pytest

from geoflac_amd import abi, inp, state, sweep

BASE = "tests/golden/inputs/box60x20.inp"


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _variants():
    import os
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    p = inp.read_inp(os.path.join(root, BASE))
    return [sweep.variant(p, 60, 20), sweep.variant(p, 37, 11, 22.0, 3.0e7, 2.0e-10),
            sweep.variant(p, 45, 27, 31.0, 5.5e7, 0.7e-10)]


def test_sweep_variants_are_consistent():
    import os
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    p = inp.read_inp(os.path.join(root, BASE))
    ps = sweep.sweep_params(p, n=16, seed=3)
    for q in ps:
        assert 100 <= q["nelx"] <= 600 and 40 <= q["nelz"] <= 150
        assert q["nx"] == q["nelx"] + 1 and all(b["nbc2"] in (q["nx"], q["nz"]) for b in q["bcs"])
        assert all(10.0 <= ph["fric1"] <= 35.0 and 10e6 <= ph["cohesion1"] <= 60e6 for ph in q["phases"])
    assert [q["nelx"] for q in ps] == [q["nelx"] for q in sweep.sweep_params(p, n=16, seed=3)]


def _run(precision, nsteps, graph=False, tune=False):
    models, states = [], []
    for q in _variants():
        gp, st, m = sweep.make_model(q, precision)
        models.append(m)
    singles = []
    for q in _variants():  # the same models, each alone
        gp, st, m = sweep.make_model(q, precision)
        m.step(nsteps)
        m.download()
        singles.append(st)
        m.close()
    with sweep.Batch(models) as b:
        if graph:
            b.set_graph(True)
        done = 0
        if tune:
            sizes = b.autotune(2)
            _say("  autotuned workgroup sizes (rheology, node, element move): %s" % sizes)
            done = 6
        b.step(nsteps - done)
    for m in models:
        m.download()
        states.append(m.state)
    for m in models:
        m.close()
    return states, singles


@pytest.mark.gpu
@pytest.mark.parametrize("graph,tune", [(False, False), (True, False), (True, True)])
def test_batch_equals_single_runs(graph, tune):
    n = 40
    _say("case: batch of 3 models (60x20, 37x11, 45x27), %d steps, graph %s, autotune %s" % (n, graph, tune))
    batch, single = _run(abi.PREC_FP64, n, graph, tune)
    for k, (a, b) in enumerate(zip(batch, single)):
        diff = [f for f in state.State.FIELDS if not np.array_equal(getattr(a, f), getattr(b, f))]
        _say("  model %d: nloop %d, differing fields %s" % (k, a.clock.nloop, diff))
        assert not diff
        assert a.clock.nloop == b.clock.nloop == n
        assert a.clock.time == b.clock.time and a.clock.dt == b.clock.dt


@pytest.mark.gpu
def test_batch_mixed_precision_equals_single_runs():
    n = 20
    _say("case: batch of 3 models in mixed precision, %d steps with graph replay" % n)
    batch, single = _run(abi.PREC_MIXED, n, graph=True)
    for a, b in zip(batch, single):
        diff = [f for f in state.State.FIELDS if not np.array_equal(getattr(a, f), getattr(b, f))]
        assert not diff, diff


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [abi.PREC_FP64, abi.PREC_MIXED])
def test_batch_then_single_steps_equal_single_runs(precision):
    """A batch leaves its last element move pending on the models (the rotation, and in
    FP32-arithmetic mode the areas); stepping the models on their own afterwards, and
    batching them again, ends bit for bit where single runs end."""
    n1, n2, n3 = 13, 4, 10
    _say("case: precision %d, batch %d steps (graph), single %d, batch %d" % (precision, n1, n2, n3))
    models, singles = [], []
    for q in _variants():
        gp, st, m = sweep.make_model(q, precision)
        models.append(m)
    for q in _variants():
        gp, st, m = sweep.make_model(q, precision)
        m.step(n1 + n2 + n3)
        m.download()
        singles.append(st)
        m.close()
    with sweep.Batch(models) as b:
        b.set_graph(True)
        b.step(n1)
    for m in models:
        m.step(n2)
    with sweep.Batch(models) as b:
        b.set_graph(True)
        b.step(n3)
    for m, ref in zip(models, singles):
        m.download()
        diff = [f for f in state.State.FIELDS if not np.array_equal(getattr(m.state, f), getattr(ref, f))]
        _say("  nloop %d, differing fields %s" % (m.state.clock.nloop, diff))
        assert not diff
        m.close()


@pytest.mark.gpu
@pytest.mark.parametrize("therm", ["tail", "multi"])
def test_thermal_fallback_passes_bitwise_equal_to_oracle(therm, monkeypatch):
    """Cells 1.7 times taller than wide (200x40 elements over the 300x100 km box)
    couple neighbouring columns strongly, so that the thermal tiles' halos do not
    settle every output node and fallback passes run (the pipelined tile kernel from
    the assembled iterate), either inside the tail launch with grid barriers between
    the phases (the default, whose result becomes the new temperature in place: after
    an odd pass it is formed from the second iterate buffer) or as a launch per phase
    (GF_THERM=multi, copied to temp).  The run stays bitwise equal to the correctly
    rounded oracle started from the same state."""
    import os
    from tests import oracle_lib as ol
    if therm == "multi":
        monkeypatch.setenv("GF_THERM", "multi")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    q = sweep.variant(inp.read_inp(os.path.join(root, BASE)), 200, 40)
    n = 30
    _say("case: FP64 200x40 variant, %d steps, thermal %s, device vs oracle" % (n, therm))
    gp, st, m = sweep.make_model(q, abi.PREC_FP64)
    m.download()
    ref = st.copy()
    m.therm_stats()
    m.step(n)
    m.download()
    unconv, extra, mism, mulp = m.therm_stats()
    m.close()
    assert ol.step(gp, ref, n, cr=True) == 0
    diff = [f for f in state.State.FIELDS if not np.array_equal(getattr(st, f), getattr(ref, f))]
    _say("  thermal: unproven %d, extra passes %d, border mismatches %d (max %d ulp); differing fields %s"
         % (unconv, extra, mism, mulp, diff))
    assert extra > 0, "the case no longer exercises the fallback passes"
    assert unconv == 0
    assert not diff


@pytest.mark.gpu
def test_fp32_arithmetic_single_thermal_pass():
    """In FP32-arithmetic mode the thermal step takes the single tile pass without the
    fixed-point proof.  On the 200x40 variant, whose proof needs fallback passes in
    FP64, the temperature stays within 1e-9 relative L2 of the FP64 run after 30
    steps."""
    import os
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    q = sweep.variant(inp.read_inp(os.path.join(root, BASE)), 200, 40)
    n = 30
    _say("case: 200x40 variant, %d steps, FP32 arithmetic vs FP64 temperature" % n)
    out = {}
    for prec in (abi.PREC_FP64, abi.PREC_MIXED):
        gp, st, m = sweep.make_model(q, prec)
        m.therm_stats()
        m.step(n)
        m.download()
        out[prec] = (st, m.therm_stats())
        m.close()
    err = state.rel_l2(out[abi.PREC_MIXED][0].temp, out[abi.PREC_FP64][0].temp)
    _say("  temp rel-L2 %.2e; FP32-mode thermal stats %s" % (err, out[abi.PREC_MIXED][1]))
    assert out[abi.PREC_MIXED][1][0] == 0 and out[abi.PREC_MIXED][1][1] == 0
    assert err <= 1e-9
