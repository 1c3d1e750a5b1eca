"""The viscosity surrogate on the device: training kernels against the CPU reference
(tests/mlp_ref.py, pinned by finite differences), the BF16 matrix-core inference
against the same network rounded to BF16 on the CPU, and the surrogate inside the
step (gate, audit, error against the exact Eff_visc)."""
import sys

import numpy as np
import pytest

from geoflac_amd import abi
from tests import mlp_ref
from tests import oracle_lib as ol

CASE = "box60x20"


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.gpu
def test_training_gradient_matches_cpu_reference():
    from geoflac_amd import gpu
    _say("case: surrogate loss/gradient on 1000 random samples, device FP32 vs CPU float64")
    rng = np.random.default_rng(7)
    w = (mlp_ref.init(seed=2) + rng.normal(0, 0.05, mlp_ref.NPARAM)).astype(np.float32)
    x = rng.normal(size=(1000, 7)).astype(np.float32)
    t = rng.normal(size=1000).astype(np.float32)
    loss, g = gpu.mlp_grad(x, t, w)
    lref, gref = mlp_ref.loss_grad(w.astype(np.float64), x.astype(np.float64), t.astype(np.float64))
    rel = np.linalg.norm(g - gref) / np.linalg.norm(gref)
    _say("  loss %.6e vs %.6e, gradient rel-L2 %.2e, max abs %.2e" % (loss, lref, rel, np.abs(g - gref).max()))
    assert abs(loss - lref) <= 1e-5 * lref
    assert rel <= 1e-5
    for name, sl in (("W1", slice(0, 448)), ("b1", slice(448, 512)), ("W2", slice(512, 4608)),
                     ("b2", slice(4608, 4672)), ("W3", slice(4672, 4736)), ("b3", slice(4736, 4737))):
        r = np.linalg.norm(g[sl] - gref[sl]) / max(np.linalg.norm(gref[sl]), 1e-30)
        assert r <= 1e-4, (name, r)


@pytest.mark.gpu
def test_matrix_core_inference_matches_bf16_reference():
    from geoflac_amd import gpu
    _say("case: surrogate BF16 MFMA inference on 3000 random inputs vs the CPU BF16 network")
    rng = np.random.default_rng(11)
    w = (mlp_ref.init(seed=5) + rng.normal(0, 0.1, mlp_ref.NPARAM)).astype(np.float32)
    x = rng.normal(size=(3000, 7)).astype(np.float32)
    y = gpu.mlp_infer(x, w)
    yb = mlp_ref.forward_bf16(w, x)
    y64, _ = mlp_ref.forward(w.astype(np.float64), x.astype(np.float64))
    eb, e64 = np.abs(y - yb).max(), np.abs(y - y64).max()
    _say("  max |y - bf16 ref| %.2e, max |y - fp64 ref| %.2e, |y| max %.2f" % (eb, e64, np.abs(y64).max()))
    assert eb <= 2e-3 * max(1.0, np.abs(yb).max())
    assert e64 <= 5e-2 * max(1.0, np.abs(y64).max())


def _trained_model(precision=abi.PREC_FP64, steps=3000):
    """The 60x20 box after 20 steps; the surrogate trained on the viscosities of
    every step up to there."""
    from geoflac_amd import gpu
    p = ol.golden_params(CASE)
    st = ol.golden_state(CASE, 0)
    m = gpu.Model(p, st, precision=precision)
    for k in range(20):
        m.step(1)
        m.surrogate_sample(100000)
    l0, l1 = m.surrogate_train(steps, batch=1024, lr=2e-3, seed=1)
    return m, l0, l1


@pytest.mark.gpu
def test_surrogate_in_the_step():
    from geoflac_amd import gpu, state
    _say("case: surrogate trained on 20 steps of %s, viscosity update with and without it" % CASE)
    m, l0, l1 = _trained_model()
    _say("  training loss (normalised) %.3e -> %.3e" % (l0, l1))
    assert l1 < 0.2 * l0
    st = m.state
    m.download()
    saved = st.copy()
    st.clock.nloop = 10                 # a viscosity-update step
    m.set_clock()
    m.call("fl_rheol")
    m.download()
    exact = st.visn.copy()
    for f in state.State.FIELDS:        # same state again, surrogate on
        getattr(st, f)[...] = getattr(saved, f)
    st.clock.nloop = 10
    m.upload()
    m.set_clock()
    m.surrogate_enable(True, tol=1.0, audit_stride=7)
    m.call("fl_rheol")
    m.download()
    s = m.surrogate_stats()
    lerr = np.abs(np.log10(st.visn) - np.log10(exact))
    _say("  stats %s; |log10 visn error| max %.3e mean %.3e" % (s, lerr.max(), lerr.mean()))
    assert s["evals"] > 0 and s["off"] == 0
    assert s["evals"] + s["fallbacks"] + s["audits"] >= 1200
    assert s["audits"] > 0
    # audited runs of 64 elements (run = 10 mod 7, in the reference's element order)
    # get the exact value
    e_order = lerr.reshape(-1, order="F")
    runs = np.arange(e_order.size) // 64
    assert np.all(e_order[runs % 7 == 3] == 0.0)
    assert np.any(e_order[runs % 7 != 3] != 0.0)
    # the Arrhenius factor spans ~10 decades over the temperature range, so the BF16
    # input (8 significant bits) alone limits the error to a few 0.01 decades
    assert lerr.mean() <= 0.1
    assert lerr.max() <= 0.6
    m.close()


@pytest.mark.gpu
def test_surrogate_gate_switches_off_on_audit():
    from geoflac_amd import gpu
    _say("case: surrogate audit with a zero tolerance switches it off")
    m, _, _ = _trained_model(steps=300)
    st = m.state
    st.clock.nloop = 10
    m.set_clock()
    m.surrogate_enable(True, tol=0.0, audit_stride=1)
    m.call("fl_rheol")
    s1 = m.surrogate_stats()
    st.clock.nloop = 20
    m.set_clock()
    m.call("fl_rheol")
    s2 = m.surrogate_stats()
    _say("  after the first call %s; after the second %s" % (s1, s2))
    assert s1["off"] == 1
    # switched off: the second call evaluates no element on the network
    assert s2["evals"] == s1["evals"] and s2["audits"] == s1["audits"]
    assert s2["fallbacks"] > s1["fallbacks"]
    m.close()


@pytest.mark.gpu
def test_surrogate_gate_takes_failed_cells_off():
    """An audit above the tolerance (but below four times it) takes its (1/T, log10
    strain rate) cell off the surrogate: the next call sends that cell's elements to
    the exact formula, and the surrogate stays on elsewhere."""
    from geoflac_amd import gpu
    _say("case: surrogate audit takes the cells of failed audits off")
    m, _, _ = _trained_model(steps=300)
    st = m.state
    st.clock.nloop = 10
    m.set_clock()
    m.surrogate_enable(True, tol=10.0, audit_stride=1)
    m.call("fl_rheol")
    probe = m.surrogate_stats()
    assert probe["audits"] > 0 and probe["failed_audits"] == 0
    tol = 0.5 * probe["max_audit_err"]
    st.clock.nloop = 20
    m.set_clock()
    m.surrogate_enable(True, tol=tol, audit_stride=1)
    m.call("fl_rheol")
    s1 = m.surrogate_stats()
    st.clock.nloop = 30
    m.set_clock()
    m.call("fl_rheol")
    s2 = m.surrogate_stats()
    _say("  tol %.3e; first call %s; second %s" % (tol, s1, s2))
    assert s1["off"] == 0 and s1["failed_audits"] > 0 and s1["cells_off"] > 0
    assert s2["off"] == 0
    # fewer elements inside the gate in the second call, and still some (with
    # audit_stride 1 every element inside the gate is audited)
    n1 = s1["evals"] + s1["audits"]
    n2 = s2["evals"] + s2["audits"] - n1
    assert 0 < n2 < n1
    m.close()
