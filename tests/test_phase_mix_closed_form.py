"""Phase mixing of the oracle against closed forms, on one element with fractional
phase ratios (four phases active).

The reference mixes per-phase properties by the element's phase ratios: the
viscosity and the friction angle and cohesion harmonically (src/matprops.f90:191-203,
src/rh_plastic.f90:250-278), dilation, hardening and density arithmetically
(src/rh_plastic.f90:274-275, src/matprops.f90:60-74), skipping phases with a ratio
below 0.01.  Each case below picks parameters for which the per-phase value is a
round number, so the expected mix is a textbook mean that an arithmetic mean in
place of a harmonic one (or the reverse), a dropped phase or a wrong weight fails.
"""
import ctypes as C

import numpy as np
import pytest

from geoflac_amd import abi, state
from tests import oracle_lib as ol


def one_element(nphase=8):
    """A 1x1-element state 10 km x 10 km, top at 0, and a parameter block."""
    p = abi.GfParams()
    p.nx, p.nz, p.nphase = 2, 2, nphase
    p.v_min, p.v_max = 1e18, 1e27
    p.g, p.rxbo, p.vbc = 10.0, 1e4, 1e-10
    st = state.State(2, 2, nphase)
    st.cord[:, :, 0] = np.array([[0.0, 1e4], [0.0, 1e4]])
    st.cord[:, :, 1] = np.array([[0.0, 0.0], [-1e4, -1e4]])
    st.temp[...] = 500.0
    st.e2sr[...] = 1e-15
    return p, st


def set_ratios(st, ratios):
    st.phase_ratio[...] = 0.0
    for k, r in ratios.items():
        st.phase_ratio[k - 1, 0, 0] = r
    st.iphase[0, 0] = max(ratios, key=lambda k: (ratios[k], -k))


def eff_visc(p, st):
    return ol.lib().gfo_eff_visc(C.byref(p), C.byref(st.host_arrays()), 1, 1)


def eff_dens(p, st):
    return ol.lib().gfo_eff_dens(C.byref(p), C.byref(st.host_arrays()), 1, 1)


def pre_plast(p, st):
    out = (C.c_double * 4)()
    ol.lib().gfo_pre_plast(C.byref(p), C.byref(st.host_arrays()), 1, 1, out)
    return tuple(out)


def linear_creep(p, k, eta):
    """Phase k with n = 1 and no activation energy: the creep law reduces to
    eta = 0.25 * (0.75 A)^-1 * 1e6 = 1e6 / (3 A), independent of T, P and strain rate."""
    p.pln[k - 1] = 1.0
    p.eactiv[k - 1] = 0.0
    p.vactiv[k - 1] = 0.0
    p.acoef[k - 1] = 1e6 / (3.0 * eta)


def test_viscosity_mix_is_harmonic_over_four_phases():
    p, st = one_element()
    etas = {2: 1e20, 4: 1e21, 6: 1e22, 8: 1e23}
    for k, e in etas.items():
        linear_creep(p, k, e)
    ratios = {2: 0.25, 4: 0.125, 6: 0.375, 8: 0.25}
    set_ratios(st, ratios)
    want = 1.0 / sum(r / etas[k] for k, r in ratios.items())
    arith = sum(r * etas[k] for k, r in ratios.items())
    got = eff_visc(p, st)
    assert got == pytest.approx(want, rel=1e-14)
    assert abs(got - arith) > 0.5 * arith   # the arithmetic mean is two decades off


def test_viscosity_mix_skips_ratios_below_one_percent_and_clips_each_phase():
    p, st = one_element()
    linear_creep(p, 2, 1e20)
    linear_creep(p, 4, 1e16)         # below v_min = 1e18: clipped to 1e18 before mixing
    linear_creep(p, 6, 1e10)         # ratio 0.005 < 0.01: ignored, or it would dominate
    set_ratios(st, {2: 0.6, 4: 0.395, 6: 0.005})
    want = 1.0 / (0.6 / 1e20 + 0.395 / 1e18)
    assert eff_visc(p, st) == pytest.approx(want, rel=1e-14)


def test_viscosity_temperature_uses_the_element_mean_plus_the_adiabat():
    """n = 1 with activation energy E: eta = 1e6/(3A) * exp(E / (R (T + 273))), where T
    is the mean of the four corner temperatures plus 0.3 K/km times the depth of the
    element centre below the surface (5 km here: +1.5 K)."""
    p, st = one_element()
    linear_creep(p, 4, 1e20)
    p.eactiv[3] = 1.0e5
    set_ratios(st, {4: 1.0})
    st.temp[...] = np.array([[400.0, 400.0], [600.0, 600.0]])
    t = 500.0 + 1.5
    want = 1e20 * np.exp(1.0e5 / (8.31448 * (t + 273.0)))
    assert eff_visc(p, st) == pytest.approx(want, rel=1e-12)


def test_friction_and_cohesion_mix_harmonically_dilation_and_hardening_arithmetically():
    p, st = one_element()
    # phase 2: phi 30 -> 10, c 40 -> 4 MPa, psi 10 -> 0 over plastic strain 0 -> 1
    # phase 4: phi 10 -> 10, c 10 -> 10 MPa, psi 0 -> 0
    for k, (f1, f2, c1, c2, d1, d2) in {2: (30, 10, 4e7, 4e6, 10, 0), 4: (10, 10, 1e7, 1e7, 0, 0)}.items():
        p.fric1[k - 1], p.fric2[k - 1] = f1, f2
        p.cohesion1[k - 1], p.cohesion2[k - 1] = c1, c2
        p.dilat1[k - 1], p.dilat2[k - 1] = d1, d2
        p.plstrain1[k - 1], p.plstrain2[k - 1] = 0.0, 1.0
    set_ratios(st, {2: 0.5, 4: 0.5})
    st.aps[...] = 0.0   # aps not below plstrain1 = 0: softening branch at dpl = 0
    coh, phi, psi, hardn = pre_plast(p, st)
    assert phi == pytest.approx(1 / (0.5 / 30 + 0.5 / 10), rel=1e-15)    # 15, not 20
    assert coh == pytest.approx(1 / (0.5 / 4e7 + 0.5 / 1e7), rel=1e-15)  # 1.6e7, not 2.5e7
    assert psi == pytest.approx(0.5 * 10, rel=1e-15)
    assert hardn == pytest.approx(0.5 * (4e6 - 4e7) + 0.5 * 0.0, rel=1e-15)
    st.aps[...] = 0.5   # half-way: phase 2 at phi 20, c 22 MPa, psi 5
    coh, phi, psi, hardn = pre_plast(p, st)
    assert phi == pytest.approx(1 / (0.5 / 20 + 0.5 / 10), rel=1e-14)
    assert coh == pytest.approx(1 / (0.5 / 2.2e7 + 0.5 / 1e7), rel=1e-14)
    assert psi == pytest.approx(2.5, rel=1e-14)


def test_cohesion_scale_multiplies_the_mixed_cohesion():
    p, st = one_element()
    for k in (2, 4):
        p.fric1[k - 1] = p.fric2[k - 1] = 30.0
        p.cohesion1[k - 1] = p.cohesion2[k - 1] = 2e7 if k == 2 else 5e7
        p.plstrain1[k - 1], p.plstrain2[k - 1] = 0.0, 1.0
    set_ratios(st, {2: 0.5, 4: 0.5})
    base = pre_plast(p, st)[0]
    st.coh_scale = np.full((1, 1), 1.01, dtype=np.float32, order="F")
    assert pre_plast(p, st)[0] == base * float(np.float32(1.01))


def test_density_mix_is_arithmetic_and_carries_the_pressure_between_phases():
    """Without thermal expansion and compressibility the mix is sum(r_k rho_k).  With
    compressibility beta the reference overwrites the pressure inside the phase loop
    (press = dens * g * zcord, src/matprops.f90:64-68), so the second phase starts
    from the first phase's pressure: the result depends on which phases are mixed."""
    p, st = one_element()
    p.den[1], p.den[3], p.den[5], p.den[7] = 2800.0, 3300.0, 2900.0, 3400.0
    st.iphase[0, 0] = 2
    ratios = {2: 0.25, 4: 0.25, 6: 0.125, 8: 0.375}
    set_ratios(st, ratios)
    st.iphase[0, 0] = 2          # a crustal iphase: no mantle density amplifier
    want = sum(r * p.den[k - 1] for k, r in ratios.items())
    assert eff_dens(p, st) == pytest.approx(want, rel=1e-15)
    # compressibility: hand-evaluate the carried pressure for two phases
    p.beta[1] = p.beta[3] = 1e-11
    set_ratios(st, {2: 0.5, 4: 0.5})
    st.iphase[0, 0] = 2
    st.stress0[...] = -1e8       # press = -(s11+s22+s33)*4/12 = 1e8
    zc = -5e3
    press = 1e8
    ed = 0.0
    for k in (2, 4):
        d = p.den[k - 1] * (1 + 1e-11 * press)
        press = d * 10.0 * zc
        d = p.den[k - 1] * (1 + 1e-11 * press)
        ed += 0.5 * d
    got = eff_dens(p, st)
    assert got == pytest.approx(ed, rel=1e-14)
    fresh = sum(0.5 * p.den[k - 1] * (1 + 1e-11 * (p.den[k - 1] * (1 + 1e-11 * 1e8) * 10.0 * zc)) for k in (2, 4))
    assert got != pytest.approx(fresh, rel=1e-12)
