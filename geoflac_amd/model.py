"""Host-side set-up of a layered-box model, following the reference's new-start path.

`build_box` mirrors src/setflac.f90:4-100 for the inputs this project runs
(uniform mesh zones, horizontal phase layers, `ictherm` 21 geotherm, velocity
boundary conditions): init_cord (src/init_cord.f90:6-80), init_temp
(src/init_temp.f90:26-60,205-213), init_stress (src/init_stress.f90:5-60),
init_bc/velbc/vbcal (src/init_bc.f90:4-330,376-417).  Phases come from the layer
depths sampled at the 9 regular marker positions of each element
(src/init_marker.f90:44-118) without the random jitter, and an optional seeded
perturbation of the layer interfaces.  The device then computes areas, real and
inertial masses, viscosity and dt (init_areas, rmasses, init_visc, dt_mass), as
the reference does at the end of setflac.
"""
import math

import numpy as np

from . import abi, inp
from .state import State

SEC_YEAR = inp.SEC_YEAR


def mesh1(x1, x2, n, nelz, sizez):
    """src/init_cord.f90:136-179 (one-dimensional graded mesh)."""
    nzon = len(nelz)
    xm = [0.0]
    all_ratio = 0.0
    for k in range(0, nzon - 1, 2):
        all_ratio = all_ratio + nelz[k] * sizez[k]
        for _ in range(nelz[k]):
            xm.append(xm[-1] + sizez[k])
        ratio = (sizez[k + 2] / sizez[k]) ** (1.0 / float(nelz[k + 1] + 1))
        for ii in range(1, nelz[k + 1] + 1):
            xm.append(xm[-1] + sizez[k] * (ratio ** ii))
            all_ratio = all_ratio + sizez[k] * (ratio ** ii)
    for _ in range(nelz[nzon - 1]):
        xm.append(xm[-1] + sizez[nzon - 1])
    all_ratio = all_ratio + nelz[nzon - 1] * sizez[nzon - 1]
    mag = (x2 - x1) / all_ratio
    return np.array([x1 + v * mag for v in xm[:n]])


def _bc_fun(b, x):
    """src/init_bc.f90:293-295 (velbc function of the normalised coordinate)."""
    pi2 = 2.0 * 3.14159
    return (b["bca"] + b["bcb"] * x + b["bcc"] * x * x
            + (b["bcd"] * math.cos(pi2 * b["bce"] * x) + b["bcf"] * math.sin(pi2 * b["bcg"] * x))
            * math.exp(-((x - b["bci"]) * b["bch"]) ** 2))


def init_bc(p, st):
    """Velocity boundary conditions (src/init_bc.f90:57-200 velbc part) and vbc (vbcal)."""
    nx, nz = p["nx"], p["nz"]
    cord = st.cord
    for b in p["bcs"]:
        side, n1, n2, nbc = b["nofside"], b["nbc1"], b["nbc2"], b["nbc"]
        if nbc in (2, 20, 40):
            raise NotImplementedError("stress boundary conditions (nystressbc) are not provided")
        ndbc = n2 - n1 + 1
        if side == 1:
            x1, x2 = cord[n1 - 1, 0, 1], cord[n2 - 1, 0, 1]
            nodes = [(n + n1 - 1, 1) for n in range(1, ndbc + 1)]
            xs = [(cord[jj - 1, 0, 1] - x1) / (x2 - x1) for jj, _ in nodes]
        elif side == 3:
            x1, x2 = cord[n1 - 1, nx - 1, 1], cord[n2 - 1, nx - 1, 1]
            nodes = [(n + n1 - 1, nx) for n in range(1, ndbc + 1)]
            xs = [(cord[jj - 1, nx - 1, 1] - x1) / (x2 - x1) for jj, _ in nodes]
        elif side == 2:
            x1, x2 = cord[nz - 1, n1 - 1, 0], cord[nz - 1, n2 - 1, 0]
            nodes = [(nz, n + n1 - 1) for n in range(1, ndbc + 1)]
            xs = [(cord[nz - 1, ii - 1, 0] - x1) / (x2 - x1) for _, ii in nodes]
        elif side == 4:
            x1, x2 = cord[0, n1 - 1, 0], cord[0, n2 - 1, 0]
            nodes = [(1, n + n1 - 1) for n in range(1, ndbc + 1)]
            xs = [(cord[0, ii - 1, 0] - x1) / (x2 - x1) for _, ii in nodes]
        else:
            raise NotImplementedError("boundary side %d" % side)
        for (jj, ii), x in zip(nodes, xs):
            if nbc not in (1, 10, 30, 200):
                continue
            fun = _bc_fun(b, x)
            if nbc == 10:
                st.ncod[jj - 1, ii - 1, 0] = 1
                if abs(st.bc[jj - 1, ii - 1, 0]) > 0.0:
                    fun = st.bc[jj - 1, ii - 1, 0]
                st.bc[jj - 1, ii - 1, 0] = fun
            if nbc in (1, 200):
                st.ncod[jj - 1, ii - 1, 1] = 1
                st.bc[jj - 1, ii - 1, 1] = 0.0 if nbc == 200 else fun
    vbc = 0.0
    m = (st.ncod == 1) | (st.ncod == 10) | (st.ncod == 30)
    if m.any():
        vbc = float(np.max(np.abs(st.bc[m])))
    if vbc == 0.0:
        vbc = 1.0e-10
    return vbc


def layer_phases(p, st, interface_noise=0.0, seed=0):
    """Phase ratios from the zone's layer depths, sampled at the 9 regular marker
    points of each element; interfaces optionally perturbed per column by a seeded
    uniform noise of +-interface_noise km."""
    nx, nz, nph = p["nx"], p["nz"], p["nphase"]
    ex, ez = nx - 1, nz - 1
    rng = np.random.default_rng(seed)
    x = st.cord[:, :, 0]
    z = st.cord[:, :, 1]
    f6, h = 0.8333333333333333, 0.5
    # marker points: a/b edge points then 3 along each (src/init_marker.f90:60-76)
    wts = []
    for ta in (f6, h, 1.0 / 6.0):
        for tb in (f6, h, 1.0 / 6.0):
            wts.append((ta, tb))
    counts = np.zeros((nph, ez, ex))
    for zone in p["zones"]:
        i1, i2 = zone["ixtb1"], zone["ixtb2"]
        cols = np.arange(ex)
        sel = (cols + 1 >= i1) & (cols + 1 <= i2)
        hc = np.array(zone["hc"], dtype=np.float64)
        noise = rng.uniform(-interface_noise, interface_noise, size=(ex, len(hc))) if interface_noise else 0.0
        hcol = hc[None, :] + noise  # (ex, nlayers-1), km
        for ta, tb in wts:
            # point on the left edge (a) and right edge (b), then between them
            za = z[:-1, :-1] * ta + z[1:, :-1] * (1 - ta)
            zb = z[:-1, 1:] * ta + z[1:, 1:] * (1 - ta)
            zz = za * tb + zb * (1 - tb)
            depth = zz * -1e-3
            kph = np.full((ez, ex), zone["iph_col"][-1], dtype=np.int64)
            for m in range(len(hc) - 1, -1, -1):
                kph = np.where(depth < hcol[None, :, m], zone["iph_col"][m], kph)
            for ph in set(zone["iph_col"]):
                counts[ph - 1] += np.where(sel[None, :], kph == ph, False)
    tot = counts.sum(axis=0)
    st.phase_ratio[...] = counts / np.maximum(tot, 1)[None]
    # most abundant phase, lowest id on ties (src/marker_data.f90:162-171)
    st.iphase[...] = (np.argmax(counts, axis=0) + 1).astype(np.int32)


def init_temp(p, st):
    """src/init_temp.f90:26-60 with the ictherm 21 profile (:205-213) and sources (:48-52)."""
    nx, nz = p["nx"], p["nz"]
    for zone in p["zones"]:
        if zone["ictherm"] != 21:
            raise NotImplementedError("geotherm type %d" % zone["ictherm"])
        for i in range(zone["ixtb1"], zone["ixtb2"] + 1):
            y = (st.cord[0, i - 1, 1] - st.cord[:, i - 1, 1]) * 1.0e-3
            t = y * (p["t_bot"] - p["t_top"]) / zone["age_1"]
            st.temp[:, i - 1] = np.where(t > p["t_bot"], p["t_bot"], t)
    for j in range(nz - 1):
        y = -(st.cord[j + 1, 0, 1] + st.cord[j, 0, 1]) / 2 / 1000
        st.source[j, :] = p["hs"] * math.exp(-y / p["hr"])


def init_stress(p, st):
    """Lithostatic initial stress, src/init_stress.f90:14-52 (returns pisos)."""
    nx, nz = p["nx"], p["nz"]
    ph = p["phases"]
    den = np.array([q["den"] for q in ph])
    alfa = np.array([q["alfa"] for q in ph])
    beta = np.array([q["beta"] for q in ph])
    g = p["g"]
    pisos = 0.0
    rogh = np.zeros(nx - 1)
    top = st.iphase[0, :]
    rogh = np.where((top == 3) | (top == 7), -1030.0 * g * 8.0e3, 0.0)
    for j in range(nz - 1):
        iph = st.iphase[j, :] - 1
        tmpr = 0.25 * (st.temp[j, :-1] + st.temp[j + 1, :-1] + st.temp[j, 1:] + st.temp[j + 1, 1:])
        densT = den[iph] * (1 - alfa[iph] * tmpr)
        dh1 = st.cord[j, :-1, 1] - st.cord[j + 1, :-1, 1]
        dh2 = st.cord[j, 1:, 1] - st.cord[j + 1, 1:, 1]
        dh = 0.5 * (dh1 + dh2)
        dPT = densT * g * dh
        dP = dPT * (1 - beta[iph] * rogh) / (1 + beta[iph] / 2 * dPT)
        press = rogh + 0.5 * dP
        for k in range(4):
            st.stress0[j, :, 0, k] = -press
            st.stress0[j, :, 1, k] = -press
            st.stress0[j, :, 2, k] = 0.0
            st.stress0[j, :, 3, k] = -press
        rogh = rogh + dP
    if p["nyhydro"] == 2:
        pisos = float(rogh[-1] + p["extra_pres"])
    return pisos


def build_box(p, interface_noise=0.0, seed=0):
    """State and parameter block of a new start (before the device-side set-up)."""
    nx, nz = p["nx"], p["nz"]
    st = State(nx, nz, p["nphase"])
    xm = mesh1(p["x0"], p["x0"] + p["rxbo"], nx, p["nelz_x"], p["sizez_x"])
    st.cord[:, :, 0] = xm[None, :]
    zm = mesh1(p["z0"], p["z0"] + p["rzbo"], nz, p["nelz_y"], p["sizez_y"])
    st.cord[:, :, 1] = zm[:, None]
    layer_phases(p, st, interface_noise, seed)
    ivis = int(any(p["phases"][k - 1]["irheol"] in (3,) or p["phases"][k - 1]["irheol"] >= 11
                   for k in np.unique(st.iphase)))
    init_temp(p, st)
    pisos = init_stress(p, st)
    vbc = init_bc(p, st)
    st.temp0[...] = st.temp
    st.e2sr[...] = 1e-16
    st.se2sr[...] = 1e-16
    c = st.clock
    c.nloop, c.nsrate, c.time, c.dt, c.dtavg = 0, -1, 0.0, 0.0, 0.0
    c.dt_elastic, c.dt_maxwell = 0.0, 0.0
    gp = inp.to_gf_params(p, vbc=vbc, ivis_present=ivis, pisos=pisos)
    return gp, st


def device_setup(model):
    """Device part of setflac (src/setflac.f90:35-91): areas, real masses, viscosity, dt,
    then dtavg = 0 and nsrate = -1 (src/setflac.f90:82-84)."""
    st = model.state
    model.init_areas()
    model.call("rmasses")
    if model.params.ivis_present == 1:
        model.init_visc()
    model.call("dt_mass")
    st.clock.dtavg = 0.0
    st.clock.nsrate = -1
    model.set_clock()
