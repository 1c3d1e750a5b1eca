"""Host-side set-up of a layered-box model, following the reference's new-start path.

`build_box` mirrors src/setflac.f90:4-100 for the inputs this project runs
(uniform mesh zones, horizontal phase layers, `ictherm` 21 geotherm, velocity
boundary conditions): init_cord (src/init_cord.f90:6-80), the markers and the
phases they give (init_marker, src/init_marker.f90:1-169, with the serial build's
seeded random offsets, and count_phase_ratio, src/marker_data.f90:140-174),
init_temp (src/init_temp.f90:26-60,205-213), init_stress (src/init_stress.f90:5-60),
init_bc/velbc/vbcal (src/init_bc.f90:4-330,376-417).  The device then computes
areas, real and inertial masses, viscosity and dt (init_areas, rmasses, init_visc,
dt_mass), as the reference does at the end of setflac.
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
    """Velocity and stress boundary conditions (src/init_bc.f90:57-200 with velbc and
    stressbc, :226-330) and vbc (vbcal).  Returns (vbc, nopbmax)."""
    nx, nz = p["nx"], p["nz"]
    cord = st.cord
    nnop = 0
    for b in p["bcs"]:
        side, n1, n2, nbc = b["nofside"], b["nbc1"], b["nbc2"], b["nbc"]
        ndbc = n2 - n1 + 1
        if nbc in (2, 20, 40):
            ndbc = ndbc - 1
            if n2 < nz:     # (the reference compares with nz on every side)
                ndbc = ndbc + 1
            if side in (1, 3):
                col = 0 if side == 1 else nx - 1
                x1, x2 = cord[n1 - 1, col, 1], cord[n2 - 1, col, 1]
            else:
                row = nz - 1 if side == 2 else 0
                x1, x2 = cord[row, n1 - 1, 0], cord[row, n2 - 1, 0]
            for n in range(1, ndbc + 1):
                numbp = n + n1 - 1
                if side == 1:
                    numbp1 = numbp + 1
                    x, xn = (cord[numbp - 1, 0, 1] - x1) / (x2 - x1), (cord[numbp1 - 1, 0, 1] - x1) / (x2 - x1)
                    nodes = (1, 1, numbp, numbp1)
                elif side == 3:
                    numbp1 = numbp + 1
                    x = (cord[numbp - 1, nx - 1, 1] - x1) / (x2 - x1)
                    xn = (cord[numbp1 - 1, nx - 1, 1] - x1) / (x2 - x1)
                    nodes = (nx, nx, numbp, numbp1)
                elif side == 2:
                    numbp1 = n + 1  # (the reference's second node of a bottom segment)
                    x = (cord[nz - 1, numbp - 1, 0] - x1) / (x2 - x1)
                    xn = (cord[nz - 1, numbp1 - 1, 0] - x1) / (x2 - x1)
                    nodes = (numbp, numbp1, nz, nz)
                else:
                    numbp1 = numbp + 1
                    x, xn = (cord[0, numbp - 1, 0] - x1) / (x2 - x1), (cord[0, numbp1 - 1, 0] - x1) / (x2 - x1)
                    nodes = (numbp, numbp1, 1, 1)
                nnop += 1
                xa = 0.5 * (x + xn)
                fun = _bc_fun(b, xa)
                st.nopbou[nnop - 1, :] = nodes
                if nbc == 20:
                    st.bcstress[nnop - 1, 0] = fun
                if nbc == 2:
                    st.bcstress[nnop - 1, 1] = fun
            continue
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
    return vbc, nnop


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


# ---------------------------------------------------------------- markers
_LCG_M, _LCG_A, _LCG_C = 134456, 8121, 28411   # src/myrandom_mod.f90:9-16


def _lcg(s):
    """One draw of myrandom (src/myrandom_mod.f90:9-16): the new seed (r = seed / 134456)."""
    return (_LCG_A * (s % _LCG_M) + _LCG_C) % _LCG_M


def _marker_points(cord, j, i):
    """The 9 regular points of elements (j, i) (0-based index arrays) and dx, dy
    (src/init_marker.f90:44-82)."""
    f6 = 0.8333333333333333333333
    o6 = 0.1666666666666666666666
    h = 0.5
    c = lambda jj, ii, d: cord[jj, ii, d]
    pts = []
    for wa, wb in ((f6, o6), (h, h), (o6, f6)):
        ax = c(j, i, 0) * wa + c(j + 1, i, 0) * wb
        ay = c(j, i, 1) * wa + c(j + 1, i, 1) * wb
        bx = c(j, i + 1, 0) * wa + c(j + 1, i + 1, 0) * wb
        by = c(j, i + 1, 1) * wa + c(j + 1, i + 1, 1) * wb
        for u, v in ((f6, o6), (h, h), (o6, f6)):
            pts.append((ax * u + bx * v, ay * u + by * v))
    dx = c(j, i + 1, 0) - c(j, i, 0)
    dy = c(j + 1, i, 1) - c(j, i, 1)
    return pts, dx, dy


def _check_inside(cord, j, i, x, y):
    """src/euler2bar.f90:102-161: triangle k (1 or 2, 0 if neither) of element (j, i)
    holding (x, y) and its barycentric coordinates."""
    c = cord
    out_k = np.zeros(np.shape(x), dtype=np.int64)
    out1 = np.zeros(np.shape(x))
    out2 = np.zeros(np.shape(x))
    for k in (2, 1):
        if k == 1:
            x1, x2, x3 = c[j, i, 0], c[j + 1, i, 0], c[j, i + 1, 0]
            y1, y2, y3 = c[j, i, 1], c[j + 1, i, 1], c[j, i + 1, 1]
        else:
            x1, x2, x3 = c[j, i + 1, 0], c[j + 1, i, 0], c[j + 1, i + 1, 0]
            y1, y2, y3 = c[j, i + 1, 1], c[j + 1, i, 1], c[j + 1, i + 1, 1]
        det = ((x2 * y3 - y2 * x3) - (x1 * y3 - y1 * x3) + (x1 * y2 - y1 * x2))
        m11, m12, m13 = (x2 * y3 - y2 * x3) / det, (y2 - y3) / det, (x3 - x2) / det
        m21, m22, m23 = (x3 * y1 - y3 * x1) / det, (y3 - y1) / det, (x1 - x3) / det
        b1 = m11 + m12 * x + y * m13
        b2 = m21 + m22 * x + y * m23
        ok = (b1 >= 0.0) & (b2 >= 0.0) & (b1 + b2 <= 1.0)
        # triangle 1 is tried first in the reference: assign 2, then let 1 override
        out_k = np.where(ok, k, out_k)
        out1 = np.where(ok, b1, out1)
        out2 = np.where(ok, b2, out2)
    return out_k, out1, out2


def _layer_phase(p, i1, yy, default, hshift=None):
    """Phase of a marker at depth -yy in column i1 (1-based) from the zones' layers
    (src/init_marker.f90:86-117); hshift[zone][i1 - 1, m]: km added to interface m
    in that column (seeded heterogeneity), None for the deck's depths."""
    kph = np.array(default, copy=True)
    for nzn, zone in enumerate(p["zones"]):
        if zone["iph_col_trans"] == 1:
            raise NotImplementedError("interpolated layer depths (iph_col_trans)")
        sel = (i1 >= zone["ixtb1"]) & (i1 <= zone["ixtb2"])
        yyy = yy * (-1e-3)
        nl = zone["nph_layer"]
        kz = np.full(np.shape(yy), zone["iph_col"][nl - 1])
        for m in range(nl - 2, -1, -1):   # the first layer whose base lies below the marker
            h = zone["hc"][m] if hshift is None else zone["hc"][m] + hshift[nzn][i1 - 1, m]
            kz = np.where(yyy < h, zone["iph_col"][m], kz)
        kph = np.where(sel, kz, kph)
    return kph


def _seed_chain(seed, inc, a18, b18):
    """Start seeds of consecutive elements that draw 18 numbers each: the element adds
    inc (its i + j) to the seed left by the one before.  The map from one element's
    end seed to the next is x -> (a18 x + a18 inc + b18) mod m, so the end seeds are a
    linear recurrence, evaluated by a doubling scan."""
    m = _LCG_M
    n = len(inc)
    b = (a18 * (inc % m) + b18) % m
    b[0] = (a18 * ((seed + inc[0]) % m) + b18) % m
    t, apow = b.copy(), a18
    shift = 1
    while shift < n:
        t[shift:] = (t[shift:] + apow * t[:-shift]) % m
        apow = (apow * apow) % m
        shift *= 2
    ends = t
    starts = np.empty(n, dtype=np.int64)
    starts[0] = seed + inc[0]
    starts[1:] = ends[:-1] + inc[1:]
    return starts, int(ends[-1])


def init_markers(p, st, interface_noise=0.0, seed=0):
    """Markers of a new start, as src/init_marker.f90:1-169 places them in the serial
    build: 9 per element at the regular points, each moved by a random offset of up
    to a sixth of the element from the serial myrandom sequence (seed 0, advanced by
    i + j per element, i outer), the phase from the layer depths, the rectangular
    and diagonal inhomogeneities (newphase2marker, src/marker_data.f90:110-136), then
    the phase ratios and iphase of every element (count_phase_ratio,
    src/marker_data.f90:140-174).  A point that falls outside its element (possible
    only on its boundary) is drawn again, which shifts the rest of the sequence: the
    elements are evaluated in chunks, and a chunk is evaluated again from the element
    after such a redraw.  interface_noise > 0 moves each layer interface by a seeded
    uniform offset of +-interface_noise km per element column (a build-side
    extension; 0 places the layers as the reference does)."""
    nx, nz = p["nx"], p["nz"]
    ez, ex = nz - 1, nx - 1
    ne = ez * ex
    cord = st.cord
    e_all = np.arange(ne)
    J, I = e_all % ez, e_all // ez          # 0-based element indices, Fortran order
    mark_a1 = np.zeros((ne, 9))
    mark_a2 = np.zeros((ne, 9))
    ntri = np.zeros((ne, 9), dtype=np.int64)
    phase = np.zeros((ne, 9), dtype=np.int64)
    # affine map of 18 draws: L^18(s) = (A18 * (s mod m) + B18) mod m
    a18, b18 = 1, 0
    for _ in range(18):
        a18, b18 = (_LCG_A * a18) % _LCG_M, (_LCG_A * b18 + _LCG_C) % _LCG_M
    inc_ij = (I + 1 + J + 1).astype(np.int64)
    hshift = None
    if interface_noise:
        rng = np.random.default_rng(seed)
        hshift = [rng.uniform(-interface_noise, interface_noise, size=(ex, max(1, z["nph_layer"] - 1)))
                  for z in p["zones"]]
    CH = 1 << 16
    seed, e0 = 0, 0
    while e0 < ne:
        e1 = min(ne, e0 + CH)
        idx = np.arange(e0, e1)
        starts, end_seed = _seed_chain(seed, inc_ij[e0:e1], a18, b18)
        jj, ii = J[idx], I[idx]
        pts, dx, dy = _marker_points(cord, jj, ii)
        sd = starts.copy()
        rej = np.zeros(len(idx), dtype=bool)
        for l in range(9):
            sd = _lcg(sd); rx = sd * 1.0 / _LCG_M
            sd = _lcg(sd); ry = sd * 1.0 / _LCG_M
            rx = 0.5 - rx
            ry = 0.5 - ry
            xx = pts[l][0] + dx * rx / 3
            yy = pts[l][1] + dy * ry / 3
            k, b1, b2 = _check_inside(cord, jj, ii, xx, yy)
            rej |= k == 0
            mark_a1[idx, l], mark_a2[idx, l] = b1, b2
            ntri[idx, l] = 2 * (ez * ii + jj) + k
            phase[idx, l] = _layer_phase(p, ii + 1, yy, st.iphase[jj, ii], hshift)
        if not rej.any():
            seed, e0 = end_seed, e1
            continue
        # the first element with a redraw: serially, then go on after it
        q = int(np.argmax(rej))
        e = e0 + q
        j0, i0 = int(J[e]), int(I[e])
        pts1, dx1, dy1 = _marker_points(cord, np.array([j0]), np.array([i0]))
        s = int(starts[q])
        l = 0
        while l < 9:
            s = _lcg(s); rx = 0.5 - s * 1.0 / _LCG_M
            s = _lcg(s); ry = 0.5 - s * 1.0 / _LCG_M
            xx = pts1[l][0] + dx1 * rx / 3
            yy = pts1[l][1] + dy1 * ry / 3
            k, b1, b2 = _check_inside(cord, np.array([j0]), np.array([i0]), xx, yy)
            if k[0] == 0:
                continue
            mark_a1[e, l], mark_a2[e, l] = b1[0], b2[0]
            ntri[e, l] = 2 * (ez * i0 + j0) + k[0]
            phase[e, l] = _layer_phase(p, np.array([i0 + 1]), yy, np.array([st.iphase[j0, i0]]), hshift)[0]
            l += 1
        seed, e0 = s, e + 1
    mark_id = np.zeros((32, ez, ex), dtype=np.int32, order="F")
    ids = (np.arange(ne * 9) + 1).reshape(ne, 9)          # markers numbered element by element
    mark_id[:9] = ids.T.reshape(9, ez, ex, order="F")
    nmark = np.full((ez, ex), 9, dtype=np.int32, order="F")
    mphase = phase.ravel()
    # initial inhomogeneities (src/init_marker.f90:128-158)
    for ih in p.get("inhoms", []):
        if ih["inphase"] < 0:
            continue
        cells = []
        if ih["igeom"] == 0:
            cells = [(j, i) for i in range(ih["ix1"], ih["ix2"] + 1) for j in range(ih["iy1"], ih["iy2"] + 1)]
        elif ih["igeom"] in (3, 4):
            f32 = np.float32
            slope = f32(ih["iy2"] - ih["iy1"]) / f32(ih["ix2"] - ih["ix1"])
            for i in range(ih["ix1"], ih["ix2"] + 1):
                v = slope * f32(i - ih["ix1"])
                if ih["igeom"] == 3:
                    k = int(np.rint(v)) + ih["iy1"]   # nint: halves away from zero
                    if abs(v - np.floor(v) - 0.5) == 0.0:
                        k = int(np.floor(v) + 1) + ih["iy1"]
                    cells += [(k, i)]
                else:
                    cells += [(j, i) for j in range(int(np.floor(v)) + ih["iy1"], int(np.ceil(v)) + ih["iy1"] + 1)]
        for j, i in cells:
            e = (j - 1) + ez * (i - 1)
            mphase[e * 9:(e + 1) * 9] = ih["inphase"]
    # with surface processes resurface appends markers: room for max_markers of them
    # (src/marker_data.f90:27), ages 0 as init_marker leaves them
    cap = int(p["nz"]) * int(p["nx"]) * abi.MAX_MARKERS_PER_ELEM if float(p.get("topo_kappa", 0.0)) > 0.0 else None
    st.set_markers(mark_a1.ravel(), mark_a2.ravel(), ntri.ravel(), mphase.astype(np.int8),
                   np.ones(ne * 9, dtype=np.int8), mark_id, nmark, capacity=cap)
    # count_phase_ratio of every element (marker2elem, src/marker2elem.f90:60)
    nph = p["nphase"]
    npc = max(nph, int(mphase.max())) + 1
    counts = np.bincount(np.repeat(e_all, 9) * npc + mphase, minlength=ne * npc).reshape(ne, npc)
    ratio = counts[:, 1:nph + 1] / (9 * 1.0)
    st.phase_ratio[...] = ratio.T.reshape(nph, ez, ex, order="F")
    first_max = np.argmax(counts[:, 1:nph + 1], axis=1) + 1
    st.iphase[...] = first_max.reshape(ez, ex, order="F").astype(np.int32)


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


def perturb_layers(p, interface_noise, seed):
    """A copy of p whose layer interfaces are moved per zone by a seeded uniform
    offset of +-interface_noise km (one offset per interface; a build-side extension
    for heterogeneous benchmarks)."""
    import copy
    q = copy.deepcopy(p)
    rng = np.random.default_rng(seed)
    for z in q["zones"]:
        z["hc"] = [h + float(rng.uniform(-interface_noise, interface_noise)) for h in z["hc"]]
    return q


def inhom_aps(p, st):
    """The initial plastic strain of the inhomogeneities (init_phase,
    src/init_phase.f90:21-47): xinitaps in rectangles (igeom 0) and in the 45-degree
    bands of igeom 4 (single-precision slope, floor and ceiling, as the reference)."""
    for ih in p.get("inhoms", []):
        ix1, ix2, iy1, iy2 = ih["ix1"], ih["ix2"], ih["iy1"], ih["iy2"]
        if ih["igeom"] == 0:
            st.aps[iy1 - 1:iy2, ix1 - 1:ix2] = ih["xinitaps"]
        elif ih["igeom"] == 4:
            slope = np.float32(np.float32(iy2 - iy1) / np.float32(ix2 - ix1))
            for j in range(ix1, ix2 + 1):
                v = np.float32(slope * np.float32(j - ix1))
                k1, k2 = int(np.floor(v)) + iy1, int(np.ceil(v)) + iy1
                st.aps[k1 - 1:k2, j - 1] = ih["xinitaps"]


def build_box(p, interface_noise=0.0, seed=0, markers=True, coh_noise=0.0):
    """State and parameter block of a new start (before the device-side set-up).

    markers: place the reference's markers (init_markers) and take the phases from
    them, as setflac does; otherwise phases come from the 9 regular points of each
    element (layer_phases).  interface_noise: seeded per-column offsets of the layer
    interfaces, +-interface_noise km.  coh_noise: a seeded per-element cohesion factor 1 + coh_noise*U,
    U uniform in [-1, 1] (a build-side extension; 0: none)."""
    nx, nz = p["nx"], p["nz"]
    st = State(nx, nz, p["nphase"])
    xm = mesh1(p["x0"], p["x0"] + p["rxbo"], nx, p["nelz_x"], p["sizez_x"])
    st.cord[:, :, 0] = xm[None, :]
    zm = mesh1(p["z0"], p["z0"] + p["rzbo"], nz, p["nelz_y"], p["sizez_y"])
    st.cord[:, :, 1] = zm[:, None]
    inhom_aps(p, st)
    if markers:
        init_markers(p, st, interface_noise, seed)
    else:
        layer_phases(p, st, interface_noise, seed)
    if coh_noise:
        rng = np.random.default_rng(seed + 1)
        st.coh_scale = np.asfortranarray(
            (1.0 + coh_noise * rng.uniform(-1.0, 1.0, size=(nz - 1, nx - 1))).astype(np.float32))
    ivis = int(any(p["phases"][k - 1]["irheol"] in (3,) or p["phases"][k - 1]["irheol"] >= 11
                   for k in np.unique(st.iphase)))
    init_temp(p, st)
    pisos = init_stress(p, st)
    vbc, nopbmax = init_bc(p, st)
    st.temp0[...] = st.temp
    st.e2sr[...] = 1e-16
    st.se2sr[...] = 1e-16
    c = st.clock
    c.nloop, c.nsrate, c.time, c.dt, c.dtavg = 0, -1, 0.0, 0.0, 0.0
    c.dt_elastic, c.dt_maxwell = 0.0, 0.0
    gp = inp.to_gf_params(p, vbc=vbc, ivis_present=ivis, pisos=pisos)
    gp.nopbmax = nopbmax
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
