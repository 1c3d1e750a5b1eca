"""Host-side model state in the reference's own array layout.

Arrays are numpy arrays in Fortran (column-major) order with the shapes of
src/arrays.f90:59-100, so `state.cord[j, i, d]` is the reference's
`cord(j+1, i+1, d+1)`.  The C-ABI receives raw pointers to these buffers
(`host_arrays()`), exactly as a Fortran caller would pass its module arrays.
"""
import ctypes as C
import os

import numpy as np

from . import abi

NODE2 = ("cord", "vel", "force", "bc")
NODE1 = ("temp", "temp0", "amass", "rmass")
ELEM1 = ("aps", "visn", "e2sr", "sshrheat", "shrheat", "source")
MARKERS = ("mark_a1", "mark_a2", "mark_ntriag", "mark_phase", "mark_dead", "mark_id_elem", "nmark_elem")
# surface processes (topo_kappa > 0, src/arrays.f90:18): per surface node / element
TOPO = ("dtopo", "dhacc", "dhacc_correction")


class State:
    """Complete set of arrays the explicit step reads or writes."""

    FIELDS = [n for n, _ in abi.HOST_FIELDS]

    def __init__(self, nx, nz, nphase):
        self.nx, self.nz, self.nphase = int(nx), int(nz), int(nphase)
        ex, ez = self.nx - 1, self.nz - 1
        f8 = dict(dtype=np.float64, order="F")
        for n in NODE2:
            setattr(self, n, np.zeros((self.nz, self.nx, 2), **f8))
        for n in NODE1:
            setattr(self, n, np.zeros((self.nz, self.nx), **f8))
        self.ncod = np.zeros((self.nz, self.nx, 2), dtype=np.int32, order="F")
        self.stress0 = np.zeros((ez, ex, 4, 4), **f8)
        self.area = np.zeros((ez, ex, 4), **f8)
        self.dvol = np.zeros((ez, ex, 4), **f8)
        self.strain = np.zeros((ez, ex, 3), **f8)
        self.se2sr = np.zeros((ez, ex, 3), **f8)
        for n in ELEM1:
            setattr(self, n, np.zeros((ez, ex), **f8))
        self.iphase = np.ones((ez, ex), dtype=np.int32, order="F")
        self.phase_ratio = np.zeros((self.nphase, ez, ex), **f8)
        self.clock = abi.GfClock()
        # stress boundary segments (src/init_bc.f90:226-280): node columns/rows, stresses
        rows = 2 * (ez + ex)
        self.nopbou = np.zeros((rows, 4), dtype=np.int32, order="F")
        self.bcstress = np.zeros((rows, 3), **f8)
        # per-element cohesion factor (seeded heterogeneity; no reference counterpart)
        self.coh_scale = None
        # topography change and its accumulations (diff_topo, src/fl_move.f90:138-315)
        self.dtopo = np.zeros(self.nx, **f8)
        self.dhacc = np.zeros(self.nx - 1, **f8)
        self.dhacc_correction = np.zeros(self.nx, **f8)
        # markers (module marker_data, src/marker_data.f90:1-20), None until set; the
        # per-marker arrays may hold more entries than nmarkers (room for resurface)
        self.nmarkers = 0
        for n in MARKERS:
            setattr(self, n, None)
        self.mark_age = None

    def set_markers(self, a1, a2, ntriag, phase, dead, id_elem, nmark_elem, capacity=None, age=None):
        """Marker arrays in the reference's layout: per marker a1, a2 (barycentric),
        ntriag (triangle index), phase, dead (1 alive); per element id_elem[32, ez, ex]
        (1-based marker numbers) and nmark_elem[ez, ex].  capacity: entries the
        per-marker arrays hold (at least len(a1); the rest zero), for the markers that
        resurface appends; age: per-marker ages (kept only when given)."""
        ez, ex = self.nz - 1, self.nx - 1
        n = int(len(a1))
        cap = max(n, int(capacity or 0))
        self.nmarkers = n

        def pad(a, dt):
            out = np.zeros(cap, dtype=dt)
            out[:n] = np.asarray(a, dtype=dt)[:n]
            return out
        self.mark_a1 = pad(a1, np.float64)
        self.mark_a2 = pad(a2, np.float64)
        self.mark_ntriag = pad(ntriag, np.int32)
        self.mark_phase = pad(phase, np.int8)
        self.mark_dead = pad(dead, np.int8)
        self.mark_age = None if age is None and capacity is None else pad(age if age is not None else np.zeros(n), np.float64)
        self.mark_id_elem = np.asfortranarray(np.asarray(id_elem, dtype=np.int32).reshape(
            (abi.MAX_MARKERS_PER_ELEM, ez, ex), order="F"))
        self.nmark_elem = np.asfortranarray(np.asarray(nmark_elem, dtype=np.int32).reshape((ez, ex), order="F"))

    @property
    def has_markers(self):
        return self.mark_a1 is not None

    def marker_arrays(self):
        m = abi.GfMarkerArrays()
        m.nmarkers = self.nmarkers
        m.capacity = len(self.mark_a1)
        for name, ptype in abi.MARKER_FIELDS:
            a = getattr(self, name)
            assert a.flags.f_contiguous
            setattr(m, name, a.ctypes.data_as(ptype))
        if self.mark_age is not None:
            m.mark_age = self.mark_age.ctypes.data_as(abi._DP)
        return m

    def live_markers(self, name):
        """The first nmarkers entries of a per-marker array."""
        return getattr(self, name)[:self.nmarkers]

    @property
    def nelem(self):
        return (self.nx - 1) * (self.nz - 1)

    def host_arrays(self):
        h = abi.GfHostArrays()
        for name, ptype in abi.HOST_FIELDS:
            a = getattr(self, name)
            assert a.flags.f_contiguous
            setattr(h, name, a.ctypes.data_as(ptype))
        h.nopbou = self.nopbou.ctypes.data_as(abi._IP)
        h.bcstress = self.bcstress.ctypes.data_as(abi._DP)
        if self.coh_scale is not None:
            assert self.coh_scale.flags.f_contiguous and self.coh_scale.dtype == np.float32
            h.coh_scale = self.coh_scale.ctypes.data_as(abi._FP)
        for name in TOPO:
            setattr(h, name, getattr(self, name).ctypes.data_as(abi._DP))
        return h

    def copy(self):
        s = State(self.nx, self.nz, self.nphase)
        for n in self.FIELDS:
            getattr(s, n)[...] = getattr(self, n)
        C.memmove(C.addressof(s.clock), C.addressof(self.clock), C.sizeof(abi.GfClock))
        s.nopbou[...] = self.nopbou
        s.bcstress[...] = self.bcstress
        if self.coh_scale is not None:
            s.coh_scale = self.coh_scale.copy(order="F")
        for n in TOPO:
            getattr(s, n)[...] = getattr(self, n)
        if self.has_markers:
            arrs = [getattr(self, n).copy(order="F") for n in MARKERS]
            s.set_markers(*arrs, capacity=len(self.mark_a1),
                          age=None if self.mark_age is None else self.mark_age.copy())
            s.nmarkers = self.nmarkers
        return s

    def clock_dict(self):
        return {k: getattr(self.clock, k) for k, _ in abi.GfClock._fields_}


def _shape_for(state, name):
    return getattr(state, name).shape


def load_dump(path, nphase=None):
    """Read one state directory written by the reference harness
    (oracle/ref_harness.f90: raw little-endian column-major arrays + scalars.txt)."""
    scal = {}
    with open(os.path.join(path, "scalars.txt")) as f:
        for line in f:
            k, v = line.split()
            scal[k] = float(v) if k not in ("nloop", "nsrate") else int(v)
    pr = np.fromfile(os.path.join(path, "phase_ratio.f64"))
    ip = np.fromfile(os.path.join(path, "iphase.i32"), dtype=np.int32)
    nph = pr.size // ip.size
    # the harness writes params.txt (with nx, nz) one directory up
    ptxt = os.path.join(os.path.dirname(path.rstrip("/")), "params.txt")
    with open(ptxt) as f:
        for line in f:
            if line.startswith("nx_nz_nphase"):
                _, nx, nz, _ = line.split()
                nx, nz = int(nx), int(nz)
    st = State(nx, nz, nph if nphase is None else nphase)
    for name in State.FIELDS:
        dt = np.int32 if name in ("ncod", "iphase") else np.float64
        ext = ".i32" if dt is np.int32 else ".f64"
        arr = np.fromfile(os.path.join(path, name + ext), dtype=dt)
        tgt = getattr(st, name)
        tgt[...] = arr.reshape(tgt.shape, order="F")
    if os.path.exists(os.path.join(path, "nopbou.i32")):
        st.nopbou[...] = np.fromfile(os.path.join(path, "nopbou.i32"), dtype=np.int32).reshape(st.nopbou.shape, order="F")
        st.bcstress[...] = np.fromfile(os.path.join(path, "bcstress.f64")).reshape(st.bcstress.shape, order="F")
    for name in TOPO:
        fp = os.path.join(path, name + ".f64")
        if os.path.exists(fp):
            getattr(st, name)[...] = np.fromfile(fp)
    if os.path.exists(os.path.join(path, "mark_a1.f64")):
        rd = lambda n, dt: np.fromfile(os.path.join(path, n + (".f64" if dt is np.float64 else ".i32")), dtype=dt)
        age = rd("mark_age", np.float64) if os.path.exists(os.path.join(path, "mark_age.f64")) else None
        st.set_markers(rd("mark_a1", np.float64), rd("mark_a2", np.float64), rd("mark_ntriag", np.int32),
                       rd("mark_phase", np.int32).astype(np.int8), rd("mark_dead", np.int32).astype(np.int8),
                       rd("mark_id_elem", np.int32), rd("nmark_elem", np.int32), age=age)
    elif os.path.exists(os.path.join(path, "mark_phase.i32")):
        st.mark_phase_only = np.fromfile(os.path.join(path, "mark_phase.i32"), dtype=np.int32).astype(np.int8)
    c = st.clock
    c.nloop, c.nsrate = scal["nloop"], scal["nsrate"]
    c.time, c.dt, c.dtavg = scal["time"], scal["dt"], scal["dtavg"]
    c.dt_elastic, c.dt_maxwell = scal["dt_elastic"], scal["dt_maxwell"]
    st.scalars = scal
    return st


def save_npz(state, path, fields=None):
    fields = state.FIELDS if fields is None else fields
    d = {n: getattr(state, n) for n in fields if n != "mark_phase" and n not in MARKERS and n not in TOPO}
    if state.has_markers and (fields is state.FIELDS or "mark_a1" in fields):
        d.update({n: (state.live_markers(n) if n not in ("mark_id_elem", "nmark_elem") else getattr(state, n))
                  for n in MARKERS})
        if state.mark_age is not None:
            d["mark_age"] = state.live_markers("mark_age")
    for n in TOPO:
        if n in fields:
            d[n] = getattr(state, n)
    if fields is state.FIELDS:
        d["nopbou"], d["bcstress"] = state.nopbou, state.bcstress
    elif "mark_phase" in fields and "mark_a1" not in fields:
        mp = state.live_markers("mark_phase") if state.has_markers else getattr(state, "mark_phase_only", None)
        if mp is not None:
            d["mark_phase"] = mp
    d["_dims"] = np.array([state.nx, state.nz, state.nphase], dtype=np.int64)
    d["_clock_i"] = np.array([state.clock.nloop, state.clock.nsrate], dtype=np.int64)
    d["_clock_f"] = np.array([state.clock.time, state.clock.dt, state.clock.dtavg,
                              state.clock.dt_elastic, state.clock.dt_maxwell])
    np.savez_compressed(path, **d)


def load_npz(path):
    z = np.load(path, allow_pickle=False)
    nx, nz, nph = (int(v) for v in z["_dims"])
    st = State(nx, nz, nph)
    for n in State.FIELDS:
        if n in z.files:
            getattr(st, n)[...] = z[n]
    if "nopbou" in z.files:
        st.nopbou[...] = z["nopbou"]
        st.bcstress[...] = z["bcstress"]
    for n in TOPO:
        if n in z.files:
            getattr(st, n)[...] = z[n]
    if "mark_a1" in z.files:
        st.set_markers(*(z[n] for n in MARKERS), age=z["mark_age"] if "mark_age" in z.files else None)
    elif "mark_phase" in z.files:
        st.mark_phase_only = z["mark_phase"]
    st.clock.nloop, st.clock.nsrate = (int(v) for v in z["_clock_i"])
    (st.clock.time, st.clock.dt, st.clock.dtavg,
     st.clock.dt_elastic, st.clock.dt_maxwell) = (float(v) for v in z["_clock_f"])
    st.present = set(z.files)
    return st


def rel_l2(a, b):
    """Relative L2 difference ||a-b|| / ||b|| (0 when both are zero)."""
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    nd = np.linalg.norm(a - b)
    if nb == 0.0:
        return 0.0 if nd == 0.0 else float("inf")
    return float(nd / nb)
