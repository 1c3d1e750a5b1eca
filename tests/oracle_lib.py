"""Loader of the CPU oracle (oracle/liboracle.so) for tests: the checker only."""
import ctypes as C
import os

from geoflac_amd import abi, inp, state

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
_LIBS = {}


def lib(cr=False):
    """The oracle with the host libm (cr=False: what the reference calls) or with the
    correctly rounded functions the device uses (cr=True)."""
    if cr not in _LIBS:
        name = "liboracle_cr.so" if cr else "liboracle.so"
        path = os.path.join(ROOT, "oracle", name)
        if not os.path.exists(path):
            raise RuntimeError("oracle/%s is not built (make -C oracle)" % name)
        L = C.CDLL(path)
        L.gfo_eff_visc.restype = C.c_double
        L.gfo_eff_dens.restype = C.c_double
        L.gfo_math.restype = C.c_double
        L.gfo_math.argtypes = [C.c_int, C.c_double, C.c_double]
        L.gfo_math_fast.restype = C.c_long
        L.gfo_math_fast.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_long, C.POINTER(C.c_long)]
        _LIBS[cr] = L
    return _LIBS[cr]


def golden_params(case):
    return inp.read_params_txt(os.path.join(GOLDEN, case + "_params.txt"))


def golden_state(case, step):
    return state.load_npz(os.path.join(GOLDEN, "%s_s%d.npz" % (case, step)))


def _bind_markers(L, st):
    """change_phase in the oracle runs on the state's markers (none: it does nothing)."""
    if st.has_markers:
        m = st.marker_arrays()
        L.gfo_bind_markers(C.byref(m))
        return m
    L.gfo_bind_markers(None)
    return None


def step(p, st, nsteps, cr=False):
    h = st.host_arrays()
    L = lib(cr)
    keep = _bind_markers(L, st)  # (kept alive over the call)
    rc = L.gfo_step(C.byref(p), C.byref(h), C.byref(st.clock), int(nsteps))
    L.gfo_bind_markers(None)
    if keep is not None:
        st.nmarkers = keep.nmarkers  # (resurface appends markers)
    return rc


def call(name, p, st, cr=False):
    h = st.host_arrays()
    L = lib(cr)
    keep = _bind_markers(L, st)
    rc = getattr(L, name)(C.byref(p), C.byref(h), C.byref(st.clock))
    L.gfo_bind_markers(None)
    if keep is not None:
        st.nmarkers = keep.nmarkers
    return rc


def golden_full(case, step, marker_room=False):
    """The complete state after `step` reference steps: the fields the step never
    changes from the step-0 golden, the others (markers included, when the golden
    holds them all) from the step-N golden.  marker_room: room for the markers
    resurface appends, max_markers entries (nz nx 32, src/marker_data.f90:27)."""
    s0 = golden_state(case, 0)
    sn = s0 if step == 0 else golden_state(case, step)
    if step:
        for f in sn.present:
            if f.startswith("_") or f in state.MARKERS or f == "mark_age":
                continue
            if f == "mark_phase":
                s0.mark_phase[...] = sn.mark_phase_only
            else:
                getattr(s0, f)[...] = getattr(sn, f)
        C.memmove(C.addressof(s0.clock), C.addressof(sn.clock), C.sizeof(abi.GfClock))
    src = sn if (step and sn.has_markers) else s0
    if src.has_markers:
        n = src.nmarkers
        arrs = [getattr(src, k)[:n] if k not in ("mark_id_elem", "nmark_elem") else getattr(src, k)
                for k in state.MARKERS]
        s0.set_markers(*arrs, capacity=s0.nz * s0.nx * abi.MAX_MARKERS_PER_ELEM if marker_room else None,
                       age=None if src.mark_age is None else src.mark_age[:n])
    return s0


def itest_mesh(p, st, mode_rem, dx_rem, angle_rem, ratio_mantle_mzone=1.0):
    """The oracle's itest_mesh (src/rem_test.f90:9-139) on a state: (need, angle, value)."""
    L = lib()
    f = L.gfo_itest_mesh
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_int),
                  C.POINTER(C.c_double), C.POINTER(C.c_double)]
    h = st.host_arrays()
    need, amin, val = C.c_int(), C.c_double(), C.c_double()
    rc = f(C.byref(p), C.byref(h), int(mode_rem), float(dx_rem), float(angle_rem), float(ratio_mantle_mzone),
           C.byref(need), C.byref(amin), C.byref(val))
    assert rc == 0
    return need.value, amin.value, val.value
