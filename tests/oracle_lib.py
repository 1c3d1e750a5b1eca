"""Loader of the CPU oracle (oracle/liboracle.so) for tests: the checker only."""
import ctypes as C
import os

from geoflac_amd import abi, inp, state

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(ROOT, "oracle", "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        _LIB = C.CDLL(path)
        _LIB.gfo_eff_visc.restype = C.c_double
        _LIB.gfo_eff_dens.restype = C.c_double
    return _LIB


def golden_params(case):
    return inp.read_params_txt(os.path.join(GOLDEN, case + "_params.txt"))


def golden_state(case, step):
    return state.load_npz(os.path.join(GOLDEN, "%s_s%d.npz" % (case, step)))


def step(p, st, nsteps):
    h = st.host_arrays()
    return lib().gfo_step(C.byref(p), C.byref(h), C.byref(st.clock), int(nsteps))


def call(name, p, st):
    h = st.host_arrays()
    return getattr(lib(), name)(C.byref(p), C.byref(h), C.byref(st.clock))
