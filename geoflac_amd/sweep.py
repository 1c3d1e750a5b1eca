"""Parameter sweeps: many independent models advanced together (gf_batch_* of the C-ABI).

`variant` derives a model from a base .inp parameter set (mesh size, friction,
cohesion, extension velocity), `sweep_params` draws a seeded sweep, and `Batch`
runs a list of device models through the batched launches.
"""
import copy
import ctypes as C

import numpy as np

from geoflac_amd import abi, gpu, model


def variant(p, nelx, nelz, fric1=None, coh1=None, vside=None):
    """A copy of parameter set p with a uniform nelx x nelz mesh of the same domain,
    optionally another initial friction angle (degrees) and cohesion (Pa) for every
    phase and another side-wall velocity (m/s, each wall)."""
    q = copy.deepcopy(p)
    q["nelx"], q["nelz"] = int(nelx), int(nelz)
    q["nx"], q["nz"] = int(nelx) + 1, int(nelz) + 1
    q["nzonx"], q["nelz_x"], q["sizez_x"] = 1, [int(nelx)], [1.0]
    q["nzony"], q["nelz_y"], q["sizez_y"] = 1, [int(nelz)], [1.0]
    for bc in q["bcs"]:
        if bc["nofside"] in (1, 3):
            bc["nbc1"], bc["nbc2"] = 1, q["nz"]
            if vside is not None and bc["nbc"] == 10:
                bc["bca"] = -abs(vside) if bc["nofside"] == 1 else abs(vside)
        else:
            bc["nbc1"], bc["nbc2"] = 1, q["nx"]
    for z in q["zones"]:
        z["ixtb1"], z["ixtb2"] = 1, q["nx"]
    for ph in q["phases"]:
        if fric1 is not None:
            ph["fric1"] = float(fric1)
        if coh1 is not None:
            ph["cohesion1"] = float(coh1)
    return q


def sweep_params(base, n=256, seed=0, nx_range=(100, 600), nz_range=(40, 150), fric=(10.0, 35.0),
                 coh=(10e6, 60e6), vel_cm_yr=(0.2, 3.0)):
    """n seeded variants: mesh nx, nz uniform in the element ranges, friction, cohesion
    and side velocity uniform in theirs."""
    rng = np.random.default_rng(seed)
    sec_year = 3.1558e7
    out = []
    for _ in range(n):
        nelx = int(rng.integers(nx_range[0], nx_range[1] + 1))
        nelz = int(rng.integers(nz_range[0], nz_range[1] + 1))
        v = rng.uniform(*vel_cm_yr) * 0.01 / sec_year
        out.append(variant(base, nelx, nelz, rng.uniform(*fric), rng.uniform(*coh), v))
    return out


def make_model(p, precision=abi.PREC_FP64):
    """Host new start and device set-up of one model: (params block, state, Model)."""
    gp, st = model.build_box(p)
    m = gpu.Model(gp, st, precision=precision)
    model.device_setup(m)
    return gp, st, m


class Batch:
    """gf_batch over a list of gpu.Model objects (they stay owned by the caller)."""

    def __init__(self, models):
        L = gpu.lib()
        self.models = list(models)
        arr = (C.c_void_p * len(self.models))(*[m._ctx for m in self.models])
        self._b = C.c_void_p()
        rc = L.gf_batch_create(arr, len(self.models), C.byref(self._b))
        if rc != 0:
            raise RuntimeError("gf_batch_create failed (%d): %s" % (rc, L.gf_batch_last_error(self._b)))

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError("batch call failed (%d): %s" % (rc, gpu.lib().gf_batch_last_error(self._b)))
        return rc

    def step(self, nsteps):
        self._check(gpu.lib().gf_batch_step(self._b, int(nsteps)))
        for m in self.models:
            gpu.lib().gf_get_clock(m._ctx, C.byref(m.state.clock))

    def set_graph(self, on=True):
        self._check(gpu.lib().gf_batch_set_graph(self._b, int(on)))

    def autotune(self, steps=2):
        v = (C.c_int32 * 3)()
        self._check(gpu.lib().gf_batch_autotune(self._b, int(steps), v))
        return list(v)

    def close(self):
        if self._b:
            gpu.lib().gf_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
