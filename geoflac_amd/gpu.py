"""ctypes binding of libgfhip.so (include/gf_abi.h) for Python hosts and tests.

This is a thin caller of the C-ABI: every computation runs in the HIP kernels of
the library.  Importing it without the built library, or calling it without a
HIP device, raises at once; there is no CPU fallback.
"""
import ctypes as C
import os

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgfhip.so")
_lib = None

EXPORTS = [
    "gf_create", "gf_destroy", "gf_last_error", "gf_error_site", "gf_bind_host", "gf_bind_markers", "gf_upload",
    "gf_download", "gf_set_clock", "gf_get_clock", "gf_step", "gf_flac", "gf_fl_therm",
    "gf_fl_rheol", "gf_fl_node", "gf_fl_move", "gf_rmasses", "gf_dt_mass", "gf_change_phase", "gf_bc_update",
    "gf_init_areas", "gf_outflac", "gf_saveflac",
    "gf_init_visc", "gf_therm_stats", "gf_set_therm_passes", "gf_sync",
    "gf_set_timing", "gf_kernel_time", "gf_math_check",
    "gf_surrogate_sample", "gf_surrogate_train", "gf_surrogate_enable", "gf_surrogate_stats",
    "gf_surrogate_weights", "gf_mlp_grad", "gf_mlp_infer", "gf_mlp_grad_noalloc",
    "gf_batch_create", "gf_batch_destroy", "gf_batch_last_error", "gf_batch_step", "gf_batch_set_graph",
    "gf_batch_autotune", "gf_stream_probe", "gf_marker_count", "gf_itest_mesh",
]
MLP_NPARAM = 4737


class GfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("libgfhip status %d: %s" % (code, msg))
        self.code = code


def lib():
    """Load libgfhip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libgfhip.so is not built; run geoflac_amd/build.py (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.gf_create.argtypes = [C.POINTER(abi.GfParams), C.c_int32, C.POINTER(vp)]
        L.gf_destroy.argtypes = [vp]
        L.gf_destroy.restype = None
        L.gf_last_error.argtypes = [vp]
        L.gf_last_error.restype = C.c_char_p
        L.gf_error_site.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.gf_bind_host.argtypes = [vp, C.POINTER(abi.GfHostArrays)]
        L.gf_bind_markers.argtypes = [vp, C.POINTER(abi.GfMarkerArrays)]
        L.gf_upload.argtypes = [vp, C.c_uint32]
        L.gf_download.argtypes = [vp, C.c_uint32]
        L.gf_marker_count.argtypes = [vp, C.POINTER(C.c_int32)]
        L.gf_itest_mesh.argtypes = [vp, C.c_int32, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_int32),
                                    C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.gf_set_clock.argtypes = [vp, C.POINTER(abi.GfClock)]
        L.gf_get_clock.argtypes = [vp, C.POINTER(abi.GfClock)]
        L.gf_step.argtypes = [vp, C.c_int32, C.POINTER(abi.GfClock)]
        L.gf_flac.argtypes = [vp, C.POINTER(abi.GfClock)]
        for n in ("gf_fl_therm", "gf_fl_rheol", "gf_fl_node", "gf_fl_move", "gf_rmasses", "gf_dt_mass",
                  "gf_change_phase", "gf_bc_update"):
            getattr(L, n).argtypes = [vp, C.POINTER(abi.GfClock)]
        L.gf_init_areas.argtypes = [vp]
        for n in ("gf_outflac", "gf_saveflac"):
            getattr(L, n).argtypes = [vp, C.POINTER(abi.GfClock), C.c_char_p, C.POINTER(abi.GfOutputOpts)]
        L.gf_init_visc.argtypes = [vp]
        L.gf_therm_stats.argtypes = [vp, C.POINTER(C.c_int64)]
        L.gf_sync.argtypes = [vp]
        L.gf_set_therm_passes.argtypes = [vp, C.c_int32]
        L.gf_set_timing.argtypes = [vp, C.c_int32]
        L.gf_kernel_time.argtypes = [vp, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        dp = C.POINTER(C.c_double)
        L.gf_math_check.argtypes = [C.c_int32, dp, dp, dp, C.c_int64]
        fp = C.POINTER(C.c_float)
        L.gf_surrogate_sample.argtypes = [vp, C.c_int64, C.POINTER(C.c_int64)]
        L.gf_surrogate_train.argtypes = [vp, C.c_int32, C.c_int32, C.c_float, C.c_uint64, fp, fp]
        L.gf_surrogate_enable.argtypes = [vp, C.c_int32, C.c_float, C.c_int32]
        L.gf_surrogate_stats.argtypes = [vp, dp]
        L.gf_surrogate_weights.argtypes = [vp, fp, fp, C.c_int32]
        L.gf_mlp_grad.argtypes = [fp, fp, C.c_int64, fp, fp, fp]
        L.gf_mlp_infer.argtypes = [fp, C.c_int64, fp, fp]
        L.gf_mlp_grad_noalloc.argtypes = [vp, C.c_int64, vp, vp, vp]
        L.gf_batch_create.argtypes = [C.POINTER(vp), C.c_int32, C.POINTER(vp)]
        L.gf_batch_destroy.argtypes = [vp]
        L.gf_batch_destroy.restype = None
        L.gf_batch_last_error.argtypes = [vp]
        L.gf_batch_last_error.restype = C.c_char_p
        L.gf_batch_step.argtypes = [vp, C.c_int32]
        L.gf_batch_set_graph.argtypes = [vp, C.c_int32]
        L.gf_batch_autotune.argtypes = [vp, C.c_int32, C.POINTER(C.c_int32)]
        L.gf_stream_probe.argtypes = [C.c_int64, C.c_int32, C.POINTER(C.c_double)]
        _lib = L
    return _lib


class Model:
    """One model on the device, bound to a geoflac_amd.state.State's host arrays."""

    def __init__(self, params, state, precision=abi.PREC_FP64, upload=True):
        self.params = params
        self.state = state
        self._ctx = C.c_void_p()
        L = lib()
        rc = L.gf_create(C.byref(params), int(precision), C.byref(self._ctx))
        if rc != 0:
            raise GfError(rc, "gf_create rejected the parameters")
        self._host = state.host_arrays()
        self._check(L.gf_bind_host(self._ctx, C.byref(self._host)))
        self._marks = None
        if state.has_markers:
            self._marks = state.marker_arrays()
            self._check(L.gf_bind_markers(self._ctx, C.byref(self._marks)))
        if upload:
            self.upload()

    def _check(self, rc):
        if rc != 0:
            msg = lib().gf_last_error(self._ctx)
            raise GfError(rc, msg.decode() if msg else "")
        return rc

    def _mask(self, mask):
        """GF_ALL also carries the markers and the cohesion factors the state holds."""
        if mask == abi.GF_ALL:
            if self.state.has_markers:
                mask |= abi.GF_MARKERS
            if self.state.coh_scale is not None:
                mask |= abi.GF_COH_SCALE
            if self.params.nystressbc:
                mask |= abi.GF_STRESSBC
            if self.params.topo_kappa > 0.0:
                mask |= abi.GF_TOPO
        return mask

    def upload(self, mask=abi.GF_ALL):
        self._check(lib().gf_upload(self._ctx, self._mask(mask)))
        self._check(lib().gf_set_clock(self._ctx, C.byref(self.state.clock)))

    def set_clock(self):
        self._check(lib().gf_set_clock(self._ctx, C.byref(self.state.clock)))

    def download(self, mask=abi.GF_ALL):
        mask = self._mask(mask)
        self._check(lib().gf_download(self._ctx, mask & ~(abi.GF_COH_SCALE | abi.GF_STRESSBC)))
        self._check(lib().gf_get_clock(self._ctx, C.byref(self.state.clock)))
        if mask & abi.GF_MARKERS:  # (resurface appends markers with surface processes)
            self.state.nmarkers = self.marker_count()

    def marker_count(self):
        n = C.c_int32()
        self._check(lib().gf_marker_count(self._ctx, C.byref(n)))
        return n.value

    def itest_mesh(self, mode_rem, dx_rem, angle_rem, ratio_mantle_mzone=1.0):
        """The remeshing test itest_mesh (src/rem_test.f90:9-139) on the device's
        coordinates: (need_remeshing, smallest angle in degrees, printed value)."""
        need, amin, val = C.c_int32(), C.c_double(), C.c_double()
        self._check(lib().gf_itest_mesh(self._ctx, int(mode_rem), float(dx_rem), float(angle_rem),
                                        float(ratio_mantle_mzone), C.byref(need), C.byref(amin), C.byref(val)))
        return need.value, amin.value, val.value

    def step(self, nsteps):
        """nsteps iterations of the reference's inner loop (flac; nloop+=1; time+=dt)."""
        return self._check(lib().gf_step(self._ctx, int(nsteps), C.byref(self.state.clock)))

    def call(self, name):
        """One reference subroutine: 'fl_therm', 'change_phase', 'fl_rheol', 'fl_node',
        'fl_move', 'rmasses', 'dt_mass', 'flac'."""
        fn = getattr(lib(), "gf_" + name)
        return self._check(fn(self._ctx, C.byref(self.state.clock)))

    def init_areas(self):
        return self._check(lib().gf_init_areas(self._ctx))

    @staticmethod
    def output_opts(io_flags=None, lastout=0, hs=0.0, **arrays):
        """gf_output_opts: the 19 frame switches (all on by default), lastout, hs and the
        host-owned arrays by name (abi.OUTPUT_ARRAYS; missing ones are written as zeros).
        The arrays must stay alive while the options are used."""
        import numpy as np
        o = abi.GfOutputOpts()
        for n, v in zip(abi.IO_FLAGS, io_flags if io_flags is not None else [1] * len(abi.IO_FLAGS)):
            setattr(o, n, int(v))
        o.lastout, o.hs = int(lastout), float(hs)
        keep = []
        for n, a in arrays.items():
            if n not in abi.OUTPUT_ARRAYS:
                raise KeyError(n)
            if a is None:
                continue
            a = np.asfortranarray(a, dtype=np.float64)
            keep.append(a)
            setattr(o, n, a.ctypes.data_as(C.POINTER(C.c_double)))
        o._keep = keep
        return o

    def outflac(self, directory, opts):
        """outflac's frame files of the current device state (src/outflac.f90), into directory."""
        return self._check(lib().gf_outflac(self._ctx, C.byref(self.state.clock), str(directory).encode(),
                                            C.byref(opts)))

    def saveflac(self, directory, opts):
        """saveflac's restart files of the current device state (src/saveflac.f90), into directory."""
        return self._check(lib().gf_saveflac(self._ctx, C.byref(self.state.clock), str(directory).encode(),
                                             C.byref(opts)))

    def init_visc(self):
        return self._check(lib().gf_init_visc(self._ctx))

    def therm_stats(self):
        """(unproven thermal calls, extra passes, border mismatches, max mismatch ulp)
        since the last query."""
        v = (C.c_int64 * 4)()
        self._check(lib().gf_therm_stats(self._ctx, v))
        return tuple(int(x) for x in v)

    def set_therm_passes(self, n):
        """Most thermal fixed-point passes per call; an unproven call raises GfError 1004."""
        return self._check(lib().gf_set_therm_passes(self._ctx, int(n)))

    def therm_unconverged(self):
        return self.therm_stats()[0]

    # -------- viscosity surrogate
    def surrogate_sample(self, capacity):
        """Append one sample per element of the current state; returns the count."""
        n = C.c_int64()
        self._check(lib().gf_surrogate_sample(self._ctx, int(capacity), C.byref(n)))
        return n.value

    def surrogate_train(self, steps, batch=1024, lr=1e-3, seed=0):
        """Train on the stored samples; returns (first, last) minibatch losses."""
        a, b = C.c_float(), C.c_float()
        self._check(lib().gf_surrogate_train(self._ctx, int(steps), int(batch), float(lr), int(seed),
                                             C.byref(a), C.byref(b)))
        return a.value, b.value

    def surrogate_enable(self, on=True, tol=0.05, audit_stride=64):
        self._check(lib().gf_surrogate_enable(self._ctx, int(on), float(tol), int(audit_stride)))

    def surrogate_stats(self):
        """dict(evals, fallbacks, audits, max_audit_err, off, worst_element, failed_audits,
        cells_off) since the last enable."""
        v = (C.c_double * 8)()
        self._check(lib().gf_surrogate_stats(self._ctx, v))
        return {"evals": int(v[0]), "fallbacks": int(v[1]), "audits": int(v[2]), "max_audit_err": v[3],
                "off": int(v[4]), "worst_element": int(v[5]), "failed_audits": int(v[6]), "cells_off": int(v[7])}

    def surrogate_weights(self, w=None, cfg=None):
        """Without arguments: (weights[4737], cfg[66]); with them: load them."""
        import numpy as np
        fp = C.POINTER(C.c_float)
        if w is None:
            w = np.zeros(MLP_NPARAM, dtype=np.float32)
            cfg = np.zeros(66, dtype=np.float32)
            self._check(lib().gf_surrogate_weights(self._ctx, w.ctypes.data_as(fp), cfg.ctypes.data_as(fp), 1))
            return w, cfg
        w = np.ascontiguousarray(w, dtype=np.float32)
        cp = None if cfg is None else np.ascontiguousarray(cfg, dtype=np.float32)
        self._check(lib().gf_surrogate_weights(self._ctx, w.ctypes.data_as(fp),
                                               None if cp is None else cp.ctypes.data_as(fp), 0))

    def error_site(self):
        j, i = C.c_int32(), C.c_int32()
        lib().gf_error_site(self._ctx, C.byref(j), C.byref(i))
        return j.value, i.value

    def set_timing(self, on):
        return self._check(lib().gf_set_timing(self._ctx, int(bool(on))))

    def kernel_time(self, name):
        """(total ms, launches) of one kernel since timing was switched on."""
        ms, n = C.c_double(), C.c_int64()
        self._check(lib().gf_kernel_time(self._ctx, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def sync(self):
        return self._check(lib().gf_sync(self._ctx))

    def close(self):
        if self._ctx:
            lib().gf_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def math_check(fn, x, y=None):
    """Device values of the correctly rounded sin (0), tan (1), exp (2), pow (3)."""
    import numpy as np
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    yy = None if y is None else np.ascontiguousarray(y, dtype=np.float64)
    dp = C.POINTER(C.c_double)
    rc = lib().gf_math_check(int(fn), x.ctypes.data_as(dp),
                             None if yy is None else yy.ctypes.data_as(dp),
                             out.ctypes.data_as(dp), x.size)
    if rc != 0:
        raise GfError(rc, "gf_math_check failed")
    return out


def mlp_grad(x, t, w):
    """Loss and gradient of the surrogate network by the device's training kernels."""
    import numpy as np
    fp = C.POINTER(C.c_float)
    x = np.ascontiguousarray(x, dtype=np.float32)
    t = np.ascontiguousarray(t, dtype=np.float32)
    w = np.ascontiguousarray(w, dtype=np.float32)
    g = np.zeros(MLP_NPARAM, dtype=np.float32)
    loss = C.c_float()
    rc = lib().gf_mlp_grad(x.ctypes.data_as(fp), t.ctypes.data_as(fp), x.shape[0], w.ctypes.data_as(fp),
                           g.ctypes.data_as(fp), C.byref(loss))
    if rc != 0:
        raise RuntimeError("gf_mlp_grad failed (%d)" % rc)
    return loss.value, g


def mlp_infer(x, w):
    """The surrogate network's BF16 matrix-core forward pass on host inputs."""
    import numpy as np
    fp = C.POINTER(C.c_float)
    x = np.ascontiguousarray(x, dtype=np.float32)
    w = np.ascontiguousarray(w, dtype=np.float32)
    y = np.zeros(x.shape[0], dtype=np.float32)
    rc = lib().gf_mlp_infer(x.ctypes.data_as(fp), x.shape[0], w.ctypes.data_as(fp), y.ctypes.data_as(fp))
    if rc != 0:
        raise RuntimeError("gf_mlp_infer failed (%d)" % rc)
    return y


def stream_probe(nbytes=1 << 30, reps=10):
    """Average duration (ms) of one plain device copy of nbytes bytes (read + write)."""
    ms = C.c_double()
    rc = lib().gf_stream_probe(int(nbytes), int(reps), C.byref(ms))
    if rc != 0:
        raise RuntimeError("gf_stream_probe failed (%d)" % rc)
    return ms.value
