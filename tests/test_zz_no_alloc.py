"""gf_mlp_grad_noalloc, the allocation-free variant of gf_mlp_grad: called through the
C-ABI on device buffers twice in a row (the second time with fewer samples, so stale
partial gradients of the first call would show), each result bitwise equal to
gf_mlp_grad on the same data."""
import ctypes as C
import sys

import numpy as np
import pytest

from tests import mlp_ref


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _hip():
    """The HIP runtime libgfhip.so is linked against, for the caller's own device
    buffers (torch would bring a second HIP runtime into the process)."""
    import os
    import re
    import subprocess
    from geoflac_amd import gpu
    out = subprocess.run(["ldd", gpu.LIB_PATH], capture_output=True, text=True).stdout
    m = re.search(r"libamdhip64\.so\S*\s+=>\s+(\S+)", out)
    H = C.CDLL(m.group(1) if m else "libamdhip64.so")
    H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    H.hipFree.argtypes = [C.c_void_p]
    H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return H


def _to_dev(H, a):
    p = C.c_void_p()
    assert H.hipMalloc(C.byref(p), a.nbytes) == 0
    assert H.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0   # host to device
    return p


@pytest.mark.gpu
def test_noalloc_gradient_equals_allocating_entry_point():
    from geoflac_amd import gpu
    L = gpu.lib()
    H = _hip()
    rng = np.random.default_rng(23)
    w = (mlp_ref.init(seed=3) + rng.normal(0, 0.05, mlp_ref.NPARAM)).astype(np.float32)
    wd = _to_dev(H, w)
    grad_d = _to_dev(H, np.full(mlp_ref.NPARAM + 1, np.nan, dtype=np.float32))
    bufs = [wd, grad_d]
    try:
        for n in (20000, 3000):
            _say("case: gf_mlp_grad_noalloc on %d device samples vs gf_mlp_grad" % n)
            x = rng.normal(size=(n, 7)).astype(np.float32)
            t = rng.normal(size=n).astype(np.float32)
            t[::97] = np.nan  # samples outside the domain: weight 0 in both entry points
            samples = np.ascontiguousarray(np.concatenate([x, t[:, None]], axis=1), dtype=np.float32)
            sd = _to_dev(H, samples)
            bufs.append(sd)
            rc = L.gf_mlp_grad_noalloc(sd, n, wd, grad_d, None)
            assert rc == 0, rc
            assert H.hipDeviceSynchronize() == 0
            got = np.empty(mlp_ref.NPARAM + 1, dtype=np.float32)
            assert H.hipMemcpy(got.ctypes.data, grad_d, got.nbytes, 2) == 0   # device to host
            loss, g = gpu.mlp_grad(x, t, w)
            _say("  loss %.6e vs %.6e, gradient equal %s" % (got[-1] / n, loss, np.array_equal(got[:-1], g)))
            assert np.array_equal(got[:-1], g)
            assert np.float32(got[-1] / np.float32(n)) == np.float32(loss)
        assert L.gf_mlp_grad_noalloc(sd, 0, wd, grad_d, None) != 0
    finally:
        for b in bufs:
            H.hipFree(b)
