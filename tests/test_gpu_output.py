"""outflac's frames and saveflac's restart files written from the device state
(gf_outflac, gf_saveflac) against the files the reference's own outflac and saveflac
wrote from the same state (tests/golden/output_<case>_s<N>.npz, raw bytes of each
file): every file byte for byte.

The arrays the explicit step never changes (xoriginal, dhacc, the magma and dike
fields, the marker ages) belong to the caller; the test passes the reference's own
values of them, read from its restart files."""
import os
import sys

import numpy as np
import pytest

from geoflac_amd import inp
from tests import oracle_lib as ol

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
CASES = [("box60x20", 10), ("hetero80x40", 10), ("topo60x20", 10)]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _host_arrays(ref, nmarkers):
    f8 = lambda n: np.frombuffer(ref[n].tobytes(), dtype=np.float64)  # noqa: E731
    out = {n: f8(n + ".rs") for n in ("xoriginal", "zoriginal", "dhacc", "extr_acc", "q_init", "vel_flex_old",
                                         "fmagma", "fmagma2", "dike_backlog", "dike_marker_vol", "mor_marker_vol")}
    out["dhacc_correction"] = f8("dhacc_corr.rs")
    out["mark_age"] = f8("marker1.rs")[2 * nmarkers:3 * nmarkers]
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("case,step", CASES)
def test_output_files_byte_equal_to_reference(case, step, tmp_path):
    from geoflac_amd import gpu
    _say("case: outflac and saveflac of the %d-step state of %s" % (step, case))
    ref = dict(np.load(os.path.join(GOLDEN, "output_%s_s%d.npz" % (case, step))))
    deck = inp.read_inp(os.path.join(GOLDEN, "inputs", case + ".inp"))
    p = ol.golden_params(case)
    st = ol.golden_full(case, step)
    arrays = _host_arrays(ref, st.nmarkers)
    with gpu.Model(p, st) as m:
        opts = m.output_opts(io_flags=deck["io_flags"], lastout=deck["lastout"], hs=deck["hs"], **arrays)
        m.outflac(tmp_path, opts)
        m.saveflac(tmp_path, opts)
    written = sorted(os.listdir(tmp_path))
    assert written == sorted(ref), (set(written) ^ set(ref))
    bad = {}
    for name in sorted(ref):
        with open(os.path.join(tmp_path, name), "rb") as f:
            mine = np.frombuffer(f.read(), dtype=np.uint8)
        want = ref[name]
        if mine.size != want.size:
            bad[name] = "size %d, reference %d" % (mine.size, want.size)
        elif not np.array_equal(mine, want):
            bad[name] = "%d bytes differ" % int(np.count_nonzero(mine != want))
    _say("  %d files, differing %s" % (len(ref), bad))
    assert not bad, bad
