"""The Fortran drop-in: the reference's own driver code (read_params, setflac, its
arrays and parameters; compiled from the reference sources by oracle/Makefile) calling
geoflac_amd/fortran/flac_hip.f90 in place of src/flac.f90, through the ISO_C_BINDING
module gf_hip.f90 and the C-ABI.  Its dumps must match the reference's goldens to the
same bar as the Python-driven runs.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from geoflac_amd import state
from tests import oracle_lib as ol

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "dropin_harness")
CASE = "box60x20"
TOL = 1e-10


def test_fortran_interface_mirrors_c_structs():
    """gf_hip.f90 declares the C structs' fields in the header's order (no GPU)."""
    import re
    hdr = open(os.path.join(ROOT, "include", "gf_abi.h")).read()
    f90 = open(os.path.join(ROOT, "geoflac_amd", "fortran", "gf_hip.f90")).read()

    def c_fields(struct):
        body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (struct, struct), hdr, re.S).group(1)
        body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
        out = []
        for decl in body.split(";"):
            decl = re.sub(r"^const\s+", "", decl.strip())
            if not decl:
                continue
            names = decl.split(None, 1)[1]
            out += [n.strip().lstrip("*").split("[")[0] for n in names.split(",")]
        return out

    def f_fields(typ):
        body = re.search(r"type, bind\(C\) :: %s\n(.*?)end type" % typ, f90, re.S).group(1)
        out = []
        for line in body.splitlines():
            if "::" in line:
                names = line.split("::")[1]
                out += [re.sub(r"\(.*?\)", "", n).strip() for n in re.split(r",(?![^(]*\))", names)]
        return out

    assert c_fields("gf_params") == f_fields("gf_params_t")
    assert c_fields("gf_host_arrays") == f_fields("gf_host_arrays_t")
    assert c_fields("gf_clock") == f_fields("gf_clock_t")
    assert c_fields("gf_marker_arrays") == f_fields("gf_marker_arrays_t")
    assert c_fields("gf_output_opts") == f_fields("gf_output_opts_t")


def _run_dropin(tmp_path, case, dumps, env=None):
    """Runs the drop-in harness on tests/golden/inputs/<case>.inp and checks its dumps
    against the reference's goldens: marker phases, iphase and phase ratios equal, the
    floating-point fields within TOL rel-L2, nloop equal.  Returns the harness's stdout."""
    if not os.path.exists(EXE):
        pytest.skip("oracle/_ref/dropin_harness is built only where the reference sources are present")
    inp = os.path.join(ROOT, "tests", "golden", "inputs", case + ".inp")
    out = str(tmp_path / "out")
    r = subprocess.run([EXE, inp, out, str(max(dumps))] + [str(d) for d in dumps], cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=300, env=dict(os.environ, **(env or {})))
    sys.stdout.write(r.stdout[-2000:])
    assert r.returncode == 0, r.stderr[-2000:]
    for step in dumps:
        st = state.load_dump(os.path.join(out, "step_%07d" % step))
        ref = ol.golden_state(case, step)
        if ref.has_markers:  # (surface processes: resurface appends and removes markers)
            assert st.nmarkers == ref.nmarkers
            for f in ("mark_phase", "mark_dead", "mark_ntriag", "mark_id_elem", "nmark_elem"):
                assert np.array_equal(getattr(st, f), getattr(ref, f)), f
        else:
            assert np.array_equal(st.mark_phase_only, ref.mark_phase_only)
        assert np.array_equal(st.iphase, ref.iphase)
        assert np.array_equal(st.phase_ratio, ref.phase_ratio)
        worst = {}
        topo = ("dtopo", "dhacc", "dhacc_correction", "mark_a1", "mark_a2") if ref.has_markers else ()
        for f in ("cord", "vel", "temp", "stress0", "aps", "strain", "area", "visn", "e2sr", "se2sr",
                  "rmass", "amass", "temp0") + topo:
            a, b = getattr(st, f), getattr(ref, f)
            worst[f] = 0.0 if np.array_equal(a, b) else state.rel_l2(a, b)
        print("  step %d: %s" % (step, {k: "%.1e" % v for k, v in worst.items() if v}), flush=True)
        assert max(worst.values()) <= TOL, worst
        assert st.clock.nloop == ref.clock.nloop
        assert abs(st.clock.dt - ref.clock.dt) <= 1e-12 * ref.clock.dt
    return r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("case,dumps", [("box60x20", (10, 100)), ("phases60x30", (11, 100)), ("bcs60x20", (10, 100)),
                                        ("topo60x20", (10, 1000))])
def test_fortran_dropin_matches_reference(tmp_path, case, dumps):
    """box60x20; phases60x30, whose markers change phase on the device (the marker
    phases, phase ratios and iphase come back to the Fortran arrays); bcs60x20 with
    stress boundaries; topo60x20 with surface processes (the topography arrays come
    back to the Fortran module)."""
    print("case: Fortran drop-in, %d steps of %s (dumps at %s)" % (max(dumps), case, dumps), flush=True)
    _run_dropin(tmp_path, case, dumps)


@pytest.mark.gpu
def test_fortran_dropin_resurface_markers_match_the_oracle(tmp_path):
    """topo60x20 through the drop-in for 8000 steps: resurface appends sediment markers
    on the device from step ~5000 on, and flac_hip_sync_host brings them back to the
    Fortran arrays with their count.  Over 8000 steps the device's correctly rounded
    libm and the reference's host libm take different paths (the run is chaotic), so
    the dump is compared with the correctly rounded oracle from the reference's step-0
    state, bit for bit: every field, the topography arrays and the markers."""
    import re
    if not os.path.exists(EXE):
        pytest.skip("oracle/_ref/dropin_harness is built only where the reference sources are present")
    case, n = "topo60x20", 8000
    inp = os.path.join(ROOT, "tests", "golden", "inputs", case + ".inp")
    print("case: Fortran drop-in, %d steps of %s against the correctly rounded oracle" % (n, case), flush=True)
    r = subprocess.run([EXE, inp, str(tmp_path / "out"), str(n), str(n)], cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    st = state.load_dump(os.path.join(str(tmp_path / "out"), "step_%07d" % n))
    p = ol.golden_params(case)
    ref = ol.golden_full(case, 0, marker_room=True)
    assert ol.step(p, ref, n, cr=True) == 0
    assert st.nmarkers == ref.nmarkers > ol.golden_state(case, 0).nmarkers, (st.nmarkers, ref.nmarkers)
    bad = [f for f in state.State.FIELDS + list(state.TOPO) + ["mark_id_elem", "nmark_elem"]
           if not np.array_equal(getattr(st, f), getattr(ref, f))]
    bad += [f for f in ("mark_a1", "mark_a2", "mark_ntriag", "mark_phase", "mark_dead", "mark_age")
            if not np.array_equal(st.live_markers(f), ref.live_markers(f))]
    print("  markers %d, differing %s" % (st.nmarkers, bad), flush=True)
    assert not bad, bad


@pytest.mark.gpu
def test_fortran_dropin_with_host_remeshing_matches_reference(tmp_path):
    """Remeshing stays with the driver, on the host: every ntest_rem steps the harness
    (GF_REM=1) brings the device state to the Fortran arrays (flac_hip_sync_host), runs
    the reference's itest_mesh and, when it asks for it, bar2euler, remesh, lpeuler2bar
    and marker2elem (src/par.f90:94-110), then hands the remeshed arrays, the markers
    and the new dt back to the device (flac_hip_sync_device).  remesh60x20 remeshes
    five times in 300 steps; the states after 100 and 300 steps match the reference
    run of the same driver loop (tests/golden/remesh60x20_s*.npz)."""
    print("case: Fortran drop-in with the driver's remeshing, 300 steps of remesh60x20", flush=True)
    out = _run_dropin(tmp_path, "remesh60x20", (100, 300), env={"GF_REM": "1"})
    import re
    m = re.search(r"remeshings\s+(\d+)", out)
    assert m and int(m.group(1)) == 5, out[-500:]


@pytest.mark.gpu
def test_fortran_dropin_writes_the_reference_files(tmp_path):
    """The drop-in's writers (flac_hip_outflac / flac_hip_saveflac: outflac's frames
    evaluated on the device, saveflac's dumps, without copying the state to the
    Fortran arrays) after 10 steps of box60x20 write the reference writers' files byte
    for byte (tests/golden/output_box60x20_s10.npz)."""
    if not os.path.exists(EXE):
        pytest.skip("oracle/_ref/dropin_harness is built only where the reference sources are present")
    inp = os.path.join(ROOT, "tests", "golden", "inputs", "box60x20.inp")
    print("case: Fortran drop-in writers after 10 steps of box60x20", flush=True)
    env = dict(os.environ, GF_HIP_WRITE="1")
    r = subprocess.run([EXE, inp, str(tmp_path / "out"), "10"], cwd=str(tmp_path), capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0, r.stderr[-2000:]
    ref = dict(np.load(os.path.join(ROOT, "tests", "golden", "output_box60x20_s10.npz")))
    bad = {}
    for name, want in sorted(ref.items()):
        path = tmp_path / name
        if not path.exists():
            bad[name] = "missing"
            continue
        mine = np.frombuffer(path.read_bytes(), dtype=np.uint8)
        if mine.size != want.size or not np.array_equal(mine, want):
            bad[name] = "differs"
    print("  %d files, differing %s" % (len(ref), bad), flush=True)
    assert not bad, bad
