"""Write the golden fixtures of tests/golden from the compiled reference.

Runs oracle/_ref/ref_harness (the reference's own subroutines compiled from
/root/reference/src by `make -C oracle ref`, with a dump driver of this
project) on the inputs in tests/golden/inputs, serially, and stores:

  <case>_s0.npz        the complete state after setflac (step 0)
  <case>_s<N>.npz      the fields the step changes, after N steps
  output_<case>_s<N>.npz  the files the reference's own outflac and saveflac write
                       after N steps (GF_REF_WRITE=1), each as its raw bytes

Only run where the reference sources and oracle/_ref are present; the committed
fixtures are what the tests read elsewhere.
usage: python tests/golden/make_golden.py [case ...]
"""
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from geoflac_amd import state  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")

MUTABLE = ["cord", "vel", "temp", "temp0", "force", "amass", "rmass", "stress0", "area", "dvol",
           "strain", "se2sr", "aps", "visn", "e2sr", "sshrheat", "shrheat", "iphase", "phase_ratio",
           "mark_phase"]

CASES = {
    # 60x20 elements, 300 km x 100 km, crust/mantle, extension 0.5 cm/yr per side
    "box60x20": dict(inp="box60x20.inp", steps=[1, 10, 100, 1000, 10000]),
    # 400x100 elements over 8000 km x 200 km, four layers; the 35 km interface lies
    # inside an element row (fractional phase ratios)
    "layers400x100": dict(inp="layers400x100.inp", steps=[1, 10, 100],
                          fields=["cord", "vel", "temp", "stress0", "area", "dvol", "strain", "aps", "visn",
                                  "e2sr", "iphase", "phase_ratio", "mark_phase"]),
    # 80x40, four layers of distinct rheology (irheol 6, 12, 12, 3) plus an elastic
    # block (irheol 1), every interface inside an element row
    "hetero80x40": dict(inp="hetero80x40.inp", steps=[1, 10, 100, 1000]),
    # 60x30 with marker phase changes (sediment, oceanic crust, serpentinite)
    "phases60x30": dict(inp="phases60x30.inp", steps=[1, 11, 21, 100, 1000]),
    # stress boundary conditions (bc_update): hydrostatic side walls, a normal stress on
    # part of the surface, a Winkler base and the sea-water load
    "bcs60x20": dict(inp="bcs60x20.inp", steps=[1, 10, 100, 1000]),
    # dt_elastic fixed by dt_scale (idt_scale 1) and the prestress freeze (i_prestress 1)
    "idt1_60x20": dict(inp="idt1_60x20.inp", steps=[1, 10, 100]),
    # box60x20 with the driver's remeshing (GF_REM=1 of the harness) every few tens of
    # steps: five remeshings in 300 steps
    "remesh60x20": dict(inp="remesh60x20.inp", steps=[100, 300], env={"GF_REM": "1"}),
    # surface processes (diff_topo with resurface): topography diffusion over a weakened
    # block; resurface appends sediment markers from step ~5000 and removes eroded ones
    # by 12000
    "topo60x20": dict(inp="topo60x20.inp", steps=[10, 100, 1000, 5000, 8000, 12000],
                      fields=MUTABLE + ["dtopo", "dhacc", "dhacc_correction", "mark_a1", "mark_a2", "mark_ntriag",
                                        "mark_dead", "mark_id_elem", "nmark_elem"]),
}


# files outflac (src/outflac.f90) and saveflac (src/saveflac.f90) write; Amin.rs (a
# list-directed scalar) and the files setflac writes are not among them
OUTPUT_CASES = {"box60x20": 10, "hetero80x40": 10, "topo60x20": 10}
SETUP_FILES = {"nxnz.0", "pisos.rs", "vbc.s", "output.asc", "Amin.rs"}


def run_output(name, nsteps):
    import numpy as np
    inp = os.path.join(HERE, "inputs", CASES[name]["inp"])
    with tempfile.TemporaryDirectory() as tmp:
        out = os.path.join(tmp, "dump")
        env = dict(os.environ, GF_REF_WRITE="1")
        subprocess.run([HARNESS, inp, out, str(nsteps)], cwd=tmp, check=True, stdout=subprocess.DEVNULL, env=env)
        files = {}
        for f in sorted(os.listdir(tmp)):
            path = os.path.join(tmp, f)
            if os.path.isfile(path) and f not in SETUP_FILES:
                with open(path, "rb") as fh:
                    files[f] = np.frombuffer(fh.read(), dtype=np.uint8)
        np.savez_compressed(os.path.join(HERE, "output_%s_s%d.npz" % (name, nsteps)), **files)
        print(name, "output files", len(files))


def run_case(name, spec):
    inp = os.path.join(HERE, "inputs", spec["inp"])
    steps = spec["steps"]
    with tempfile.TemporaryDirectory() as tmp:
        out = os.path.join(tmp, "dump")
        cmd = [HARNESS, inp, out, str(max(steps)), "0"] + [str(s) for s in steps]
        subprocess.run(cmd, cwd=tmp, check=True, stdout=subprocess.DEVNULL, env=dict(os.environ, **spec.get("env", {})))
        with open(os.path.join(out, "params.txt")) as f:
            ptxt = f.read()
        with open(os.path.join(HERE, name + "_params.txt"), "w") as f:
            f.write(ptxt)
        s0 = state.load_dump(os.path.join(out, "step_0000000"))
        state.save_npz(s0, os.path.join(HERE, "%s_s0.npz" % name))
        for n in steps:
            sn = state.load_dump(os.path.join(out, "step_%07d" % n))
            state.save_npz(sn, os.path.join(HERE, "%s_s%d.npz" % (name, n)), fields=spec.get("fields", MUTABLE))
            print(name, n, "written")


if __name__ == "__main__":
    names = sys.argv[1:] or list(CASES)
    for n in names:
        if n.startswith("output_"):
            run_output(n[len("output_"):], OUTPUT_CASES[n[len("output_"):]])
        else:
            run_case(n, CASES[n])
            if n in OUTPUT_CASES:
                run_output(n, OUTPUT_CASES[n])
