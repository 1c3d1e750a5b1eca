#!/usr/bin/env python3
"""Throughput of the GeoFLAC explicit time step on one MI355X (element-steps/s).

Workload: the 4000x1000-element four-layer box (geoflac_amd/configs/
layers4000x1000.inp), synthetic new-start state built on the host, all arrays
resident in HBM before timing.  A step is one iteration of the reference's inner
loop (src/par.f90:81-91: flac; nloop += 1; time += dt).  Prints one JSON line.

roofline: per-kernel durations are measured with HIP events recorded on the
library's own stream (gf_set_timing / gf_kernel_time) over a separate profiled
run; the dominant kernel's algorithmic bytes per launch (alg_bytes) over its
average duration is `achieved`, and `traffic` its counter-measured bytes
(profiles/traffic.json).  cpu_baseline: the reference's own OpenMP program on the
host cores, on the benchmark deck cut to a bounded number of columns, timed before
this process touches the GPU; the C oracle on one core is a second field.
--mixed-report: the FP32 modes' per-field errors after 100, 1000 and 10000 steps
beside the FP64 run's one-ulp floor.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "element-steps/s on 1 GPU: mixed/learned vs plain FP64, at stated rel-L2 error"


def area_pending_fraction(p, n0, nsteps):
    """Fraction of the steps n0 .. n0+nsteps-1 whose fl_rheol forms the areas and dvol of
    the previous step's move (the device defers them to it), assuming none pending at
    n0: a step after rmasses or dt_mass finds them formed."""
    pend, cnt = False, 0
    for n in range(n0, n0 + nsteps):
        cnt += pend
        pend = not (n % p.ifreq_rmasses == 0 or n % p.ifreq_imasses == 0)
    return cnt / max(nsteps, 1)


def alg_bytes(kernel, E, N, nact, p, area_frac=0.0, mixfrac=1.0):
    """Algorithmic (compulsory) HBM bytes of one launch: each array the kernel must
    read once and write once, FP64 = 8 B, int32 = 4 B.  The nact phase-ratio planes
    are read only for the fraction mixfrac of elements holding more than one phase
    (an element of one phase is flagged beside its phase and reads none)."""
    kernel = kernel.replace("<double>", "").replace("<0>", "")
    if kernel.startswith("k_rheol"):
        per_e = (32 + 32 + 128 + 8 + 8 + 8 * nact * mixfrac + 4 + 24 + 8 + 24)  # area dvol stress0 aps visn pr iphase se2sr sshrheat strain
        per_e += 128 + 8 + 24 + 24 + 8                                 # writes stress0 aps strain se2sr sshrheat
        per_n = 16 + 16                                                # vel, cord
        # with the move's areas pending: no dvol read, area and dvol written
        return E * (per_e + 32 * area_frac) + N * per_n
    if kernel in ("k_therm_tile", "k_therm_pipe"):
        return N * (8 + 8 + 8 + 16) + E * (16 + 4 + 8)                # temp, temp0 out, x out, cord | area(1:2) iphase source
    if kernel in ("k_therm_tile_extra", "k_therm_pipe_extra"):
        return N * (8 + 8 + 8 + 16) + E * (16 + 4 + 8)                # temp0, x in, x out, cord | area(1:2) iphase source
    if kernel == "k_node":
        return E * (96 + 32) + N * (16 + 16 + 16 + 8 + 16 + 16 + 16)  # stress0(1:3) area | cord rmass+amass vel ncod bc, vel+force out
    if kernel == "k_node_t":  # the step's fl_node also writes the moved coordinates
        return E * (96 + 32) + N * (16 + 16 + 16 + 8 + 16 + 16 + 16 + 16)
    if kernel == "k_move_elem":
        return E * (32 + 96 + 24 + 32 + 32 + 96 + 24) + N * 32
    if kernel == "k_move_area":
        return E * (32 + 32 + 32) + N * 16                             # area in, area dvol out | cord
    if kernel == "k_move_rot":
        return E * (96 + 24 + 96 + 24) + N * 32                        # stress0(1:3) strain in and out | cord vel
    if kernel == "k_visc":
        return E * (96 + 8 + 8 * nact * mixfrac + 4 + 8) + N * 8       # stress0(1,2,4) e2sr pr iphase, visn out | temp
    if kernel == "k_dens":
        return E * (96 + 8 * nact * mixfrac + 4 + 8) + N * 8           # stress0(1,2,4) pr iphase, dens out | temp
    # (the step's masses and dt minima form the move's pending areas from the coordinates)
    if kernel == "k_rmass":
        return E * 8 + N * (8 + 16)                                    # dens | rmass out, cord
    if kernel == "k_amass":
        return E * 4 + N * (8 + 16)                                    # iphase | amass out, cord
    if kernel == "k_masses":
        return E * (8 + 4) + N * (16 + 16)                             # dens iphase | rmass amass out, cord
    if kernel == "k_dt_reduce":
        return E * (8 + 4) + N * 32                                    # visn iphase | cord vel
    if kernel == "k_move_cord":
        return N * 48
    if kernel == "k_therm_bc":  # the boundary conditions, copied to temp
        return N * 16
    if kernel == "k_therm_tail":
        # the boundary conditions in place on the boundary nodes of the result, which
        # becomes temp (a fallback pass inside k_therm_tail is added where it ran)
        return (2 * p.nx + 2 * p.nz) * 16
    if kernel == "k_area_check":
        return N * 16                                                  # cord (the move's negative-area test)
    if kernel == "k_therm_check":
        # the border check: 78 border values of each 64x16-node tile (54x12 outputs)
        # against the assembled iterate; fallback passes are added where they ran
        ntiles = -(-p.nz // 54) * -(-p.nx // 12)
        return ntiles * 78 * 16
    return 0


def traffic_from_profile(kernel):
    """HBM bytes per launch of `kernel` measured with profiler byte counters
    (profiles/traffic.json: FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by the
    same counter's reading for the copy kernel of known size), when that file holds
    this kernel; otherwise traffic stays null."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        t = json.load(f)
    k = t.get("kernels", {}).get(kernel)
    if not k:
        return {}
    return {"traffic": k["bytes_per_launch"], "traffic_source": "profiles/traffic.json (%s)" % t.get("command", "")}


INTERFACE_NOISE_KM = 2.0   # seeded per-column offsets of the layer interfaces
COH_NOISE = 0.01           # seeded per-element cohesion factor 1 +- 1%


def build_model(cfg, precision, seed):
    from geoflac_amd import inp, model, gpu
    p = inp.read_inp(os.path.join(ROOT, "geoflac_amd", "configs", cfg + ".inp"))
    gp, st = model.build_box(p, interface_noise=INTERFACE_NOISE_KM, coh_noise=COH_NOISE, seed=seed)
    m = gpu.Model(gp, st, precision=precision)
    model.device_setup(m)
    return p, gp, st, m


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _ref_deck(cfg_text, ncols, nsteps):
    """The benchmark deck cut to its first ncols element columns (the same cells,
    layers and boundary conditions) running exactly nsteps steps (ntest_rem = nsteps,
    time_max = 0; output intervals beyond the run)."""
    lines = cfg_text.splitlines()
    nelx, nelz = (int(v) for v in lines[7].split(","))
    x0, z0 = (float(v) for v in lines[8].split(","))
    rx, rz = (float(v) for v in lines[9].split(","))
    lines[7] = "%d,%d" % (ncols, nelz)
    lines[9] = "%r,%r" % (rx * ncols / nelx, rz)
    out = []
    for ln in lines:
        t = ln.split()
        if len(t) >= 4 and t[0] == "2" and t[2] == str(nelx + 1):          # bottom boundary nodes
            ln = ln.replace(" %d " % (nelx + 1), " %d " % (ncols + 1), 1)
        elif ln.strip().startswith("21,") and ln.rstrip().endswith(str(nelx + 1)):  # zone columns
            ln = ln[:ln.rstrip().rfind(str(nelx + 1))] + str(ncols + 1)
        elif ln.strip() == "0  3  10  11.0":                                # ny_rem, mode_rem, ntest_rem
            ln = "0  3  %d  11.0" % nsteps
        out.append(ln)
    return "\n".join(out) + "\n"


def ref_cpu_baseline(cfg, ncols=500, n1=10, n2=60, repeats=3):
    """The reference's own OpenMP program (oracle/_ref/ref_flac_omp, built from the
    reference sources with flang -O2 -fopenmp) on the host cores: loop-only time per
    step t = [t(n2) - t(n1)] / (n2 - n1) from runs of the benchmark deck cut to its
    first ncols element columns (same cells and physics; init, output and the final
    restart dump cancel in the difference), each run length timed `repeats` times and
    the shortest kept.  Runs as child processes before this process touches the GPU."""
    import resource
    import shutil
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_flac_omp")
    if not os.path.exists(exe):
        return None
    with open(os.path.join(ROOT, "geoflac_amd", "configs", cfg + ".inp")) as f:
        text = f.read()
    nelz = int(text.splitlines()[7].split(",")[1])
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PLACES="cores", OMP_PROC_BIND="close",
               OMP_STACKSIZE="1G")

    def unlimited_stack():
        resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))

    base = "/dev/shm" if os.path.isdir("/dev/shm") else None
    times = {n1: [], n2: []}
    for _ in range(repeats):
        for n in (n1, n2):
            d = tempfile.mkdtemp(prefix="gf_refomp_", dir=base)
            try:
                with open(os.path.join(d, "deck.inp"), "w") as f:
                    f.write(_ref_deck(text, ncols, n))
                t0 = time.perf_counter()
                r = subprocess.run([exe, "deck.inp"], cwd=d, env=env, preexec_fn=unlimited_stack,
                                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=600)
                times[n].append(time.perf_counter() - t0)
                if r.returncode != 0:
                    return None
            finally:
                shutil.rmtree(d, ignore_errors=True)
            print("bench: reference OpenMP run of %d steps: %.2f s" % (n, times[n][-1]), file=sys.stderr,
                  flush=True)
    times = {n: min(v) for n, v in times.items()}
    t_step = (times[n2] - times[n1]) / (n2 - n1)
    cpu = _cpu_model()
    E = ncols * nelz
    return {"value": E / t_step, "unit": "element-steps/s", "cores": threads, "kind": "reference",
            "cpu_model": cpu,
            "sample": "reference OpenMP program (flang -O2 -fopenmp), %d threads, OMP_PLACES=cores "
                      "OMP_PROC_BIND=close: the %s deck cut to its first %d of %d element columns "
                      "(%d elements, same cells and layers), loop time [t(%d)-t(%d)]/%d = %.3f s/step "
                      "(shortest of %d runs each: %.1f s and %.1f s)" % (threads, cfg, ncols, int(text.splitlines()[7].split(",")[0]),
                                                        E, n2, n1, n2 - n1, t_step, repeats, times[n1],
                                                        times[n2])}


def cpu_baseline(gp, st, seconds_cap=20.0):
    """The oracle (serial C restatement) on a copy of the same model, a few steps."""
    sys.path.insert(0, ROOT)
    from tests import oracle_lib as ol
    s = st.copy()
    E = (gp.nx - 1) * (gp.nz - 1)
    n, t0 = 0, time.perf_counter()
    while True:
        rc = ol.step(gp, s, 1)
        if rc != 0:
            break
        n += 1
        el = time.perf_counter() - t0
        if el > seconds_cap / 2 or n >= 5:
            break
    el = time.perf_counter() - t0
    return {"value": E * n / el, "unit": "element-steps/s", "cores": 1, "kind": "port",
            "cpu_model": _cpu_model(),
            "sample": "%d step(s) of the same %dx%d model by the oracle (serial C restatement "
                      "of the reference step), %.1f s" % (n, gp.nx - 1, gp.nz - 1, el)}


def mixed_run(gp, host_state, fp64_state, args, E, precision, learned=None):
    """The same model and step count in a mixed-precision mode (FP32 planes for stress,
    velocity and force; FP32 or FP64 arithmetic): throughput and per-field error
    against the FP64 run."""
    import numpy as np
    from geoflac_amd import abi, gpu, state
    st = host_state.copy()
    m = gpu.Model(gp, st, precision=precision)
    if learned is not None:  # the surrogate trained on the FP64 run's samples
        m.surrogate_weights(learned[0], learned[1])
        m.surrogate_enable(True, tol=0.5, audit_stride=64)
    m.step(args.warmup)
    m.sync()
    t0 = time.perf_counter()
    m.step(args.steps)
    m.sync()
    el = time.perf_counter() - t0
    m.download()
    sur = m.surrogate_stats() if learned is not None else None
    # time of the viscosity update per call (every ifreq_visc steps), exact or learned
    m.set_timing(True)
    m.step(10)
    vnames = ["k_visc_mlp<float>", "k_visc_list<float>"] if learned is not None \
        else ["k_visc<float>"]
    vt = {k: m.kernel_time(k) for k in vnames}
    calls = max(1, vt[vnames[-1]][1])
    m.set_timing(False)
    m.close()
    err = {}
    for f in ("vel", "stress0", "temp", "cord", "aps"):
        a, b = getattr(st, f), getattr(fp64_state, f)
        r = state.rel_l2(a, b)
        # rel-L2 is undefined where the FP64 field is zero (no plastic strain yet): null
        err[f] = {"rel_l2": float("%.3e" % r) if np.isfinite(r) else None,
                  "max_abs": float("%.3e" % np.max(np.abs(a - b)))}
    dtype = "f32 arithmetic, f32 planes" if precision == abi.PREC_MIXED else "f64 arithmetic, f32 planes"
    out = {"value": E * args.steps / el, "ms_per_step": el / args.steps * 1e3, "dtype": dtype,
           "steps_compared": args.warmup + args.steps, "error_vs_fp64": err,
           "viscosity_update_ms_per_call": dict({k: round(v[0] / calls, 4) for k, v in vt.items()},
                                                 total=round(sum(v[0] for v in vt.values()) / calls, 4))}
    if learned is not None:
        tot = max(1, sur["evals"] + sur["fallbacks"] + sur["audits"])
        out["surrogate"] = dict(learned[2], fallback_rate=round(sur["fallbacks"] / tot, 4),
                                audits=sur["audits"], max_audit_log10_err=round(sur["max_audit_err"], 5),
                                worst_element=sur["worst_element"], switched_off=sur["off"],
                                failed_audits=sur["failed_audits"], cells_taken_off=sur["cells_off"],
                                tolerance_log10=0.5)
    return out


def mixed_report(args):
    """Per-field error of both FP32 modes against the FP64 run after 100, 1000 and
    10000 steps of the benchmark model, beside the FP64 run's own sensitivity: the same
    FP64 run from an initial stress perturbed by one ulp (every stress value moved to
    the next double up), compared after 100 and 1000 steps (the chaos floor)."""
    import numpy as np
    from geoflac_amd import abi, gpu, state
    checkpoints = [c for c in (100, 1000, 10000) if c <= args.report_steps]
    fields = ("vel", "stress0", "temp", "cord", "aps")
    p, gp, st, m = build_model(args.config, abi.PREC_FP64, args.seed)
    m.download()
    host0 = st.copy()

    def run(precision, start, points):
        s = start.copy()
        mm = gpu.Model(gp, s, precision=precision)
        snaps, done, t = {}, 0, 0.0
        for c in points:
            mm.sync()
            t0 = time.perf_counter()
            mm.step(c - done)
            mm.sync()
            t += time.perf_counter() - t0
            done = c
            mm.download()
            snaps[c] = {f: getattr(s, f).copy() for f in fields}
            print("bench: mixed report, precision %d, %d steps" % (precision, c), file=sys.stderr, flush=True)
        mm.close()
        return snaps, t / max(done, 1) * 1e3

    def errs(a, b):
        out = {}
        for f in fields:
            r = state.rel_l2(a[f], b[f])
            out[f] = {"rel_l2": float("%.3e" % r) if np.isfinite(r) else None,
                      "max_abs": float("%.3e" % np.max(np.abs(a[f] - b[f])))}
        return out

    m.close()
    ref, ms64 = run(abi.PREC_FP64, host0, checkpoints)
    pert = host0.copy()
    pert.stress0[...] = np.nextafter(pert.stress0, np.inf)
    fl, _ = run(abi.PREC_FP64, pert, [c for c in checkpoints if c <= 1000])
    out = {"config": args.config, "seed": args.seed, "fp64_ms_per_step": round(ms64, 4),
           "fp64_one_ulp_floor": {str(c): errs(fl[c], ref[c]) for c in fl}}
    for name, prec in (("fp32_arithmetic", abi.PREC_MIXED), ("fp32_storage", abi.PREC_MIXED_STORAGE)):
        mx, ms = run(prec, host0, checkpoints)
        out[name] = {"ms_per_step": round(ms, 4), "speedup_vs_fp64": round(ms64 / ms, 3),
                     "error_vs_fp64": {str(c): errs(mx[c], ref[c]) for c in checkpoints}}
    print(json.dumps(out), flush=True)


def sweep_run(args):
    """Batched parameter sweep: seeded variants of the 60x20 box (mesh nx 100-600 and
    nz 40-150 elements over the same domain, friction 10-35 deg, cohesion 10-60 MPa,
    side velocity 0.2-3 cm/yr), mixed precision, advanced together by the batched
    launches with graph replay; per-model error against the same models run alone in
    FP64."""
    import numpy as np
    from geoflac_amd import abi, inp, state, sweep
    base = inp.read_inp(os.path.join(ROOT, "geoflac_amd", "configs", "box60x20.inp"))
    ps = sweep.sweep_params(base, n=args.sweep_models, seed=0)
    t0 = time.perf_counter()
    models = [sweep.make_model(q, abi.PREC_MIXED)[2] for q in ps]
    t_setup = time.perf_counter() - t0
    E_tot = sum((m.params.nx - 1) * (m.params.nz - 1) for m in models)
    with sweep.Batch(models) as b:
        b.set_graph(True)
        tuned = b.autotune(2)                       # 6 real steps
        b.step(4)                                   # to step 10: graph cycles from here
        t1 = time.perf_counter()
        b.step(args.sweep_steps)
        el = time.perf_counter() - t1
    done = 10 + args.sweep_steps
    errs = []
    for k in range(min(args.sweep_check, len(models))):
        m = models[k]
        m.download()
        gp, st, ref = sweep.make_model(ps[k], abi.PREC_FP64)
        ref.step(done)
        ref.download()
        ref.close()
        errs.append({"model": k, "elements": (gp.nx - 1) * (gp.nz - 1),
                     "vel_rel_l2": float("%.3e" % state.rel_l2(m.state.vel, st.vel)),
                     "stress_rel_l2": float("%.3e" % state.rel_l2(m.state.stress0, st.stress0)),
                     "temp_rel_l2": float("%.3e" % state.rel_l2(m.state.temp, st.temp))})
    for m in models:
        m.close()
    return {"value": E_tot * args.sweep_steps / el, "unit": "element-steps/s", "models": len(models),
            "elements_total": E_tot, "steps": args.sweep_steps, "ms_per_step": el / args.sweep_steps * 1e3,
            "dtype": "f32 arithmetic, f32 planes", "graph": True, "autotuned_workgroup_sizes": tuned,
            "setup_s": round(t_setup, 1), "error_vs_fp64_alone": errs}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--config", default="layers4000x1000")
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-precision runs (FP32 stress/strain rate/force/velocity)")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned-viscosity run (surrogate trained on this model's samples)")
    ap.add_argument("--no-sweep", action="store_true", help="skip the batched parameter sweep")
    ap.add_argument("--sweep-models", type=int, default=256)
    ap.add_argument("--sweep-steps", type=int, default=2000)
    ap.add_argument("--sweep-check", type=int, default=4,
                    help="models of the sweep also run alone in FP64 for the error report")
    ap.add_argument("--seed", type=int, default=0,
                    help="seed of the layer-interface offsets and the per-element cohesion factors")
    ap.add_argument("--mixed-report", action="store_true",
                    help="only the FP32 modes' per-field error after 100/1000/10000 steps and the FP64 one-ulp "
                         "floor (one JSON line)")
    ap.add_argument("--report-steps", type=int, default=10000)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="short copy probe only (for profiler counter runs, where k_copy over 1 GiB is the "
                         "kernel of known traffic that scales FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this benchmark runs on one GPU only (one model per device); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.mixed_report:
        mixed_report(args)
        return
    # the reference's OpenMP program on the host cores, before this process touches the GPU
    ref_cpu = None if args.no_cpu_baseline else ref_cpu_baseline(args.config)

    from geoflac_amd import abi, gpu
    p, gp, st, m = build_model(args.config, abi.PREC_FP64, args.seed)
    E = (gp.nx - 1) * (gp.nz - 1)
    N = gp.nx * gp.nz
    m.download()
    host_state = st.copy()  # complete new-start state, for the CPU baseline

    m.step(args.warmup)
    m.sync()
    t0 = time.perf_counter()
    m.step(args.steps)       # returns after the device finished (clock read-back)
    m.sync()
    t1 = time.perf_counter()
    el = t1 - t0
    value = E * args.steps / el
    m.download()
    fp64_state = st.copy()   # after warmup + steps, for the mixed-precision comparison
    # the viscosity network: trained beforehand on a separate run (a different seed),
    # stored by geoflac_amd/train_surrogate.py
    learned_w, learned_note = None, None
    if not args.no_learned:
        from geoflac_amd import train_surrogate
        wpath = os.path.join(ROOT, "geoflac_amd", "weights", "visc_mlp_%s.npz" % args.config)
        if os.path.exists(wpath):
            w, wcfg, meta = train_surrogate.load(wpath)
            learned_w = (w, wcfg, dict(meta, weights=os.path.relpath(wpath, ROOT)))
        else:
            learned_note = "no stored network for %s (python -m geoflac_amd.train_surrogate)" % args.config

    # profiled run: HIP events around every launch on the library stream
    args.profile_steps = max(args.profile_steps, 1)
    m.therm_stats()
    m.set_timing(True)
    n0 = m.state.clock.nloop
    m.step(args.profile_steps)
    afrac = area_pending_fraction(gp, n0, args.profile_steps)
    kernels = ["k_rheol1<false>", "k_rheol1<true>", "k_rheol", "k_visc<double>", "k_therm_pipe",
               "k_therm_pipe_extra", "k_therm_check", "k_node<0>", "k_node_t<0>", "k_move_elem<0>",
               "k_move_area", "k_move_rot<0>", "k_move_cord<double>", "k_therm_bc", "k_therm_tail", "k_dens<double>", "k_rmass", "k_dt_reduce<double>", "k_amass", "k_masses",
               "k_area_check", "k_clock"]
    times = {k: m.kernel_time(k) for k in kernels}
    m.set_timing(False)
    th = m.therm_stats()
    dom = max(times, key=lambda k: times[k][0])
    ms_tot, nl = times[dom]
    nact = int((host_state.phase_ratio >= 0.01).any(axis=(1, 2)).sum())
    mixfrac = float((host_state.phase_ratio.max(axis=0) < 1.0).mean())  # elements of more than one phase
    ab = alg_bytes(dom, E, N, nact, gp, afrac, mixfrac)
    # a fallback pass and its check count only where the pass ran (the launches after a
    # proven pass exit at once)
    step_bytes = sum(alg_bytes(k, E, N, nact, gp, afrac, mixfrac) * v[1] for k, v in times.items()
                     if k not in ("k_therm_pipe_extra", "k_therm_check"))
    step_bytes += alg_bytes("k_therm_pipe_extra", E, N, nact, gp) * th[1]
    step_bytes += alg_bytes("k_therm_check", E, N, nact, gp) * (args.profile_steps + th[1])
    step_bytes //= args.profile_steps
    step_ms = sum(v[0] for v in times.values()) / args.profile_steps
    avg_s = ms_tot / 1e3 / max(nl, 1)
    achieved = ab / avg_s / 1e9 if ab else 0.0
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "alg_bytes_per_launch": ab, "mixed_element_fraction": round(mixfrac, 5),
                "avg_launch_us": round(avg_s * 1e6, 2),
                "kernel_ms_per_step": {k: round(v[0] / args.profile_steps, 4) for k, v in times.items() if v[1]},
                "therm_extra_passes_per_step": th[1] / args.profile_steps,
                # the whole step: algorithmic bytes of every launch over the kernels' time
                "step_alg_bytes": step_bytes, "step_kernel_ms": round(step_ms, 4),
                "step_achieved": round(step_bytes / (step_ms * 1e-3) / 1e9, 1) if step_ms else None,
                "therm_border_mismatches": th[2], "therm_mismatch_max_ulp": th[3]}
    m.close()
    mixed = None
    if not args.no_mixed:
        mixed = {"fp32_arithmetic": mixed_run(gp, host_state, fp64_state, args, E, abi.PREC_MIXED),
                 "fp32_storage": mixed_run(gp, host_state, fp64_state, args, E, abi.PREC_MIXED_STORAGE)}
    if learned_w is not None:
        mixed = mixed or {}
        lr = mixed_run(gp, host_state, fp64_state, args, E, abi.PREC_MIXED, learned=learned_w)
        ex = mixed.get("fp32_arithmetic")
        if ex:  # the network's viscosity update against the exact one of the same mode
            lr["visc_update_gain_vs_exact"] = round(ex["viscosity_update_ms_per_call"]["total"] /
                                                    lr["viscosity_update_ms_per_call"]["total"], 3)
            lr["step_gain_vs_exact"] = round(lr["value"] / ex["value"], 3)
        mixed["fp32_arithmetic_learned_visc"] = lr
    elif learned_note:
        mixed = mixed or {}
        mixed["fp32_arithmetic_learned_visc"] = {"skipped": learned_note}
    # measured copy bandwidth (k_copy over 1 GiB: 2 GiB of traffic per launch)
    copy_ms = gpu.stream_probe(1 << 30, 10 if not args.copy_calibration else 3)
    roofline["copy_bw_GBs"] = round(2.0 * (1 << 30) / (copy_ms * 1e-3) / 1e9, 1)
    roofline["frac_of_copy_bw"] = round(achieved / roofline["copy_bw_GBs"], 4)
    roofline.update(traffic_from_profile(dom))

    out = {"metric": METRIC, "value": value, "unit": "element-steps/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": el / args.steps * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": "%s: %dx%d-element four-layer extension box, plain FP64 drop-in path"
                                  % (args.config, gp.nx - 1, gp.nz - 1),
                      "elements": E, "nodes": N, "precision": "fp64", "parallelism": "single GPU",
                      "markers": int(st.nmarkers), "change_phase": "every %d steps" % gp.ifreq_rmasses,
                      "interface_noise_km": INTERFACE_NOISE_KM, "cohesion_noise": COH_NOISE, "seed": args.seed},
           "roofline": roofline}
    if mixed:
        for v in mixed.values():
            if "value" in v:
                v["speedup_vs_fp64"] = round(v["value"] / value, 3)
        out["mixed"] = mixed
    if not args.no_sweep:
        out["sweep"] = sweep_run(args)
    if not args.no_cpu_baseline:
        port = cpu_baseline(gp, host_state)
        if ref_cpu is not None:
            ref_cpu["port_1core"] = port
            out["cpu_baseline"] = ref_cpu
        else:
            out["cpu_baseline"] = port
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
