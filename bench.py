"""Benchmark of the ps3D model time step on one MI355X.

Workload (the config the metric is quoted on): two balanced zonal jets on a
256 x 256 x 64 grid, rigid lid, FP64, AB3, harmonic (Ah, Av, Kh, Kv) and
horizontal biharmonic (Ahbi, Khbi) friction -- 8 friction passes per step.
A step is one call of integrate (integrate.f90:2-115) with the state already
resident in HBM.  value = grid points x steps / seconds over exactly --steps
steps, bracketed by device synchronisation.

The JSON line also carries:
* roofline: the kernel with the largest share of the step, timed with HIP
  events around each of its launches on the model's stream during the timed
  steps; achieved = its algorithmic bytes per launch / mean launch time;
  traffic = its measured HBM bytes per launch (rocprofv3 FETCH_SIZE +
  WRITE_SIZE, calibrated on a device copy; profiles/).
* cpu_baseline: the CPU oracle (a C port of the reference step), one
  independent instance per host core (at most 16), all stepped at once
  on a bounded sample of the same workload.
* mixed: the mixed-precision form (FP32 stencils and transforms, FP64
  spectral solve and (hi, lo) accumulation) on the same workload: its
  grid-pt steps/s, the speed-up over FP64, and its relative L2 error per
  field and of the total energy (diagnose.F90:14-24) against the FP64 path
  at steps 1, 100 and 1000 from the same initial state.
* ensemble: 48 baroclinic-eddy members, 16 each of 64x64x32, 96x96x48 and
  128x128x64 (eddy radius drawn per member), batched (one launch per stage,
  hipGraph replay, autotuned tile shapes) against the same members run one
  after another as single models: member-steps/s of both and their ratio.
* ensemble_mixed: the same 48 members in the mixed-precision form, batched
  against single mixed models run one after another.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "grid-pt steps/s FP64 vs mixed @≤1e-5 rel L2; ensemble member-steps/s"
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    """A progress line on stderr (the JSON result is the one stdout line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(name: str, prm, order: int = 3) -> float:
    """Compulsory HBM bytes of one launch of a kernel (element size 8 B)."""
    N = prm.nx * prm.ny * prm.nz
    S = (prm.nx // 2 + 1) * prm.ny * prm.nz
    e = 8
    table = {
        "tendency": 8 * e * N,                 # R u,v,w,b  W du,dv,dw,db(tau)
        "harmonic": 2 * e * N,                 # R + W one field
        "biharmonic_h": 2 * e * N,
        "biharmonic_v": 2 * e * N,
        "ab_update": 4 * (2 + order) * e * N,  # per field R u, q slots  W u
        "friction_ab": 4 * (2 + order) * e * N,  # friction chain fused into the AB pass: same I/O
        # tendency + friction + AB: per field R f and the q-1 older slots,
        # W slot tau and the new f -- the same count
        "tend_friction_ab": 4 * (2 + order) * e * N,
        "fft_x_r2c": 3 * e * N + 2 * e * S,    # R u,v,w (divergence)  W FS
        "fft_y_fwd": 4 * e * S,                # R + W complex FS
        "fft_y_inv": 4 * e * S,
        "thomas": 4 * e * S,
        "z_solve": 4 * e * S,
        "fft_x_c2r": 2 * e * S + e * N,        # R FS  W p
        "project": 7 * e * N,                  # R u,v,w,p  W u,v,w
        "halo": 0.0,
    }
    return float(table[name])


def measured_traffic(kernel: str):
    """HBM bytes per launch of the kernel on this workload, from the committed
    rocprofv3 runs of the TCC request counters by request size (profiles/;
    exact on the library's streaming copy); None where no profile exists."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_fp64_256x256x64_jets.json")
    try:
        with open(path) as f:
            return json.load(f)["traffic_bytes_per_launch"].get(kernel)
    except (OSError, ValueError, KeyError):
        return None


def hbm_probe(mib: int = 1024, reps: int = 10) -> dict:
    """Rates of the library's hand-written streaming kernels (ps3d_hbm_probe,
    probe.hip) on this GPU: copy, read and write of a MIB-MiB buffer, plain
    and nontemporal, best of reps launches; 'ceiling' is the faster copy,
    the achievable rate for kernels that read and write."""
    import ctypes
    import numpy as np
    from ps3d_amd import _lib
    lib = _lib.load()
    out = np.zeros(8)
    _lib.check(lib.ps3d_hbm_probe(mib << 20, reps, 0, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    names = ("copy", "copy_nt", "read", "read_nt", "write", "write_nt")
    r = {n: float(v) for n, v in zip(names, out[:6])}
    r["ceiling"] = max(r["copy"], r["copy_nt"])
    r["workgroups"] = int(out[6])
    r["buffer_mib"] = mib
    r["unit"] = "GB/s"
    return r


def fma_probe(reps: int = 5) -> dict:
    """Vector FMA issue rates of the library's probe (ps3d_fma_probe,
    probe.hip), TFLOP/s: FP32 as the library compiles it (without the packed
    FP32 instructions) and FP64."""
    import ctypes
    import numpy as np
    from ps3d_amd import _lib
    lib = _lib.load()
    out = np.zeros(2)
    _lib.check(lib.ps3d_fma_probe(reps, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return {"fp32_fma": float(out[0]), "fp64_fma": float(out[1]), "unit": "TFLOP/s"}


def build_workload(precision=0):
    from ps3d_amd import configs
    return configs.jets(256, 256, 64, seed=2, precision=precision)


def time_steps(model, steps: int) -> float:
    model.synchronize()
    t0 = time.perf_counter()
    model.run(steps)
    model.synchronize()
    return time.perf_counter() - t0


def time_steps_graph(model, steps: int) -> float:
    model.synchronize()
    t0 = time.perf_counter()
    model.run(steps, use_graph=True)
    model.synchronize()
    return time.perf_counter() - t0


def mixed_leg(prm, fields, steps: int, warmup: int, fp64_value: float, check_steps=(1, 100, 1000)) -> dict:
    from ps3d_amd import diagnose
    from ps3d_amd.model import Model
    from ps3d_amd.params import PS3D_FP64, PS3D_MIXED
    pm = prm.replace(precision=PS3D_MIXED)
    m = Model(pm)
    for n, a in fields.items():
        m.f[n][...] = a
    m.push()
    m.run(max(warmup, 3))
    m.tune(reps=5)
    plans = m.fft_plan
    tuning = m.tuning().strip().splitlines()
    el = time_steps(m, steps)
    value = prm.npoints * steps / el
    m.close()
    # accuracy: both forms from the same initial state, the mixed one with
    # the radix plans its tuner picked
    g64 = Model(prm.replace(precision=PS3D_FP64))
    gmx = Model(pm)
    for n, a in fields.items():
        g64.f[n][...] = a
        gmx.f[n][...] = a
    g64.push()
    gmx.push()
    gmx.fft_plan = plans
    errs = {}
    done = 0
    for upto in check_steps:
        g64.run(upto - done)
        gmx.run(upto - done)
        done = upto
        g64.pull()
        gmx.pull()
        errs[str(upto)] = diagnose.compare(gmx.f, g64.f, prm)
    g64.close()
    gmx.close()
    worst = max(max(e.values()) for e in errs.values())
    return {"value": value, "unit": "grid-pt steps/s", "ms_per_step": 1e3 * el / steps,
            "speedup_vs_fp64": value / fp64_value, "rel_l2_vs_fp64": errs, "max_rel_l2": worst,
            "tuning": tuning,
            "dtype": "f32 stencils/FFT, f64 solve + accumulation"}


def ensemble_leg(steps: int, warm: int = 3, precision: int = 0) -> dict:
    import numpy as np
    from ps3d_amd import configs
    from ps3d_amd.ensemble import Ensemble
    from ps3d_amd.model import Model
    members = configs.ensemble_c4(precision=precision)
    n = len(members)
    # batched: warm-up steps, then the autotuner (tile shapes; radix plans of
    # the mixed form's FP32 transforms)
    ens = Ensemble([p for p, _ in members])
    for i, (_, fields) in enumerate(members):
        for k, a in fields.items():
            ens.set(i, k, a)
    ens.run(warm)
    t0 = time.perf_counter()
    ens.tune(os.path.join("/tmp", f"ps3d_ensemble_tune_{precision}.txt"), reps=5)
    tune_s = time.perf_counter() - t0
    plans = [ens.fft_plan(i) for i in range(n)]
    # one after another: each member alone through the single-model path,
    # with the plans the ensemble's tuner picked
    seq = 0.0
    for i, (prm, fields) in enumerate(members):
        m = Model(prm)
        for k, a in fields.items():
            m.f[k][...] = a
        m.push()
        if precision:
            m.fft_plan = plans[i]
        m.run(warm)
        seq += time_steps(m, steps)
        if i % 16 == 15:
            log(f"ensemble ({'mixed' if precision else 'FP64'}): {i + 1} of {n} members run one after another")
        m.close()
    ens.run(1, use_graph=True)   # capture
    ens.synchronize()
    t0 = time.perf_counter()
    ens.run(steps, use_graph=True)
    ens.synchronize()
    bat = time.perf_counter() - t0
    # the first and last members against single models that take the same
    # path: warm steps before the tuning, the tuned plans for the rest
    same = True
    for i in (0, n - 1):
        prm, fields = members[i]
        m = Model(prm)
        for k, a in fields.items():
            m.f[k][...] = a
        m.push()
        m.run(warm)
        if precision:
            m.fft_plan = plans[i]
        m.run(1 + steps)
        m.pull(["u", "w", "p"])
        for k in ("u", "w", "p"):
            same = same and bool(np.array_equal(m.f[k], ens.get(i, k)))
        m.close()
    desc = ens.describe()
    ens.close()
    v_seq = n * steps / seq
    v_bat = n * steps / bat
    return {"value": v_bat, "unit": "member-steps/s", "members": n, "steps": steps,
            "sequential_value": v_seq, "speedup_vs_sequential": v_bat / v_seq,
            "ms_per_ensemble_step": 1e3 * bat / steps, "sequential_ms_per_ensemble_step": 1e3 * seq / steps,
            "points_per_ensemble_step": int(sum(p.npoints for p, _ in members)),
            "bitwise_equal_to_single_runs": same, "tuning_s": tune_s,
            "tiles": desc.strip().splitlines(),
            "dtype": "f32 stencils/FFT, f64 solve + accumulation" if precision else "f64",
            "workload": "16x(64x64x32) + 16x(96x96x48) + 16x(128x128x64) baroclinic eddies, rigid lid, AB3, "
                        "horizontal biharmonic friction; one launch per stage, hipGraph replay"}


def configs_leg(steps: int = 50, warm: int = 5, cases=("eddy", "lee_wave", "jets_eddies")) -> dict:
    """Throughput of the other configurations, each form from the same
    initial state after warm-up steps and tuning: the 64x64x32 baroclinic
    eddy, the 512x128x128 lee wave (background flow and body force on the
    device), the 1024x1024x128 jets and eddies and the 256x256x64 jets at
    0.5 m/s with noise on every field."""
    from ps3d_amd import configs
    from ps3d_amd.model import Model
    from ps3d_amd.params import PS3D_FP64, PS3D_MIXED
    table = {"eddy": ("eddy 64x64x32, rigid lid, AB3, biharmonic friction",
                      lambda: configs.eddy(64, 64, 32, seed=1) + (None,)),
             "lee_wave": ("lee wave 512x128x128, rigid lid, background flow + body force",
                          lambda: configs.lee_wave(512, 128, 128, seed=3)),
             "jets_eddies": ("jets and eddies 1024x1024x128, rigid lid",
                             lambda: configs.jets_eddies(1024, 1024, 128, seed=5) + (None,)),
             "jets_c2": ("jets 256x256x64 at 0.5 m/s over Ly/10, noise on u, v, w, b, N0 raised for stability",
                         lambda: configs.jets_c2(256, 256, 64) + (None,))}
    out = {}
    for key in cases:
        name, make = table[key]
        prm, fields, forcing = make()
        row = {"grid": [prm.nx, prm.ny, prm.nz]}
        for tag, prec in (("fp64", PS3D_FP64), ("mixed", PS3D_MIXED)):
            m = Model(prm.replace(precision=prec))
            for n, a in fields.items():
                m.f[n][...] = a
            m.push()
            if forcing is not None:
                m.set_background_flow(forcing["u_back"])
                m.set_body_force(forcing["force"])
            m.run(warm)
            m.tune(reps=3)
            el = time_steps(m, steps)
            m.close()
            row[tag] = {"grid_pt_steps_per_s": prm.npoints * steps / el, "ms_per_step": 1e3 * el / steps}
            print(f"configs: {name} {tag}: {1e3 * el / steps:.3f} ms per step", file=sys.stderr, flush=True)
        row["mixed_speedup"] = row["mixed"]["grid_pt_steps_per_s"] / row["fp64"]["grid_pt_steps_per_s"]
        out[name] = row
        del fields
    return out


def cpu_baseline(prm, fields, budget_s: float = 15.0, threads: int = 0) -> dict:
    """The C oracle stepping the same workload on the host: one independent
    instance per thread, all stepped at once (ctypes releases the GIL inside
    each step), so the aggregate is the host's throughput on this workload
    with every core busy -- what a one-rank-per-core run could reach at best.
    The single-thread rate is reported beside it."""
    import threading
    import oracle
    if threads <= 0:
        try:
            threads = len(os.sched_getaffinity(0))
        except AttributeError:
            threads = os.cpu_count() or 1
        threads = max(1, min(16, threads))   # the host share of one GPU
    models = []
    for _ in range(threads):
        orc = oracle.OracleModel(prm)
        for n, a in fields.items():
            orc.f[n][...] = a
        models.append(orc)
    steps = [0] * threads
    elapsed = [0.0] * threads
    start = threading.Barrier(threads)

    def work(q):
        orc = models[q]
        orc.run(1)   # untimed first step (Euler)
        start.wait()
        t0 = time.perf_counter()
        while True:
            orc.run(1)
            steps[q] += 1
            el = time.perf_counter() - t0
            if el >= budget_s or steps[q] >= 200:
                break
        elapsed[q] = el

    pool = [threading.Thread(target=work, args=(q,)) for q in range(threads)]
    for t in pool:
        t.start()
    for t in pool:
        t.join()
    for orc in models:
        orc.close()
    wall = max(elapsed)
    total = prm.npoints * sum(steps) / wall
    return {"value": total, "unit": "grid-pt steps/s", "cores": threads, "kind": "port",
            "per_thread_value": total / threads,
            "sample": f"{threads} independent instances of the same 256x256x64 jets workload stepped at once, "
                      f"one thread each, {sum(steps)} steps in all (after 1 untimed step each), C oracle "
                      f"oracle/ps3d_oracle.c, {wall:.1f} s"}


def _reference_worker(args):
    """One process: the reference's own integrate.f90 (compiled unchanged into
    oracle/_ref/libps3d_ref.so; module globals, so one model per process)
    stepped up to its pressure solve, and the port doing the same part, on
    the same state; returns (steps, seconds) of each."""
    budget_s, small = args
    import oracle
    if small:   # a test of this leg on a small grid
        from ps3d_amd import configs
        prm, fields = configs.jets(32, 32, 16, seed=2)
    else:
        prm, fields = build_workload()
    ref = oracle.ReferenceModel(prm)
    orc = oracle.OracleModel(prm)
    for n, a in fields.items():
        ref.f[n][...] = a
        orc.f[n][...] = a
    out = {}
    for tag, mdl in (("reference", ref), ("port", orc)):
        mdl.integrate_pre()   # untimed first step (Euler)
        mdl.rotate()
        n, t0 = 0, time.perf_counter()
        while True:
            mdl.integrate_pre()
            mdl.rotate()
            n += 1
            el = time.perf_counter() - t0
            if el >= budget_s or n >= 100:
                break
        out[tag] = (n, el)
    # the port's pressure solve on its current state
    n, t0 = 0, time.perf_counter()
    while True:
        orc.pressure_solver()
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s / 2 or n >= 100:
            break
    out["port_pressure"] = (n, el)
    orc.close()
    return out


def reference_cpu_leg(prm, procs: int, budget_s: float = 4.0, small: bool = False) -> dict | None:
    """The reference's own Fortran timed on the host beside the port: its
    integrate.f90 up to the pressure solve (pressure_solver.f90 and my_fft.f90
    need P3DFFT, absent from the image, so they are not built) and the port's
    pressure solve, one process per core, all at once."""
    import multiprocessing as mp
    import oracle
    if not oracle.reference_available():
        return None
    import resource
    from concurrent.futures import ProcessPoolExecutor, as_completed
    # the friction routines keep grid-sized automatic arrays (mixing.f90:11-12,
    # ~75 MB at 256x256x64), which amdflang places on the stack: the workers
    # are started with the stack limit raised to the hard limit (a process's
    # main stack is sized from the limit it starts with)
    soft, hard = resource.getrlimit(resource.RLIMIT_STACK)
    need = 256 << 20
    if hard != resource.RLIM_INFINITY and hard < need and not small:
        log(f"CPU baseline: stack hard limit {hard} bytes is below what the reference needs; skipped")
        return None
    resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))
    try:
        # a pool that reports a worker's death (multiprocessing.Pool would wait
        # for it for ever)
        with ProcessPoolExecutor(procs, mp_context=mp.get_context("spawn")) as pool:
            futs = [pool.submit(_reference_worker, (budget_s, small)) for _ in range(procs)]
            res = []
            for f in as_completed(futs):
                res.append(f.result())
                log(f"CPU baseline (reference): {len(res)} of {procs} processes done")
    finally:
        resource.setrlimit(resource.RLIMIT_STACK, (soft, hard))
    N = prm.npoints

    def rate(tag):
        return sum(N * r[tag][0] / r[tag][1] for r in res)

    ref_pre, port_pre, port_p = rate("reference"), rate("port"), rate("port_pressure")
    # per-point seconds of the parts add up to a step
    step = 1.0 / (1.0 / ref_pre + 1.0 / port_p)
    return {"value": step, "unit": "grid-pt steps/s", "cores": procs, "kind": "reference",
            "integrate_before_pressure": {"reference": ref_pre, "port": port_pre,
                                          "port_over_reference": port_pre / ref_pre},
            "pressure_solve_port": port_p,
            "sample": f"{procs} processes at once, each one {prm.nx}x{prm.ny}x{prm.nz} jets model: the reference's own "
                      f"integrate.f90 (for_src/, amdflang -O2, unchanged; oracle/_ref) up to its pressure "
                      f"solve, and the C port's pressure solve (the reference's needs P3DFFT, absent), "
                      f"~{budget_s:.0f} s each; value = aggregate step rate of reference part + port "
                      f"pressure solve"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the CPU baseline (0: the cores available, at most 16)")
    ap.add_argument("--profile-all", action="store_true", help="print every kernel's share")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the ensemble leg")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--ens-steps", type=int, default=100, help="timed steps of the ensemble leg")
    ap.add_argument("--configs", nargs="?", const="eddy,lee_wave,jets_eddies,jets_c2", default=None,
                    help="also time the eddy, lee-wave, 1024x1024x128 and 0.5 m/s jets configurations (both "
                         "forms); optionally a comma list of them")
    ap.add_argument("--config-steps", type=int, default=50, help="timed steps of each --configs case")
    ap.add_argument("--mixed-check", default="1,100,1000",
                    help="steps at which the mixed leg reports its error against FP64 (comma list)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run the streaming probe kernels on MIB MiB (to calibrate HBM byte counters)")
    ap.add_argument("--tendency-form", choices=("default", "separate", "fused"), default="default",
                    help="tendencies in a kernel of their own (separate, the default) or formed inside the "
                         "friction + AB kernel (fused)")
    args = ap.parse_args()
    if args.copy_probe:
        # the library's streaming kernels only (k_copy moves 2 x MIB MiB per
        # dispatch): the known-size kernels that calibrate the byte counters
        print(json.dumps({"hbm_probe": hbm_probe(args.copy_probe, reps=2),
                          "copy_bytes_per_dispatch": 2 * (args.copy_probe << 20)}), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py: this build runs the whole domain on one GPU; --gpus {args.gpus} is not "
              "supported (members of an ensemble or z-slabs of one grid would be sharded across GPUs, "
              "which is not part of this build)", file=sys.stderr)
        return 2

    from ps3d_amd.model import Model, kernel_names

    log("HBM probe")
    hbm = hbm_probe()
    log(f"probe copy {hbm['ceiling']:.0f} GB/s; building the 256x256x64 jets workload")
    prm, fields = build_workload()
    model = Model(prm)
    if args.tendency_form != "default":
        model.set_fused_tendency({"separate": 0, "fused": 1}[args.tendency_form])
    for n, a in fields.items():
        model.f[n][...] = a
    model.push()
    model.run(max(args.warmup, 3))   # past Euler and AB2: the timed steps are AB3
    if not args.no_tune:             # launch-shape autotuner (ps3d_tune), untimed set-up
        model.tune(reps=5)
    model.synchronize()

    # share of each kernel over a few untimed steps (one kernel at a time)
    shares = {}
    probe = 10
    for name in kernel_names():
        model.profile_kernel(name)
        model.run(probe)
        n, ms = model.profile_read()
        if n:
            shares[name] = (n / probe, ms / probe)
    model.profile_kernel(None)
    dominant = max(shares, key=lambda k: shares[k][1])

    log(f"kernel shares measured; timing {args.steps} steps")
    # timed region: exactly --steps steps, the dominant kernel under events
    model.profile_kernel(dominant)
    el = time_steps(model, args.steps)
    launches, kms = model.profile_read()
    model.profile_kernel(None)

    value = prm.npoints * args.steps / el
    per_launch_s = kms / launches / 1e3
    traffic_bytes = algorithmic_bytes(dominant, prm)
    achieved = traffic_bytes / per_launch_s / 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "grid-pt steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * el / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "baroclinic zonal jets 256x256x64, rigid lid, FP64, AB3, "
                               "harmonic + horizontal biharmonic friction",
                   "grid": [prm.nx, prm.ny, prm.nz], "precision": "fp64", "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dominant),
                     "launches": launches, "mean_launch_us": per_launch_s * 1e6,
                     "algorithmic_bytes_per_launch": traffic_bytes,
                     "achievable": hbm["ceiling"], "frac_of_achievable": achieved / hbm["ceiling"],
                     "probe": hbm},
    }
    if args.profile_all:
        out["kernel_shares_ms_per_step"] = {k: v[1] for k, v in sorted(shares.items(), key=lambda x: -x[1][1])}
    out["tuning"] = model.tuning().strip().splitlines()
    # the same steps replayed from ps3d_run's captured six-step hipGraph
    model.run(12, use_graph=True)   # capture
    el_graph = time_steps_graph(model, args.steps)
    out["graph_replay"] = {"ms_per_step": 1e3 * el_graph / args.steps,
                           "value": prm.npoints * args.steps / el_graph, "unit": "grid-pt steps/s"}
    model.close()
    log(f"FP64: {1e3 * el / args.steps:.4f} ms per step ({1e3 * el_graph / args.steps:.4f} as graph replays)")
    if not args.no_mixed:
        checks = tuple(int(x) for x in args.mixed_check.split(",") if x.strip())
        out["mixed"] = mixed_leg(prm, fields, args.steps, args.warmup, value, check_steps=checks)
        # the rate the FP32 stencils issue at, beside FP64's
        out["mixed"]["vector_issue"] = fma_probe()
        log(f"mixed: {out['mixed']['ms_per_step']:.4f} ms per step, worst rel L2 {out['mixed']['max_rel_l2']:.2e}")
    if not args.no_ensemble:
        out["ensemble"] = ensemble_leg(args.ens_steps)
        log(f"ensemble FP64: {out['ensemble']['speedup_vs_sequential']:.2f}x sequential")
        if not args.no_mixed:
            out["ensemble_mixed"] = ensemble_leg(args.ens_steps, precision=1)
            log(f"ensemble mixed: {out['ensemble_mixed']['speedup_vs_sequential']:.2f}x sequential")
    if args.configs:
        out["configs"] = configs_leg(steps=args.config_steps,
                                     cases=tuple(c.strip() for c in args.configs.split(",") if c.strip()))
    if not args.no_cpu:
        log("CPU baseline: the C port on the host cores")
        port = cpu_baseline(prm, fields, threads=args.cpu_threads)
        log("CPU baseline: the reference's integrate.f90 on the host cores")
        ref = reference_cpu_leg(prm, port["cores"])
        if ref is not None:
            # the reference's own code where it builds; the port's whole step beside it
            ref["port_whole_step"] = port
            out["cpu_baseline"] = ref
        else:
            out["cpu_baseline"] = port
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
