#!/usr/bin/env python3
"""Throughput of the self-consistency iteration on one MI355X.

A step is one self-consistency iteration (sc_compute: accelerator, Riccati
sweeps over every angle and Ozaki energy fused with the Green-function sums,
angle integration, gap equation, current, free energy, specular boundary
condition, FFT vector-potential solve) on the d-wave octagon, with every
input already resident in HBM.  The metric is Riccati trajectory steps per
second: S = 2 Ne #{sites inside the grain in each frame} summed over frames,
per iteration, times iterations, over the wall time of the timed iterations.

Prints one JSON line.  Besides the FP64 line it reports the mixed path
(FP32 stepping, FP64 sums) with its error against the FP64 path, the
roofline of the dominant kernel (the Riccati sweep, timed per launch with HIP
events on the stream it runs on) and the CPU oracle timed on a bounded
sample of the same workload on this host.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload octagon_dwave]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic FP64 flops per trajectory step of the sweep kernel, counted
# from the arithmetic of the step and of the Green-function sums (square
# root and reciprocal count one flop each):
#   midpoint step (RiccatiSolutions.h:65-94): 67 add/mul + 2 sqrt + 2 rcp = 71
#   Green sums per site and energy (LocalGreens.h:32-70,
#   ComputeFreeEnergy.h:95-160): 58 add/mul + 1 rcp = 59, shared by the
#   gamma and gamma~ steps of that site: 29.5 per trajectory step.
STEP_FLOPS = 71.0
GREEN_FLOPS_PER_SITE_ENERGY = 59.0
FLOPS_PER_TRAJ_STEP = STEP_FLOPS + 0.5 * GREEN_FLOPS_PER_SITE_ENERGY
# Real-frequency (Keldysh) sweep of the LDOS: the step plus 2 for the real
# energy part (scale, add), and the spectral weight (1 - |p|^2) / |1 + p|^2
# per site and energy: 15 add/mul + 1 rcp = 16, shared by the two steps.
KELDYSH_STEP_FLOPS = STEP_FLOPS + 2.0
SPECTRAL_FLOPS_PER_SITE_ENERGY = 16.0
KELDYSH_FLOPS_PER_TRAJ_STEP = KELDYSH_STEP_FLOPS + 0.5 * SPECTRAL_FLOPS_PER_SITE_ENERGY
PEAK_FP64_TFLOPS = 78.6   # MI355X vector FP64
PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 (packed)
PEAK_HBM_GBS = 8000.0
METRIC = "Riccati traj-steps/s; wall time to 1e-8 residual, mixed/NK vs plain FP64"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed_iterations(sim, precision, warmup, steps):
    for _ in range(warmup):
        sim.compute(precision)
    # sc_compute returns after its stream has drained (it reads the residual
    # back), so each call is a synchronisation point on both sides.
    sim.results()
    sweep = []
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.compute(precision)
        sweep.append(sim.results()["last_sweep_ms"])
    t1 = time.perf_counter()
    return (t1 - t0), np.array(sweep)


def roofline(traj_steps, sweep_ms, peak, flops_per_step=FLOPS_PER_TRAJ_STEP):
    t = float(np.mean(sweep_ms)) * 1e-3
    achieved = traj_steps * flops_per_step / t / 1e12
    return {"bound": "valu", "achieved": round(achieved, 3), "peak": peak,
            "unit": "TFLOP/s", "frac": round(achieved / peak, 4),
            "traffic": None, "kernel": "sweepKernel",
            "flops_per_traj_step": flops_per_step,
            "traj_steps_per_launch": int(traj_steps),
            "avg_launch_ms": round(float(np.mean(sweep_ms)), 4)}


def load_traffic(workload, precision, shape):
    """HBM bytes per sweep launch from the committed rocprofv3 counter runs,
    when one of them was taken on the sweep variant and shape (energies per
    lane, waves per SIMD, slots per CU) that this run timed; None otherwise."""
    path = os.path.join(ROOT, "profiles", "sweep_%s_%s_traffic.json" % (workload, precision))
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    want = [int(v) for v in shape]
    for r in rec.get("records", [rec]):
        if [int(v) for v in r.get("shape", [])] == want:
            return r.get("bytes_per_launch")
    log("traffic: no counter run of %s %s at shape %s: traffic left null"
        % (workload, precision, want))
    return None


def copy_probe(mib):
    """One device-to-device copy of `mib` MiB through the HIP runtime: the
    reference kernel of known traffic for scaling the HBM counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = int(mib) << 20
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, ctypes.c_size_t(n))
    hip.hipDeviceSynchronize()
    if hip.hipMemcpy(dst, src, ctypes.c_size_t(n), 3) != 0:  # device to device
        raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    log("copy probe: %d MiB device-to-device" % mib)


def start(sa, workloads, cfg, name):
    """A grain on the workload's initial state: the product's seeded guess,
    with the d + is square's s component set to i 0.2 of its d component."""
    sim = sa.Simulation(cfg)
    if name == "dplus_is_square":
        d = sim.order_parameter(0)
        for c, x in enumerate(workloads.d_plus_is_initial(d[0])):
            sim.set_order_parameter(c, x)
    return sim


def converge(sa, cfg, precision, cap, workloads=None, name=""):
    """Wall time of one solve to the config's criterion (sc_run) from the
    initial state, with the sweep shape autotuned first (outside the timed
    region)."""
    sim = start(sa, workloads, cfg, name)
    # The autotuner's sweeps leave the state as it is.
    shape64 = sim.autotune("fp64")
    shape32 = sim.autotune("mixed")
    t = time.perf_counter()
    ok = sim.run(0, 0, cap, precision)
    wall = time.perf_counter() - t
    r = sim.results()
    sim.close()
    return {"converged": bool(ok), "iterations": r["iterations"],
            "fp32_iterations": r["mixed_iterations"], "wall_s": round(wall, 4),
            "residual": r["residual"], "free_energy_per_area": r["free_energy_per_area"],
            "residual_vector_potential": r["residual_vector_potential"],
            "sweep_shapes": [list(shape64), list(shape32)]}


# Newton-Krylov stops at this fraction of the convergence criterion in its
# block-weighted norm (each block's relative residual is then below it); the
# FP64 iterations after it apply the reference's own test.
NEWTON_TOL = 0.1
# GMRES restarted every NEWTON_KRYLOV_DIM directions, at most NEWTON_CYCLES
# cycles per Newton step (the batch of small grains: 40 directions).  On the
# d + is square, 60 x 2 takes 297 maps where 30 x 4 takes 460 (FP64
# products), 389 against 685 with FP32 products.
NEWTON_KRYLOV_DIM = 60
NEWTON_CYCLES = 2


def converge_newton(sa, cfg, cap, warm_residual=1e-3, warm_max=30, jvp="fp64",
                    workloads=None, name=""):
    """Wall time to the config's criterion with Newton-Krylov: FP32-stepping
    iterations until the residual is below warm_residual (at most warm_max),
    Jacobian-free Newton-Krylov on the FP64 map (FP32 products switch to
    FP64 ones when a Newton step stalls), then FP64 iterations until the
    reference's convergence test passes."""
    sim = start(sa, workloads, cfg, name)
    sim.autotune("fp64")
    sim.autotune("mixed")
    t = time.perf_counter()
    warm = 0
    for warm in range(1, warm_max + 1):
        sim.compute("mixed")
        if sim.results()["residual"] < warm_residual:
            break
    sim.set_newton_options(restarts=NEWTON_CYCLES)
    rep = sim.newton_krylov(max_newton=20, krylov_dim=NEWTON_KRYLOV_DIM,
                            tol=NEWTON_TOL * cfg["numerics"]["convergence_criterion"],
                            jvp_precision=jvp)
    ok = sim.run(0, 0, cap, "fp64")
    wall = time.perf_counter() - t
    r = sim.results()
    sim.close()
    out = {"converged": bool(ok), "fp32_iterations": warm,
           "newton_iterations": rep["newton_iterations"],
           "krylov_iterations": rep["krylov_iterations"],
           "map_evaluations": rep["map_evaluations"], "newton_residual": rep["residual"],
           "fp64_iterations_after": r["iterations"] - warm, "wall_s": round(wall, 4),
           "residual": r["residual"], "free_energy_per_area": r["free_energy_per_area"]}
    return out


def convergence(sa, workloads, args, name=None):
    """Wall time to a 1e-8 residual from the same initial state: plain FP64
    iterations with the config's accelerator and with CongAcc (the
    reference's default), mixed refinement, Newton-Krylov with FP64 and with
    FP32 finite-difference Jacobian-vector products and with exact ones (the
    tangent-linear sweep).  Speedups are against the faster plain
    FP64 run that converged; when neither did, against the capped time of
    the faster one, as a lower bound."""
    name = name or args.converge_workload
    ccfg = workloads.WORKLOADS[name]()
    ccfg["numerics"]["convergence_criterion"] = 1e-8
    conv = {"workload": name, "criterion": 1e-8, "cap_iterations": args.converge_cap,
            "accelerator": ccfg["accelerator"]["name"]}
    plain = ["fp64"]
    conv["fp64"] = converge(sa, ccfg, "fp64", args.converge_cap, workloads, name)
    log("converge %s fp64: %s" % (name, conv["fp64"]))
    if ccfg["accelerator"]["name"] != "congacc":
        cc = dict(ccfg, accelerator={"name": "congacc"})
        conv["fp64_congacc"] = converge(sa, cc, "fp64", args.converge_cap, workloads, name)
        log("converge %s fp64 congacc: %s" % (name, conv["fp64_congacc"]))
        plain.append("fp64_congacc")
    done = [conv[k]["wall_s"] for k in plain if conv[k]["converged"]]
    base = min(done) if done else min(conv[k]["wall_s"] for k in plain)
    suffix = "" if done else "_lower_bound"
    conv["mixed_refine"] = converge(sa, ccfg, "mixed_refine", args.converge_cap, workloads, name)
    log("converge %s mixed_refine: %s" % (name, conv["mixed_refine"]))
    methods = ["mixed_refine"]
    if not args.no_newton:
        for key, jvp in (("newton_krylov", "fp64"), ("newton_krylov_mixed", "mixed"),
                         ("newton_krylov_tangent", "tangent")):
            conv[key] = converge_newton(sa, ccfg, args.converge_cap, jvp=jvp,
                                        workloads=workloads, name=name)
            log("converge %s %s: %s" % (name, key, conv[key]))
            methods.append(key)
    for key in methods:
        if conv[key]["converged"]:
            conv["speedup_" + key + suffix] = base / conv[key]["wall_s"]
    return conv


def octagon_convergence(sa, workloads, args):
    """Wall time towards the 1e-8 criterion on the d-wave octagon itself,
    each method from the same seeded start and capped: plain FP64 with the
    config's Anderson(8) and with CongAcc (the reference's default
    accelerator), mixed refinement with CongAcc, and Newton-Krylov with FP32
    Jacobian-vector products after FP32 CongAcc iterations.  A method that
    stops at its cap reports converged false and the residual it reached,
    with the residual of each field (the vector potential's is the one that
    stays highest)."""
    base = workloads.octagon_dwave()
    cap = args.octagon_converge_cap
    res = {"workload": "octagon_dwave", "criterion": 1e-8, "cap_iterations": cap}
    runs = (("fp64_anderson", "anderson", "fp64"), ("fp64_congacc", "congacc", "fp64"),
            ("mixed_refine_congacc", "congacc", "mixed_refine"))
    for key, acc, prec in runs:
        cfg = workloads.octagon_dwave()
        cfg["accelerator"] = dict(base["accelerator"]) if acc == "anderson" else {"name": acc}
        res[key] = converge(sa, cfg, prec, cap)
        log("octagon converge %s: %s" % (key, res[key]))
    if not args.no_newton:
        cfg = workloads.octagon_dwave()
        cfg["accelerator"] = {"name": "congacc"}
        sim = sa.Simulation(cfg)
        sim.autotune("fp64")
        sim.autotune("mixed")
        t = time.perf_counter()
        for _ in range(cap):
            sim.compute("mixed")
        warm = sim.results()["residual"]
        sim.set_newton_options(restarts=NEWTON_CYCLES)
        rep = sim.newton_krylov(max_newton=8, krylov_dim=NEWTON_KRYLOV_DIM,
                                tol=NEWTON_TOL * 1e-8, jvp_precision="mixed")
        ok = sim.run(0, 0, 5, "fp64")
        wall = time.perf_counter() - t
        r = sim.results()
        sim.close()
        res["newton_krylov_mixed"] = {
            "converged": bool(ok), "fp32_iterations": cap, "residual_after_fp32": warm,
            "newton_iterations": rep["newton_iterations"],
            "krylov_iterations": rep["krylov_iterations"],
            "map_evaluations": rep["map_evaluations"], "newton_residual": rep["residual"],
            "wall_s": round(wall, 3), "residual": r["residual"],
            "residuals": {k: r[k] for k in ("residual_order_parameter", "residual_current",
                                             "residual_vector_potential", "residual_boundary",
                                             "residual_free_energy")}}
        log("octagon converge newton-krylov (FP32 products): %s" % res["newton_krylov_mixed"])
    return res


def ldos_bench(sa, workloads, args):
    """The LDOS postprocess of the workload's grain with the reference's
    example spectroscopy settings (examples/*/postprocess_config.json: 720
    angles, 128 energies in [0, 0.5] with broadening 0.008, blocks of 32):
    iterations of the first energy block, FP64."""
    cfg = workloads.WORKLOADS[args.workload]()
    post = workloads.postprocess_config()
    t = time.perf_counter()
    sp = sa.Spectroscopy(cfg, post)
    setup_s = time.perf_counter() - t
    sim = sa.Simulation(cfg)  # the initial order parameter of the grain
    delta = sim.order_parameter(0)
    sim.close()
    for c, d in enumerate(delta):
        sp.set_order_parameter(c, d)
    sp.compute("fp64")
    if args.ldos_shape:
        sp.set_sweep_shape(*(int(v) for v in args.ldos_shape.split(",")), precision="fp64")
        shape = sp.sweep_shape("fp64")
    else:
        shape = sp.autotune("fp64")
    log("ldos sweep shape fp64: %s" % (shape,))
    for _ in range(max(args.warmup - 1, 0)):
        sp.compute("fp64")
    sweeps = []
    t0 = time.perf_counter()
    for _ in range(args.ldos_steps):
        sp.compute("fp64")
        sweeps.append(sp.status()["last_sweep_ms"])
    wall = time.perf_counter() - t0
    st = sp.status()
    sp.close()
    steps = st["traj_steps"]
    rf = roofline(steps, sweeps, PEAK_FP64_TFLOPS, KELDYSH_FLOPS_PER_TRAJ_STEP)
    rf["kernel"] = "sweepKernel (Keldysh, summed over angle chunks)"
    return {"workload": args.workload + " LDOS postprocess",
            "num_angles": post["numerics"]["num_fermi_momenta"],
            "num_energies": post["spectroscopy"]["num_energies"],
            "energies_per_block": post["numerics"]["num_energies_per_block"],
            "angle_chunk": st["angle_chunk"], "setup_s": round(setup_s, 2),
            "value": steps * args.ldos_steps / wall, "unit": "traj-steps/s",
            "ms_per_iteration": 1e3 * wall / args.ldos_steps,
            "traj_steps_per_iteration": steps, "dtype": "f64", "roofline": rf,
            "sweep_shape": {"energies_per_lane": shape[0], "waves_per_simd": shape[1],
                            "slots_per_cu": shape[2]}}


def batch_bench(sa, workloads, args):
    """Batched parameter sweep: one iteration of every grain, the sweep of
    all of them in one persistent launch, against the same grains iterated
    one after the other."""
    cfgs = workloads.parameter_sweep(args.batch_count)
    t = time.perf_counter()
    b = sa.Batch(cfgs)
    log("batch of %d grains set up in %.1fs" % (len(cfgs), time.perf_counter() - t))
    b.compute("fp64")
    shape = b.autotune("fp64")
    log("batch autotune fp64: %s" % (shape,))
    for _ in range(max(args.warmup - 1, 0)):
        b.compute("fp64")
    sweeps, steps = [], 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.compute("fp64")
        sw, _it, n = b.timing()
        sweeps.append(sw)
        steps += n
    wall = time.perf_counter() - t0
    b.close()
    # the same grains one at a time
    seq = 0.0
    seq_steps = 0
    for c in cfgs:
        s = sa.Simulation(c)
        s.compute("fp64")
        s.autotune("fp64")
        t1 = time.perf_counter()
        s.compute("fp64")
        seq += time.perf_counter() - t1
        seq_steps += s.dims.traj_steps
        s.close()
    value = steps / wall
    return {
        "metric": METRIC + " (batched parameter sweep)", "value": value,
        "unit": "traj-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "parameter_sweep", "grains": len(cfgs),
                   "traj_steps_per_iteration": steps // max(args.steps, 1),
                   "step": "one self-consistency iteration of every grain"},
        "roofline": roofline(steps / max(args.steps, 1), np.array(sweeps), PEAK_FP64_TFLOPS),
        "sweep_shape": {"energies_per_lane": shape[0], "waves_per_simd": shape[1],
                        "slots_per_cu": shape[2]},
        "sequential": {"value": seq_steps / seq, "ms_per_step": 1e3 * seq,
                       "speedup_of_batch": (seq_steps / seq and value / (seq_steps / seq))},
    }


def batch_convergence(sa, workloads, args):
    """The parameter sweep solved to the 1e-8 criterion as one batch: every
    grain iterates until its own residual passes (sc_batch_run), with FP32
    stepping switching to FP64 per grain (mixed refinement) and, from the
    same seeded starts, in plain FP64.  Reports the wall time of each run and
    how many grains converged within the cap."""
    cfgs = workloads.parameter_sweep(args.batch_count)
    res = {"workload": "parameter_sweep", "grains": len(cfgs), "criterion": 1e-8,
           "cap_iterations": args.batch_converge_cap}
    for prec in (("mixed_refine", "fp64") if args.workload == "parameter_sweep" else ("fp64",)):
        b = sa.Batch(cfgs)
        b.autotune("fp64")
        b.autotune("mixed")
        t = time.perf_counter()
        conv = b.run(0, 0, args.batch_converge_cap, prec)
        wall = time.perf_counter() - t
        its = [b.grain(i).results()["iterations"] for i in range(len(b))]
        fp32 = [b.grain(i).results()["mixed_iterations"] for i in range(len(b))]
        res[prec] = {"wall_s": round(wall, 3), "converged_grains": int(sum(conv)),
                     "iterations_max": int(max(its)), "iterations_total": int(sum(its)),
                     "fp32_iterations_total": int(sum(fp32))}
        b.close()
        log("batch converge %s: %s" % (prec, res[prec]))
    if "mixed_refine" in res and res["fp64"]["converged_grains"] == len(cfgs) and \
            res["mixed_refine"]["converged_grains"] == len(cfgs):
        res["speedup_mixed_refine"] = res["fp64"]["wall_s"] / res["mixed_refine"]["wall_s"]
    if not args.no_newton:
        # plain iterations (FP32 stepping, then FP64 per grain) until most
        # grains have converged, then every grain's Newton-Krylov solve with
        # FP32 products on the grains still iterating, in shared sweep
        # launches, then FP64 iterations until the reference's test passes
        b = sa.Batch(cfgs)
        b.autotune("fp64")
        b.autotune("mixed")
        t = time.perf_counter()
        conv0 = b.run(0, 0, args.batch_newton_warm, "mixed_refine")
        left = [i for i, c in enumerate(conv0) if not c]
        reps = []
        if left:
            # FP64 iterations follow: a grain whose FP32 products find no
            # descent leaves the Newton solve to them rather than going on
            # with FP64 products
            for i in left:
                b.grain(i).set_newton_options(restarts=NEWTON_CYCLES, fp64_fallback=False)
            reps = b.newton_krylov(max_newton=12, krylov_dim=40, tol=NEWTON_TOL * 1e-8,
                                   jvp_precision="mixed", grains=left)
        conv = b.run(0, 0, 30, "fp64")
        wall = time.perf_counter() - t
        its = [b.grain(i).results()["iterations"] for i in left]
        res["newton_krylov_mixed"] = {
            "wall_s": round(wall, 3), "converged_grains": int(sum(conv)),
            "plain_iterations": args.batch_newton_warm,
            "converged_by_plain_iterations": int(sum(conv0)),
            "newton_grains": len(left),
            "newton_converged_grains": int(sum(reps[i]["converged"] for i in left)),
            "newton_iterations_max": int(max([reps[i]["newton_iterations"] for i in left] or [0])),
            "map_evaluations_max": int(max([reps[i]["map_evaluations"] for i in left] or [0])),
            "fp64_iterations_after_max": int(max(its or [0]))}
        b.close()
        log("batch converge newton-krylov (FP32 products): %s" % res["newton_krylov_mixed"])
        if res["newton_krylov_mixed"]["converged_grains"] == len(cfgs):
            base = res["fp64"]
            key = "speedup_newton_krylov_mixed" if base["converged_grains"] == len(cfgs) \
                else "speedup_newton_krylov_mixed_lower_bound"
            res[key] = base["wall_s"] / res["newton_krylov_mixed"]["wall_s"]
    return res


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def nproc():
    """The CPUs this process may run on (what `nproc` prints)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "host CPU"


def cpu_swave_disc(workloads, threads, criterion=1e-6, cap=400):
    """The oracle on the whole s-wave disc: seconds per full iteration and
    the wall time to the L2 criterion 1e-6 (Picard) from the seeded initial
    state the GPU leg starts from."""
    import superconga_amd as sa
    from oracle import pyoracle
    cfg = workloads.swave_disc()
    cfg["numerics"]["convergence_criterion"] = criterion
    sim = sa.Simulation(cfg)
    d0 = sim.order_parameter(0)
    sim.close()
    o = pyoracle.Oracle(dict(cfg, num_threads=threads))
    for c, d in enumerate(d0):
        o.set_order_parameter(c, d)
    t = time.perf_counter()
    ok = o.run(0, 0, cap)
    wall = time.perf_counter() - t
    its = o.get_int("num_iterations")
    return {"workload": "swave_disc", "criterion": criterion, "converged": bool(ok),
            "iterations": its, "time_to_criterion_s": round(wall, 3),
            "seconds_per_iteration": round(wall / max(its, 1), 4)}


def gpu_swave_disc(sa, workloads, criterion=1e-6, cap=400):
    """The same solve (s-wave disc, Picard, L2 criterion 1e-6, same start) on
    the GPU in FP64, sweep shape autotuned outside the timed region."""
    cfg = workloads.swave_disc()
    cfg["numerics"]["convergence_criterion"] = criterion
    r = converge(sa, cfg, "fp64", cap)
    r["workload"] = "swave_disc"
    r["criterion"] = criterion
    r["seconds_per_iteration"] = round(r["wall_s"] / max(r["iterations"], 1), 6)
    return r


def cpu_baseline(cfg, seconds, threads):
    """The oracle (FP64 OpenMP restatement of the reference's CPU path) on
    a bounded sample: the per-angle iteration work for consecutive angles
    until `seconds` of CPU time have been spent."""
    from oracle import pyoracle
    c = dict(cfg, num_threads=threads)
    t = time.perf_counter()
    o = pyoracle.Oracle(c)
    o.initial_bulk()
    o.sweep_angles(0, 1)  # boundary initial condition and first angle, untimed
    log("cpu oracle ready in %.1fs" % (time.perf_counter() - t))
    steps, spent, k = 0, 0.0, 1
    while k < o.Na and (spent < seconds or k < 3):
        t0 = time.perf_counter()
        steps += o.sweep_angles(k, k + 1)
        spent += time.perf_counter() - t0
        k += 1
    return {"value": steps / spent, "unit": "traj-steps/s", "cores": threads,
            "kind": "port",
            "sample": "angles 1-%d of %d of the same grain (rotation, energy shift, "
                      "both Riccati sweeps, Green functions, angle integration; "
                      "%d traj-steps in %.1f s), FP64, %d OpenMP threads (nproc) on %s"
                      % (k - 1, o.Na, steps, spent, threads, cpu_model()),
            # the angle work of one whole iteration at the sample's rate (the
            # vector-potential solve and the accelerator not included)
            "seconds_per_iteration_angles": round(spent / max(k - 1, 1) * o.Na, 2)}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="octagon_dwave")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--shape", default="",
                    help="fixed sweep shape 'energies,waves,slots' for both "
                         "precisions instead of the autotuner")
    ap.add_argument("--shape-mixed", default="",
                    help="fixed sweep shape of the mixed path (default: --shape)")
    ap.add_argument("--ldos-only", action="store_true",
                    help="only the LDOS postprocess measurement")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="run one device-to-device copy of this size first "
                         "(calibrates HBM counters of a profiler run)")
    ap.add_argument("--no-converge", action="store_true")
    ap.add_argument("--converge-workload", default="swave_disc",
                    help="workload of the time-to-criterion comparison")
    ap.add_argument("--converge-cap", type=int, default=2000)
    ap.add_argument("--batch-count", type=int, default=48)
    ap.add_argument("--no-ldos", action="store_true")
    ap.add_argument("--ldos-steps", type=int, default=3)
    ap.add_argument("--ldos-shape", default="", help="fixed LDOS sweep shape")
    ap.add_argument("--no-newton", action="store_true")
    ap.add_argument("--converge-only", action="store_true",
                    help="only the time-to-1e-8 legs (for profiling them)")
    ap.add_argument("--cpu-threads", type=int, default=nproc(),
                    help="OpenMP threads of the CPU baseline (default: nproc)")
    ap.add_argument("--octagon-converge-cap", type=int, default=300,
                    help="iteration cap of each time-to-1e-8 method on the octagon")
    ap.add_argument("--no-octagon-converge", action="store_true")
    ap.add_argument("--batch-converge-cap", type=int, default=600,
                    help="iteration cap of the parameter sweep's batched solve to 1e-8")
    ap.add_argument("--no-batch-converge", action="store_true")
    ap.add_argument("--batch-newton-warm", type=int, default=150,
                    help="plain batch iterations before the batched Newton-Krylov solve")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if int(os.environ.get("RANK", "0")) != 0:
        sys.exit(0)

    import superconga_amd as sa
    from superconga_amd import workloads

    if args.copy_probe_mib > 0:
        copy_probe(args.copy_probe_mib)

    if args.workload == "parameter_sweep":
        out = batch_bench(sa, workloads, args)
        if not args.no_converge and not args.no_batch_converge:
            out["time_to_1e-8"] = batch_convergence(sa, workloads, args)
        print(json.dumps(out), flush=True)
        return
    if args.ldos_only:
        print(json.dumps(ldos_bench(sa, workloads, args)), flush=True)
        return
    if args.converge_only:
        print(json.dumps(convergence(sa, workloads, args)), flush=True)
        return

    cfg = workloads.WORKLOADS[args.workload]()
    t = time.perf_counter()
    sim = sa.Simulation(cfg)
    d = sim.dims
    log("setup %.1fs: N0=%d N1=%d Ne=%d Na=%d grain sites=%d traj-steps/iter=%d"
        % (time.perf_counter() - t, d.n0, d.n1, d.num_energies, d.num_angles,
           d.num_grain_sites, d.traj_steps))
    # First FP64 iteration from the initial state: the reference for the
    # mixed path's error.
    sim.compute("fp64")
    ref = (sim.order_parameter(1), sim.current(), sim.vector_potential(1),
           sim.results()["free_energy"])
    fixed = tuple(int(v) for v in args.shape.split(",")) if args.shape else None
    t = time.perf_counter()
    if fixed:
        sim.set_sweep_shape(*fixed, precision="fp64")
        shape = sim.sweep_shape("fp64")
    else:
        shape = sim.autotune("fp64")
    tune_s = time.perf_counter() - t
    log("autotune fp64: energies/lane %d, waves/SIMD %d, slots/CU %d (%.1fs)"
        % (shape + (tune_s,)))
    wall, sweep = timed_iterations(sim, "fp64", max(args.warmup - 1, 0), args.steps)
    value = d.traj_steps * args.steps / wall
    rf = roofline(d.traj_steps, sweep, PEAK_FP64_TFLOPS)
    rf["traffic"] = load_traffic(args.workload, "fp64", shape)
    log("fp64: %.2f ms/iter, sweep %.2f ms, %.3e traj-steps/s, %.2f TFLOP/s (%.1f%%)"
        % (1e3 * wall / args.steps, np.mean(sweep), value, rf["achieved"],
           100 * rf["frac"]))
    sim.close()

    out = {
        "metric": METRIC, "value": value, "unit": "traj-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "n0": d.n0, "n1": d.n1,
                   "num_energies": d.num_energies, "num_angles": d.num_angles,
                   "grain_sites": d.num_grain_sites,
                   "traj_steps_per_iteration": d.traj_steps,
                   "step": "one self-consistency iteration"},
        "roofline": rf,
        "sweep_shape": {"energies_per_lane": shape[0], "waves_per_simd": shape[1],
                        "slots_per_cu": shape[2], "autotune_s": round(tune_s, 2)},
    }

    if not args.no_mixed:
        sim = sa.Simulation(cfg)
        sim.compute("mixed")
        got = (sim.order_parameter(1), sim.current(), sim.vector_potential(1),
               sim.results()["free_energy"])
        mfixed = (tuple(int(v) for v in args.shape_mixed.split(","))
                  if args.shape_mixed else fixed)
        if mfixed:
            sim.set_sweep_shape(*mfixed, precision="mixed")
            mshape = sim.sweep_shape("mixed")
        else:
            mshape = sim.autotune("mixed")
        log("autotune mixed: energies/lane %d, waves/SIMD %d, slots/CU %d" % mshape)
        err = {"delta": rel(got[0], ref[0]),
               "current": rel(got[1], ref[1]) if np.any(ref[1]) else float(np.max(np.abs(got[1]))),
               "vector_potential": rel(got[2], ref[2]) if np.any(ref[2]) else float(np.max(np.abs(got[2]))),
               "free_energy": abs(got[3] - ref[3]) / abs(ref[3])}
        mwall, msweep = timed_iterations(sim, "mixed", max(args.warmup - 1, 0), args.steps)
        mval = d.traj_steps * args.steps / mwall
        mrf = roofline(d.traj_steps, msweep, PEAK_FP32_TFLOPS)
        mrf["traffic"] = load_traffic(args.workload, "mixed", mshape)
        log("mixed: %.2f ms/iter, sweep %.2f ms, %.3e traj-steps/s; error vs fp64 %s"
            % (1e3 * mwall / args.steps, np.mean(msweep), mval, err))
        out["mixed"] = {"value": mval, "unit": "traj-steps/s",
                        "ms_per_step": 1e3 * mwall / args.steps,
                        "dtype": "f32 stepping, f64 sums",
                        "error_vs_fp64_first_iteration": err, "roofline": mrf,
                        "sweep_shape": {"energies_per_lane": mshape[0],
                                        "waves_per_simd": mshape[1],
                                        "slots_per_cu": mshape[2]}}
        sim.close()

    if not args.no_converge:
        out["time_to_1e-8"] = convergence(sa, workloads, args)
        if args.workload == "octagon_dwave":
            out["time_to_1e-8_dplus_is_square"] = convergence(sa, workloads, args,
                                                              "dplus_is_square")

    if not args.no_ldos:
        out["ldos"] = ldos_bench(sa, workloads, args)
        log("ldos: %s" % out["ldos"])

    if not args.no_converge and not args.no_octagon_converge and \
            args.workload == "octagon_dwave":
        out["time_to_1e-8_octagon"] = octagon_convergence(sa, workloads, args)

    if not args.no_converge and not args.no_batch_converge and \
            args.workload == "octagon_dwave":
        out["time_to_1e-8_parameter_sweep"] = batch_convergence(sa, workloads, args)

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cfg, args.cpu_seconds, args.cpu_threads)
        log("cpu: %.3e traj-steps/s" % out["cpu_baseline"]["value"])
        out["cpu_baseline"]["swave_disc"] = cpu_swave_disc(workloads, args.cpu_threads)
        log("cpu swave_disc: %s" % out["cpu_baseline"]["swave_disc"])
        out["gpu_swave_disc"] = gpu_swave_disc(sa, workloads)
        log("gpu swave_disc: %s" % out["gpu_swave_disc"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
