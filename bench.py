#!/usr/bin/env python3
"""Throughput of the range-limited force path on one GPU.

    python bench.py [--gpus 1] [--steps 1000] [--warmup 50] [--workload 2]

A step is one full MD step of the 92,224-atom LJ + Coulomb system (108.86 x 108.86 x
77.76 A, rc = 12 A, 9 x 9 x 6 cells, dt = 1 fs): Motion_Update (velocity Verlet), rebin
(counting sort) and the half-shell force kernel with int64 fixed-point accumulation, all
resident in HBM.  Prints one JSON line: pair-evals/s (unordered in-cutoff pairs per step
over wall time per step), ns/day, the force kernel's roofline (HIP-event duration over the
timed steps), and a CPU baseline (the C model restatement, one thread, timed on a bounded
sample of the same system in the same run) together with the force error against an FP64
direct evaluation.
"""
import argparse
import copy
import json
import math
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from sc2019_md_amd import mdrl  # noqa: E402

METRIC = "pair-evals/s + ns/day at 92k atoms; batched vs plain, force err vs FP64"
PEAK_FP32_TFLOPS = 157.3          # MI355X FP32 vector peak (packed)
PEAK_HBM_GBS = 8000.0
FLOPS_PER_PAIR = {True: 29, False: 24}   # LJ + Coulomb / LJ only, FMA = 2

WORKLOADS = {
    1: "512-atom argon, 34 A box, rc 8.5 A, 4x4x4 cells, dt 2 fs",
    2: "92,224-atom LJ+Coulomb, 108.86x108.86x77.76 A, rc 12 A, 9x9x6 cells, dt 1 fs",
    3: "1,048,576-atom LJ+Coulomb, 219.0 A cube, rc 12 A, 18x18x18 cells, dt 1 fs",
    4: "512 replicas, log-uniform 256..16,384 atoms each, 0.1 atoms/A^3, rc 10 A, LJ+Coulomb, dt 1 fs",
    5: "128 replicas of 64 chains x 48 united-atom beads, bonded + LJ + Coulomb, dt 2 fs",
}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def cpu_baseline_leg(s, gpu_fix, pairs_per_step, budget_s, checks=None):
    """The C model restatement (oracle), one thread, on a bounded sample of the workload, and
    the GPU forces checked against the FP64 direct oracle.  The oracle supplies the reference
    results and the CPU baseline here, never the measured GPU path."""
    from oracle import oracle as O

    p = O.cmodel_default_params()
    p.box[:] = [float(b) for b in s.box]
    p.rc = s.rc
    part = np.zeros((s.n, 11), np.float32)
    part[:, :3] = s.particle[:, :3]
    nref = 200
    t0 = time.perf_counter()
    O.cmodel_force_sample(part, p, nref)
    dt = time.perf_counter() - t0
    nref2 = int(min(s.n, max(nref, nref * budget_s / max(dt, 1e-6))))
    t0 = time.perf_counter()
    O.cmodel_force_sample(part, p, nref2)
    dt2 = time.perf_counter() - t0
    step_s = dt2 * s.n / nref2
    base = {
        "value": pairs_per_step / step_s,
        "unit": "pair-evals/s",
        "cores": 1,
        "kind": "port",
        "sample": ("C-model restatement (all-pairs, pow() arithmetic, one thread, %s) over the "
                   "%d-atom positions with the workload's box and cutoff: reference atoms "
                   "0..%d against all atoms took %.2f s; the step time %.1f s is that x %d/%d, "
                   "with the GPU's per-step pair count as numerator; host has %d cores"
                   % (cpu_model(), s.n, nref2 - 1, dt2, step_s, s.n, nref2, os.cpu_count())),
        "ms_per_step": step_s * 1e3,
    }
    base["cmodel_agreement"] = cmodel_agreement(O)
    # force error vs the FP64 analytic oracle
    fd, _, _, _ = O.fp64_direct(s.particle[:, :3], s.type, s.eps, s.sigma, s.box, s.rc,
                                charge=s.charge if s.coulomb_k else None, ke=s.coulomb_k)
    # further GPU force sets at other positions (the precision leg's relaxed system):
    # relative L2 error of each against the FP64 direct evaluation there
    extra = {}
    if checks:
        pos, sets = checks
        fp, _, _, _ = O.fp64_direct(pos[:, :3], s.type, s.eps, s.sigma, s.box, s.rc,
                                    charge=s.charge if s.coulomb_k else None, ke=s.coulomb_k)
        for name, f in sets.items():
            extra[name] = float(np.linalg.norm(f.astype(np.float64) / 2 ** 32 * 0.1 - fp) / np.linalg.norm(fp))
    return base, force_error(gpu_fix, fd), fd, extra


def cmodel_agreement(O, iterations=10, seed=2019):
    """The reference C model's own computation (all pairs, pow() arithmetic, kick-drift) on
    the GPU entry point against its CPU restatement: 500 argon atoms placed uniformly in the
    C model's 24 A box from the seed, at rest, with the C model's defaults.  The C model
    clears its force columns after each update, so the forces are compared through the
    velocities after one iteration (v = F dt / m from rest); the whole state (positions,
    velocities, per-atom energies) and the printed energy sums after `iterations`."""
    rng = np.random.default_rng(seed)
    p = np.zeros((500, 11), np.float32)
    p[:, :3] = rng.uniform(0.0, 24.0, (500, 3)).astype(np.float32)
    out = {"input": "500 atoms uniform in the 24 A box, at rest (seed %d), C-model defaults" % seed,
           "tolerance": 1e-5}
    params = O.cmodel_default_params()
    params.iterations = 1
    want, _, _ = O.cmodel_run(p, params)
    got, _ = mdrl.cmodel_run(p, iterations=1)
    vw, vg = want[:, 3:6].astype(np.float64), got[:, 3:6].astype(np.float64)
    rel = float(np.linalg.norm(vg - vw) / max(np.linalg.norm(vw), 1e-300))
    out["force_rel_l2_1_iteration"] = rel
    out["within_tolerance"] = rel <= 1e-5
    params.iterations = iterations
    want, we, _ = O.cmodel_run(p, params)
    got, ge = mdrl.cmodel_run(p, iterations=iterations)
    out["state_bit_identical_%d_iterations" % iterations] = bool(
        np.array_equal(got.view(np.uint32), want.view(np.uint32)) and np.array_equal(ge.view(np.uint32), we.view(np.uint32)))
    return out


def force_error(gpu_fix, fd):
    """int64 GPU forces (2^-32 kJ/mol/nm) against FP64 forces (kJ/mol/A)."""
    f = gpu_fix.astype(np.float64) / 2 ** 32 * 0.1
    d = np.linalg.norm(f - fd, axis=1)
    m = np.linalg.norm(fd, axis=1)
    per = d / np.maximum(m, 1e-300)
    return {
        "vs": "FP64 analytic direct evaluation (all images within rc)",
        "global_rel_l2": float(np.linalg.norm(f - fd) / np.linalg.norm(fd)),
        "per_atom_median": float(np.median(per)),
        "per_atom_p99": float(np.percentile(per, 99)),
        "target_1e-5_met": bool(np.linalg.norm(f - fd) / np.linalg.norm(fd) <= 1e-5),
    }


def analytic_leg(s, relaxed, fd, steps, cpu_value):
    """The same workload with the FP32 direct pair function (MDRL_EVAL_ANALYTIC) in place of
    the reference's tables: throughput over `steps` graph-captured steps from the relaxed
    state, force error at the initial positions against the same FP64 reference."""
    g = mdrl.System.from_synthetic(s, flags=mdrl.EVAL_ANALYTIC)
    fix_a = g.forces_fixed()
    g.set_particles(relaxed)
    sched, _ = g.autotune(reps=2)
    g.step(20)
    g.set_graph(True)
    t0 = time.perf_counter()
    g.step(steps)
    wall = time.perf_counter() - t0
    pairs = g.stats().pair_evals
    g.set_graph(False)
    g.set_timing(True)
    g.step(steps)
    st = g.stats()
    g.close()
    ms = wall / steps * 1e3
    value = pairs / (ms * 1e-3)
    return {"value": value, "unit": "pair-evals/s", "ms_per_step": ms, "steps": steps,
            "force_kernel_ms": st.force_ms, "schedule": sched.as_dict(),
            "over_cpu_baseline": value / cpu_value, "force_error": force_error(fix_a, fd),
            "note": "FP32 direct LJ + cutoff Coulomb instead of the first-order tables"}


def cpu_celllist_leg(s, pairs_per_step, budget_s):
    """The half-shell cell-list restatement (int64 fixed point, the same pair set and table
    arithmetic as the GPU) with OpenMP on every host core: separates the algorithmic
    speedup over the all-pairs C model from the hardware speedup."""
    import math
    from oracle import oracle as O

    t = mdrl.Table.load()
    tab6 = O.interleave([*t.get(mdrl.TERM_14), *t.get(mdrl.TERM_8), *t.get(mdrl.TERM_3)])
    ncell = [max(1, int(math.floor(float(s.box[d]) / s.rc))) for d in range(3)]
    rl = O.RL(s.particle[:, :3], s.box, ncell, s.rc, s.type, s.eps, s.sigma, tab6,
              charge=s.charge if s.coulomb_k else None, coulomb_k=s.coulomb_k)
    reps, tot = 0, 0.0
    while reps < 1 or (tot < budget_s and reps < 20):
        t0 = time.perf_counter()
        _, npairs = rl.forces_fix()
        tot += time.perf_counter() - t0
        reps += 1
    step_s = tot / reps
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": npairs / step_s,
        "unit": "pair-evals/s",
        "cores": cores,
        "kind": "port",
        "sample": "half-shell cell list, int64 fixed point, OpenMP over home cells on %d threads "
                  "(%s); %d force evaluation(s) of the full %d-atom system, %.3f s each"
                  % (cores, cpu_model(), reps, s.n, step_s),
        "ms_per_step": step_s * 1e3,
    }


def batched_vs_plain(nrep, steps, seed, table):
    """The batched workload: nrep replicas (log-uniform 256..16,384 atoms, 0.1 atoms/A^3,
    rc 10 A) stepped as one batch (autotuned schedule, hipGraph-captured steps) against the
    same replicas stepped one context at a time.  Both produce bit-identical int64 forces,
    so the error is equal by construction."""
    systems = [mdrl.SyntheticSystem(4, k, seed=seed) for k in range(nrep)]
    natoms = sum(s.n for s in systems)
    b = mdrl.Batch.from_synthetic(systems, table=table)
    sched, tune_ms = b.autotune(reps=2)
    b.set_graph(True)
    b.step(2)
    pairs = b.stats().pair_evals
    t0 = time.perf_counter()
    b.step_async(steps)
    b.sync()
    tb = (time.perf_counter() - t0) / steps
    pairs_b = b.stats().pair_evals
    b.close()
    plain = [mdrl.System.from_synthetic(s, table=table) for s in systems]
    for g in plain:
        g.step(2)
    t0 = time.perf_counter()
    for _ in range(steps):
        for g in plain:
            g.step_async(1)
    for g in plain:
        g.sync()
    tp = (time.perf_counter() - t0) / steps
    pairs_p = sum(g.stats().pair_evals for g in plain)
    for g in plain:
        g.close()
    return {
        "workload": "%d replicas, %d atoms total, log-uniform 256..16,384 atoms each, "
                    "0.1 atoms/A^3, rc 10 A, dt 1 fs" % (nrep, natoms),
        "steps": steps,
        "batched": {"ms_per_step": tb * 1e3, "pair_evals_per_s": pairs_b / tb,
                    "schedule": sched.as_dict(),
                    "graph": True},
        "plain": {"ms_per_step": tp * 1e3, "pair_evals_per_s": pairs_p / tp,
                  "note": "one context per replica, one step launched per replica in turn"},
        "pair_evals_per_step": pairs_b,
        "batched_over_plain": (pairs_b / tb) / (pairs_p / tp),
        "force_error": "equal (bit-identical int64 forces)",
    }


def precision_leg(s, relax, relax_step, nve_steps, dt_list):
    """The precision configuration: int64 forces bit-identical over 10 repeat
    evaluations, FP32 atomics vs int64, FP64-reduced total energy along NVE runs (one per
    time step in dt_list).  Returns the leg's record and, for the CPU leg's comparison with
    the FP64 direct oracle, the relaxed positions and both force sets."""
    import hashlib

    g = mdrl.System.from_synthetic(s)
    if relax:
        g.minimize(relax, relax_step, 1e-4)
    digests = set()
    for _ in range(10):
        g.force_eval()
        digests.add(hashlib.sha256(g.forces_fixed().tobytes()).hexdigest())
    fix = g.forces_fixed()
    relaxed = g.particles()
    g.close()
    h = mdrl.System.from_synthetic(s, flags=mdrl.ACC_FP32)
    h.set_particles(relaxed)
    f32a = h.forces_fixed()
    h.force_eval()
    f32b = h.forces_fixed()
    h.close()
    out = {
        "repeat_force_evals": 10,
        "int64_bit_identical": len(digests) == 1,
        "fp32_atomics_runs_bit_identical": bool(np.array_equal(f32a, f32b)),
        "fp32_vs_int64_rel_l2": float(np.linalg.norm((f32a - fix).astype(np.float64)) /
                                      np.linalg.norm(fix.astype(np.float64))),
        "nve": [],
    }
    for dt in dt_list:
        sd = copy.copy(s)
        sd.dt = float(dt)
        g = mdrl.System.from_synthetic(sd)
        g.set_particles(relaxed)
        g.set_graph(True)
        pe0, ke0, _ = g.energy()
        trace = [(0, pe0 + ke0)]
        every = max(1, nve_steps // 10)
        t0 = time.perf_counter()
        done = 0
        while done < nve_steps:
            k = min(every, nve_steps - done)
            g.step_async(k)
            done += k
            pe, ke, _ = g.energy()
            trace.append((done, pe + ke))
        wall = time.perf_counter() - t0
        g.close()
        e0 = trace[0][1]
        out["nve"].append({"dt_fs": dt, "steps": nve_steps, "E0_kj_mol": e0,
                           "E_end_kj_mol": trace[-1][1],
                           "rel_drift": (trace[-1][1] - e0) / abs(e0),
                           "max_rel_dev": max(abs(e - e0) for _, e in trace) / abs(e0),
                           "trace": trace, "wall_s": wall})
    return out, relaxed, {"int64": fix, "fp32": f32a}


def bench_batch(args, which):
    """A batched workload as one graph-captured batch: the 512 uneven LJ + Coulomb replicas
    (which = 4), or the bonded workload, 128 replicas of 64 chains x 48 beads with bonded
    terms and exclusions (which = 5); then the same replicas as separate contexts."""
    nrep = args.replicas if args.replicas else (128 if which == 5 else 512)
    systems = [mdrl.SyntheticSystem(which, k, seed=args.seed) for k in range(nrep)]
    if args.dt is not None:
        for s in systems:
            s.dt = float(args.dt)
    s0 = systems[0]
    b = mdrl.Batch.from_synthetic(systems)
    relax = args.relax if args.relax else 0
    if relax:
        b.minimize(relax, args.relax_step, 1e-4)
    sched, tune_ms = b.autotune(reps=3)
    b.step(args.warmup)
    pe0, ke0, _ = b.energy()
    b.set_graph(bool(args.graph))
    t0 = time.perf_counter()
    b.step_async(args.steps)
    b.sync()
    wall = time.perf_counter() - t0
    pairs = int(b.stats().pair_evals)
    pe1, ke1, _ = b.energy()
    b.set_graph(False)
    b.set_timing(True)
    b.step_async(min(args.steps, 1000))
    b.sync()
    st = b.stats()
    b.close()
    ms = wall / args.steps * 1e3
    value = pairs / (ms * 1e-3)
    # the same replicas as separate contexts, stepped one after another
    nplain = min(args.replica_steps, args.steps)
    plain = [mdrl.System.from_synthetic(s) for s in systems] if nplain > 0 else []
    for g in plain:
        if relax:
            g.minimize(relax, args.relax_step, 1e-4)
        g.step(2)
    t0 = time.perf_counter()
    for _ in range(nplain):
        for g in plain:
            g.step_async(1)
    for g in plain:
        g.sync()
    tp = (time.perf_counter() - t0) / max(nplain, 1)
    pairs_p = sum(int(g.stats().pair_evals) for g in plain)
    for g in plain:
        g.close()
    natoms = sum(s.n for s in systems)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "pair-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded %s, seed %d)" % ("chains" if which == 5 else "lattice replicas", args.seed),
        "config": {
            "workload": ("%d replicas x 64 chains x 48 beads (%d beads), 0.033 beads/A^3, rc 10 A, "
                         "bonds/angles/torsions, LJ + Coulomb excluded within 3 bonds, dt %g fs"
                         % (nrep, natoms, s0.dt)) if which == 5 else
                        ("%d replicas (%d atoms), log-uniform 256..16,384 atoms each, 0.1 atoms/A^3, "
                         "rc 10 A, LJ + Coulomb, dt %g fs" % (nrep, natoms, s0.dt)),
            "n_atoms": natoms,
            "bonded_terms": int(sum(s.nterms for s in systems)),
            "pair_evals_per_step": pairs,
            "accumulation": "int64 fixed point (2^-32 kJ/mol/nm)",
            "pair_eval": "first-order tables, ApoA1 layout",
            "integrator": "velocity Verlet",
            "parallelism": "single GPU, one batch",
        },
        "ns_per_day": (1e3 / ms) * s0.dt * 0.0864,
        "force_kernel_ms": st.force_ms,
        "step_device_ms_eager": st.step_ms,
        "schedule": dict(sched.as_dict(), autotune_best_ms=tune_ms),
        "graph_captured_steps": bool(args.graph),
        "energy_kj_mol": {"potential_start": pe0, "kinetic_start": ke0, "potential_end": pe1,
                          "kinetic_end": ke1,
                          "total_change_rel_ke": (pe1 + ke1 - pe0 - ke0) / ke0},
        "pre_relaxation": {"iterations": relax, "max_disp_A": args.relax_step},
    }
    if nplain > 0:
        out["batched_vs_plain"] = {"batched_pair_evals_per_s": value, "plain_ms_per_step": tp * 1e3,
                                   "plain_pair_evals_per_s": pairs_p / tp,
                                   "batched_over_plain": value / (pairs_p / tp), "plain_steps": nplain}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", type=int, default=2, choices=sorted(WORKLOADS))
    ap.add_argument("--seed", type=int, default=2019)
    ap.add_argument("--fp32-atomics", action="store_true", help="FP32 atomics instead of int64")
    ap.add_argument("--analytic", action="store_true", help="FP32 direct instead of tables")
    ap.add_argument("--relax", type=int, default=200,
                    help="capped steepest-descent iterations before timing (lattice workloads)")
    ap.add_argument("--relax-step", type=float, default=0.02, help="max displacement per iteration (A)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--replicas", type=int, default=None,
                    help="replicas of the batched legs (default 512 for workloads 2 and 4, 128 for 5; 0: skip)")
    ap.add_argument("--replica-steps", type=int, default=20)
    ap.add_argument("--graph", type=int, default=1, help="capture the timed steps in hipGraphs")
    ap.add_argument("--dt", type=float, default=None, help="time step in fs (default: the workload's)")
    ap.add_argument("--nve-steps", type=int, default=10000,
                    help="NVE steps of the precision leg (workload 3)")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="also run one device copy of this many bytes (traffic counter calibration)")
    ap.add_argument("--long-range-field", action="store_true",
                    help="attach a summation-hook provider (a uniform field on the charges, copied "
                         "from a device buffer each evaluation) for the steps, to time the hook")
    args = ap.parse_args()
    if args.gpus != 1:
        sys.stderr.write("bench.py: this build runs on one GPU; --gpus %d is not supported "
                         "(replicas and spatial domains would shard one process per GPU)\n" % args.gpus)
        sys.exit(2)

    if args.replicas is None:
        args.replicas = 128 if args.workload == 5 else 512
    if args.workload in (4, 5):
        bench_batch(args, args.workload)
        return
    s = mdrl.SyntheticSystem(args.workload, seed=args.seed)
    if args.dt is not None:
        s.dt = float(args.dt)
    flags = (mdrl.ACC_FP32 if args.fp32_atomics else 0) | (mdrl.EVAL_ANALYTIC if args.analytic else 0)
    g = mdrl.System.from_synthetic(s, flags=flags)
    fix0 = g.forces_fixed()
    coul = bool(s.coulomb_k)
    pe_start, _, _ = g.energy()
    relax = args.relax if args.workload != 1 else 0
    if relax:
        g.minimize(relax, args.relax_step, 1e-4)
    pe_relaxed, ke0, _ = g.energy()
    relaxed = g.particles()

    sched, tune_ms = g.autotune(reps=3)
    field_buf = None
    if args.long_range_field:
        # F_i = q_i E with E = (0.8, -0.3, 0.45) kJ/mol/A/e, input order
        field = (s.charge.astype(np.float32)[:, None] *
                 np.array([0.8, -0.3, 0.45], np.float32)).astype(np.float32)
        field_buf = mdrl.DeviceBuffer(field)
        g.set_long_range(mdrl.device_provider(field_buf))
    g.step(args.warmup)
    # timed region: the steps as a user runs them (graph-captured unless --graph 0)
    g.set_graph(bool(args.graph))
    t0 = time.perf_counter()
    g.step_async(args.steps)
    g.sync()
    wall = time.perf_counter() - t0
    pairs_last = g.stats().pair_evals
    # kernel time: the same number of steps again, each force launch bracketed by HIP
    # events recorded on the stream it runs on
    g.set_graph(False)
    g.set_timing(True)
    g.step_async(args.steps)
    g.sync()
    st = g.stats()
    g.set_timing(False)
    pe, ke, _ = g.energy()
    g.close()
    if args.calibrate_copy:
        # a plain streaming kernel of known traffic (reads and writes calibrate_copy bytes)
        # for scaling the HBM byte counters
        import torch
        src = torch.empty(args.calibrate_copy, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        mdrl.stream_copy(dst.data_ptr(), src.data_ptr(), args.calibrate_copy)

    ms_per_step = wall / args.steps * 1e3
    pairs = int(pairs_last)
    value = pairs / (ms_per_step * 1e-3)
    ns_day = (1e3 / ms_per_step) * s.dt * 0.0864
    flops = FLOPS_PER_PAIR[coul] * pairs
    achieved = flops / (st.force_ms * 1e-3) / 1e12
    # HBM bytes per force launch from the committed counter passes (FETCH_SIZE and
    # WRITE_SIZE, scaled by a copy kernel of known traffic); null when not measured
    traffic = None
    tf = os.path.join(ROOT, "profiles", "force_traffic_workload%d.json" % args.workload)
    if os.path.exists(tf) and not (args.fp32_atomics or args.analytic):
        traffic = json.load(open(tf)).get("hbm_bytes_per_launch")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "pair-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded lattice, seed %d)" % args.seed,
        "config": {
            "workload": WORKLOADS[args.workload],
            "n_atoms": s.n,
            "pair_evals_per_step": pairs,
            "accumulation": "fp32 atomics" if args.fp32_atomics else "int64 fixed point (2^-32 kJ/mol/nm)",
            "pair_eval": "FP32 analytic" if args.analytic else "first-order tables, ApoA1 layout",
            "integrator": "velocity Verlet",
            "parallelism": "single GPU",
        },
        "ns_per_day": ns_day,
        "force_kernel_ms": st.force_ms,
        "schedule": dict(sched.as_dict(), autotune_best_ms=tune_ms),
        "dt_fs": s.dt,
        "step_device_ms_eager": st.step_ms,
        "graph_captured_steps": bool(args.graph),
        "long_range_hook": ("uniform field on the charges, device copy per evaluation"
                            if args.long_range_field else None),
        "energy_kj_mol": {"potential_initial": pe_start, "potential_relaxed": pe_relaxed,
                          "kinetic_relaxed": ke0, "potential_end": pe, "kinetic_end": ke},
        "pre_relaxation": {"iterations": relax, "max_disp_A": args.relax_step,
                           "note": "capped steepest descent before timing; the jittered lattice "
                                   "starts with r_min ~1.35 A"},
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": traffic,
            "kernel": "k_rl_halfshell",
            "flops_per_pair": FLOPS_PER_PAIR[coul],
            "hbm": None if traffic is None else {
                "achieved": traffic / (st.force_ms * 1e-3) / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": traffic / (st.force_ms * 1e-3) / 1e9 / PEAK_HBM_GBS},
        },
    }
    if args.replicas and args.workload == 2:
        out["batched_vs_plain"] = batched_vs_plain(args.replicas, args.replica_steps, args.seed,
                                                   mdrl.Table.load())
    checks = None
    if args.workload == 3:
        out["precision"], relaxed_p, sets = precision_leg(s, relax, args.relax_step, args.nve_steps,
                                                          sorted({s.dt, 0.5}))
        checks = (relaxed_p, sets)
    if not args.no_cpu:
        base, err, fd, extra = cpu_baseline_leg(s, fix0, pairs, args.cpu_seconds, checks)
        if "int64" in extra:
            out["precision"]["int64_vs_fp64_rel_l2"] = extra["int64"]
            out["precision"]["fp32_vs_fp64_rel_l2"] = extra["fp32"]
        out["cpu_baseline"] = base
        out["cpu_celllist_baseline"] = cpu_celllist_leg(s, pairs, args.cpu_seconds / 3)
        out["force_error"] = err
        out["speedup_vs_cpu_baseline"] = value / base["value"]
        if args.workload == 2 and not (args.analytic or args.fp32_atomics):
            out["analytic_mode"] = analytic_leg(s, relaxed, fd, min(args.steps, 200), base["value"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
