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


def cpu_baseline_leg(s, pairs_per_step, budget_s):
    """The C model restatement (oracle), one thread, on a bounded sample of the workload.
    The oracle supplies the CPU baseline and the reference results here, never the measured
    GPU path."""
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
    return {
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


def fp64_reference(s, xyz):
    """FP64 analytic direct forces (kJ/mol/A) of the workload at positions xyz."""
    from oracle import oracle as O
    f, _, _, _ = O.fp64_direct(xyz[:, :3], s.type, s.eps, s.sigma, s.box, s.rc,
                               charge=s.charge if s.coulomb_k else None, ke=s.coulomb_k)
    return f


def table_model_fp64(s, xyz):
    """The same table model evaluated in FP64 (FP32 coefficients, the same pair set): the
    reference for the kernel's own arithmetic error."""
    import math
    from oracle import oracle as O
    t = mdrl.Table.load()
    tab6 = O.interleave([*t.get(mdrl.TERM_14), *t.get(mdrl.TERM_8), *t.get(mdrl.TERM_3)])
    ncell = [max(1, int(math.floor(float(s.box[d]) / s.rc))) for d in range(3)]
    rl = O.RL(xyz[:, :3], s.box, ncell, s.rc, s.type, s.eps, s.sigma, tab6,
              charge=s.charge if s.coulomb_k else None, coulomb_k=s.coulomb_k)
    f, _ = rl.forces_table_fp64()
    return f


C_MODEL_TOL = {"analytic": 1e-5, "table": 1.5e-4}


def cmodel_agreement(O):
    """The half-shell path on the GPU against the reference C model's force arithmetic
    (R/C_Model_ForceEval/MD_Evaluation_Model.cpp:72-124: all pairs, pow() in double, FP32
    sums) per atom, on the 512-atom argon workload (34 A box, rc 8.5 A, sigma 3.4 A,
    eps 0.996 kJ/mol), in both pair modes; and the GPU port of the C model itself against
    its CPU restatement (500 atoms at rest in the C model's 24 A box): state bit-identical
    after 10 iterations."""
    s = mdrl.SyntheticSystem(1)
    p = O.cmodel_default_params()
    p.eps, p.sigma, p.rc = float(s.eps[0]), float(s.sigma[0]), float(s.rc)
    p.box[:] = [float(b) for b in s.box]
    part = np.zeros((s.n, 11), np.float32)
    part[:, :3] = s.particle[:, :3]
    _, fc = O.cmodel_force_sample(part, p, s.n)
    fc = fc.astype(np.float64)
    out = {"input": "512-atom argon workload (seed 2019), forces at its initial positions",
           "reference": "C-model force loop, MD_Evaluation_Model.cpp:72-124, with these constants"}
    for mode, flags in (("analytic", mdrl.EVAL_ANALYTIC), ("table", 0)):
        g = mdrl.System.from_synthetic(s, flags=flags)
        e = force_error(g.forces_fixed(), fc, "C model (FP32, pow() in double), per atom")
        g.close()
        del e["target_1e-5_met"]
        e["tolerance_global"] = C_MODEL_TOL[mode]
        e["within_tolerance"] = e["global_rel_l2"] <= C_MODEL_TOL[mode]
        out[mode + "_mode"] = e
    rng = np.random.default_rng(2019)
    q = np.zeros((500, 11), np.float32)
    q[:, :3] = rng.uniform(0.0, 24.0, (500, 3)).astype(np.float32)
    params = O.cmodel_default_params()
    params.iterations = 10
    want, we, _ = O.cmodel_run(q, params)
    got, ge = mdrl.cmodel_run(q, iterations=10)
    out["gpu_port_state_bit_identical_10_iterations"] = bool(
        np.array_equal(got.view(np.uint32), want.view(np.uint32)) and
        np.array_equal(ge.view(np.uint32), we.view(np.uint32)))
    return out


def force_error(gpu_fix, fd, vs="FP64 analytic direct evaluation (all images within rc)"):
    """int64 GPU forces (2^-32 kJ/mol/nm) against reference forces fd (kJ/mol/A)."""
    f = gpu_fix.astype(np.float64) / 2 ** 32 * 0.1
    d = np.linalg.norm(f - fd, axis=1)
    m = np.linalg.norm(fd, axis=1)
    per = d / np.maximum(m, 1e-300)
    g = float(np.linalg.norm(f - fd) / np.linalg.norm(fd))
    return {
        "vs": vs,
        "global_rel_l2": g,
        "per_atom_median": float(np.median(per)),
        "per_atom_p99": float(np.percentile(per, 99)),
        "per_atom_max": float(per.max()),
        "target_1e-5_met": bool(g <= 1e-5),
    }


def run_timed(g, warmup, steps, graph):
    """warmup untimed steps, the graphs of the timed steps captured, then `steps` timed
    steps (graph replays when graph is set); returns the wall time per step (s)."""
    g.set_graph(bool(graph))
    g.step(warmup)
    if graph and warmup > 0:
        g.prepare_graphs(steps)
    t0 = time.perf_counter()
    g.step_async(steps)
    g.sync()
    return (time.perf_counter() - t0) / steps


def analytic_leg(s, relaxed, fd_relaxed, warmup, steps, cpu_value):
    """The same workload with the FP32 direct pair function (MDRL_EVAL_ANALYTIC, cutoff
    decided exactly) in place of the reference's tables: throughput over `steps`
    graph-captured steps from the relaxed state; force error against the FP64 direct
    evaluation at the relaxed state (where the timed steps start) and after them."""
    g = mdrl.System.from_synthetic(s, flags=mdrl.EVAL_ANALYTIC)
    g.set_particles(relaxed)
    err_relaxed = force_error(g.forces_fixed(), fd_relaxed)
    sched, _, _ = g.autotune(reps=2, cache=True)
    step_s = run_timed(g, warmup, steps, True)
    pairs = g.stats().pair_evals
    fix_end, pos_end = g.forces_fixed(), g.particles()
    g.set_graph(False)
    g.set_timing(True)
    g.step(min(steps, 200))
    st = g.stats()
    g.close()
    ms = step_s * 1e3
    value = pairs / step_s
    return {"value": value, "unit": "pair-evals/s", "ms_per_step": ms, "steps": steps,
            "force_kernel_ms": st.force_ms, "schedule": sched.as_dict(),
            "over_cpu_baseline": value / cpu_value,
            "force_error": err_relaxed,
            "force_error_after_timed_steps": force_error(fix_end, fp64_reference(s, pos_end)),
            "note": "FP32 direct LJ + cutoff Coulomb instead of the first-order tables; the "
                    "cutoff test r^2 <= rc^2 is decided in FP64 for pairs within the FP32 "
                    "error margin of rc^2, so the pair set is the FP64 evaluation's"}


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


def state_digest(ctxs_and_sys):
    """SHA-256 over the int64 forces and position bits of systems, in order; items are
    (batch, k) pairs or single contexts."""
    import hashlib
    h = hashlib.sha256()
    for item in ctxs_and_sys:
        if isinstance(item, tuple):
            b, k = item
            f, p = b.system_forces_fixed(k), b.system_particles(k)
        else:
            f, p = item.forces_fixed(), item.particles()
        h.update(f.tobytes())
        h.update(p[:, :3].tobytes())
    return h.hexdigest()


def plain_contexts(systems, table, relax=0, relax_step=0.02, tune=False, flags=0):
    plain = [mdrl.System.from_synthetic(s, flags=flags, table=table) for s in systems]
    for g in plain:
        if relax:
            g.minimize(relax, relax_step, 1e-4)
        if tune:
            g.autotune(reps=1, cache=True)
    return plain


def time_plain(plain, warmup, steps, graph):
    """Every context takes `warmup` steps, then (graph: each replays the graphs of its
    `steps` steps, captured beforehand; each context on its own stream, all enqueued before
    one wait) or (eager: one step per context in turn, `steps` rounds) is timed."""
    for g in plain:
        g.set_graph(graph)
        g.step(warmup)
        if graph and warmup > 0:
            g.prepare_graphs(steps)
    t0 = time.perf_counter()
    if graph:
        for g in plain:
            g.step_async(steps)
    else:
        for _ in range(steps):
            for g in plain:
                g.step_async(1)
    for g in plain:
        g.sync()
    return (time.perf_counter() - t0) / steps


def batched_vs_plain(nrep, warmup, steps, seed, table):
    """The batched workload: nrep replicas (log-uniform 256..16,384 atoms, 0.1 atoms/A^3,
    rc 10 A) stepped as one batch (autotuned schedule, graph-captured steps) against the same
    replicas as one context each, (a) autotuned and graph-captured, every context's steps
    enqueued on its own stream, and (b) eager with the default schedule, one step per
    context in turn.  All three start from the same state and take the same steps; their
    int64 forces and positions are hashed to show the error is equal (identical bits)."""
    systems = [mdrl.SyntheticSystem(4, k, seed=seed) for k in range(nrep)]
    natoms = sum(s.n for s in systems)
    b = mdrl.Batch.from_synthetic(systems, table=table)
    sched, tune_ms, tune_hit = b.autotune(reps=2, cache=True)
    tb = run_timed(b, warmup, steps, True)
    pairs_b = b.stats().pair_evals
    dig_b = state_digest([(b, k) for k in range(nrep)])
    b.close()
    out = {
        "workload": "%d replicas, %d atoms total, log-uniform 256..16,384 atoms each, "
                    "0.1 atoms/A^3, rc 10 A, dt 1 fs" % (nrep, natoms),
        "steps": steps,
        "warmup": warmup,
        "batched": {"ms_per_step": tb * 1e3, "pair_evals_per_s": pairs_b / tb,
                    "schedule": dict(sched.as_dict(), autotune_cache_hit=tune_hit), "graph": True},
        "pair_evals_per_step": pairs_b,
    }
    digests = {"batched": dig_b}
    for name, tune, graph, note in (
            ("plain", True, True, "one context per replica, each autotuned, its steps replayed "
                                  "from captured graphs on its own stream"),
            ("plain_eager", False, False, "one context per replica, default schedule, eager, "
                                          "one step per replica in turn")):
        plain = plain_contexts(systems, table, tune=tune)
        tp = time_plain(plain, warmup, steps, graph)
        pairs_p = sum(g.stats().pair_evals for g in plain)
        digests[name] = state_digest(plain)
        for g in plain:
            g.close()
        out[name] = {"ms_per_step": tp * 1e3, "pair_evals_per_s": pairs_p / tp, "note": note}
    out["batched_over_plain"] = out["batched"]["pair_evals_per_s"] / out["plain"]["pair_evals_per_s"]
    out["batched_over_plain_eager"] = (out["batched"]["pair_evals_per_s"] /
                                       out["plain_eager"]["pair_evals_per_s"])
    out["state_sha256"] = digests
    out["force_error"] = ("equal: identical int64 forces and positions in all three runs"
                          if len(set(digests.values())) == 1 else "DIFFERENT")
    return out


def precision_leg(s, relax, relax_step, nve_steps, dt_list):
    """The precision configuration: int64 forces bit-identical over 10 repeat
    evaluations, FP32 atomics vs int64, FP64-reduced total and shifted energies along NVE
    runs (one per time step in dt_list), sampled ten times per run.  Returns the leg's record and, for the CPU leg's comparison with
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

        def sample(step):
            # raw KE + PE, and the shifted energy KE + PE - cutoff_shift (one U(rc) per
            # in-cutoff pair): the quantity velocity Verlet conserves on a potential that is
            # truncated without a shift
            t = g.energy_terms()
            return (step, t["kinetic"] + t["potential"], t["kinetic"] + t["potential"] - t["cutoff_shift"],
                    t["kinetic"])

        trace = [sample(0)]
        every = max(1, nve_steps // 10)
        t0 = time.perf_counter()
        done = 0
        while done < nve_steps:
            k = min(every, nve_steps - done)
            g.step_async(k)
            done += k
            trace.append(sample(done))
        wall = time.perf_counter() - t0
        g.close()
        _, e0, h0, k0 = trace[0]
        out["nve"].append({"dt_fs": dt, "steps": nve_steps, "KE0_kj_mol": k0,
                           "E0_kj_mol": e0, "E_end_kj_mol": trace[-1][1],
                           "shifted_E0_kj_mol": h0, "shifted_E_end_kj_mol": trace[-1][2],
                           "rel_drift": (trace[-1][1] - e0) / abs(e0),
                           "max_rel_dev": max(abs(e - e0) for _, e, _, _ in trace) / abs(e0),
                           "shifted_drift_vs_KE0": (trace[-1][2] - h0) / k0,
                           "shifted_max_dev_vs_KE0": max(abs(h - h0) for _, _, h, _ in trace) / k0,
                           "trace_step_raw_shifted_kinetic": trace, "wall_s": wall})
    return out, relaxed, {"int64": fix, "fp32": f32a}


def tune_or_pin(g, pinned):
    """The autotuner's schedule (through its cache), or the one given as "apw,tile,groups,
    persistent": (schedule, ms of the winner or None, cache hit)."""
    if not pinned:
        return g.autotune(reps=3, cache=True)
    g.set_schedule(*[int(v) for v in pinned.split(",")])
    return g.schedule(), None, False


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
    coarse = args.coarse_cells if args.coarse_cells >= 0 else int(which == 5)
    cflags = mdrl.COARSE_CELLS if coarse else 0
    b = mdrl.Batch.from_synthetic(systems, flags=cflags)
    relax = args.relax if args.relax else 0
    if relax:
        b.minimize(relax, args.relax_step, 1e-4)
    sched, tune_ms, tune_hit = tune_or_pin(b, args.schedule)
    b.step(args.warmup)
    pe0, ke0, _ = b.energy()
    b.set_graph(bool(args.graph))
    if args.graph:
        b.prepare_graphs(args.steps)
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
    # the same replicas as separate contexts (autotuned and graph-captured, then eager with
    # the default schedule), against a fresh batch taking the same steps: timing and hashes
    nplain = min(args.replica_steps, args.steps)
    versus = None
    if nplain > 0:
        b2 = mdrl.Batch.from_synthetic(systems, flags=cflags)
        if relax:
            b2.minimize(relax, args.relax_step, 1e-4)
        b2.set_schedule(sched.atoms_per_wave, sched.tile_depth, sched.offset_groups, sched.persistent)
        tb2 = run_timed(b2, args.warmup, nplain, True)
        pairs_b2 = b2.stats().pair_evals
        digests = {"batched": state_digest([(b2, k) for k in range(nrep)])}
        b2.close()
        versus = {"steps": nplain, "warmup": args.warmup,
                  "batched": {"ms_per_step": tb2 * 1e3, "pair_evals_per_s": pairs_b2 / tb2}}
        for name, tune, graph in (("plain", True, True), ("plain_eager", False, False)):
            plain = plain_contexts(systems, None, relax, args.relax_step, tune, cflags)
            tp = time_plain(plain, args.warmup, nplain, graph)
            pairs_p = sum(int(g.stats().pair_evals) for g in plain)
            digests[name] = state_digest(plain)
            for g in plain:
                g.close()
            versus[name] = {"ms_per_step": tp * 1e3, "pair_evals_per_s": pairs_p / tp,
                            "autotuned": tune, "graph": graph}
        versus["batched_over_plain"] = versus["batched"]["pair_evals_per_s"] / versus["plain"]["pair_evals_per_s"]
        versus["batched_over_plain_eager"] = (versus["batched"]["pair_evals_per_s"] /
                                              versus["plain_eager"]["pair_evals_per_s"])
        versus["state_sha256"] = digests
        versus["force_error"] = ("equal: identical int64 forces and positions in all three runs"
                                 if len(set(digests.values())) == 1 else "DIFFERENT")
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
            "cells": ("coarse: floor(L/rc) per axis, reduced while cells average < 128 atoms "
                      "(MDRL_COARSE_CELLS)" if coarse else "floor(L/rc) per axis"),
        },
        "ns_per_day": (1e3 / ms) * s0.dt * 0.0864,
        "force_kernel_ms": st.force_ms,
        "step_device_ms_eager": st.step_ms,
        "schedule": dict(sched.as_dict(), autotune_best_ms=tune_ms, autotune_cache_hit=tune_hit),
        "graph_captured_steps": bool(args.graph),
        "energy_kj_mol": {"potential_start": pe0, "kinetic_start": ke0, "potential_end": pe1,
                          "kinetic_end": ke1,
                          "total_change_rel_ke": (pe1 + ke1 - pe0 - ke0) / ke0},
        "pre_relaxation": {"iterations": relax, "max_disp_A": args.relax_step},
    }
    if versus:
        out["batched_vs_plain"] = versus
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
    ap.add_argument("--schedule", default=None,
                    help="atoms_per_wave,tile_depth,offset_groups,persistent instead of the autotuner's choice")
    ap.add_argument("--coarse-cells", type=int, default=-1,
                    help="batched workloads: 1 = MDRL_COARSE_CELLS grid, 0 = floor(L/rc) grid, "
                         "-1 = coarse for the bonded workload (sparse: 48 beads per rc-sized cell)")
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
    fix_unrelaxed = g.forces_fixed()
    coul = bool(s.coulomb_k)
    pe_start, _, _ = g.energy()
    relax = args.relax if args.workload != 1 else 0
    if relax:
        g.minimize(relax, args.relax_step, 1e-4)
    pe_relaxed, ke0, _ = g.energy()
    relaxed = g.particles()
    fix_relaxed = g.forces_fixed()     # the forces the timed steps start from

    sched, tune_ms, tune_hit = tune_or_pin(g, args.schedule)
    field_buf = None
    if args.long_range_field:
        # F_i = q_i E with E = (0.8, -0.3, 0.45) kJ/mol/A/e, input order
        field = (s.charge.astype(np.float32)[:, None] *
                 np.array([0.8, -0.3, 0.45], np.float32)).astype(np.float32)
        field_buf = mdrl.DeviceBuffer(field)
        g.set_long_range(mdrl.device_provider(field_buf))
    # warm-up with the timed configuration, every graph the timed steps replay captured
    # beforehand; timed region: the steps as a user runs them (graph replays unless --graph 0)
    step_s = run_timed(g, args.warmup, args.steps, bool(args.graph))
    wall = step_s * args.steps
    st_last = g.stats()
    pairs_last = st_last.pair_evals
    fix_end, pos_end = g.forces_fixed(), g.particles()
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
        # candidate r2 tests of the last timed step per in-cutoff pair (whole 64-lane steps,
        # idle lanes of partly filled chunks included)
        "candidates_per_pair_eval": st_last.candidates / max(st_last.pair_evals, 1),
        "schedule": dict(sched.as_dict(), autotune_best_ms=tune_ms, autotune_cache_hit=tune_hit),
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
        out["batched_vs_plain"] = batched_vs_plain(args.replicas, max(args.warmup, 2),
                                                   args.replica_steps, args.seed, mdrl.Table.load())
    checks = None
    if args.workload == 3:
        out["precision"], relaxed_p, sets = precision_leg(s, relax, args.relax_step, args.nve_steps,
                                                          sorted({s.dt, 0.5}))
        checks = (relaxed_p, sets)
    if not args.no_cpu:
        from oracle import oracle as O
        base = cpu_baseline_leg(s, pairs, args.cpu_seconds)
        base["cmodel_agreement"] = cmodel_agreement(O)
        out["cpu_baseline"] = base
        out["cpu_celllist_baseline"] = cpu_celllist_leg(s, pairs, args.cpu_seconds / 3)
        out["speedup_vs_cpu_baseline"] = value / base["value"]
        # force error where the timed steps start (the relaxed state), after them, and on
        # the unrelaxed lattice; (i) against the same table model in FP64, (ii) against the
        # FP64 analytic direct evaluation, which the 1e-5 target is judged on
        fd_relaxed = fp64_reference(s, relaxed)
        states = {"relaxed_start_of_timed_steps": (fix_relaxed, relaxed, fd_relaxed),
                  "after_timed_steps": (fix_end, pos_end, None)}
        if relax:
            states["unrelaxed_lattice"] = (fix_unrelaxed, s.particle, None)
        tabular = not (args.analytic or args.fp32_atomics)
        errs = {}
        for name, (fx, xyz, fd) in states.items():
            e = force_error(fx, fd if fd is not None else fp64_reference(s, xyz))
            if tabular:
                e["kernel_arithmetic_vs_fp64_table_model"] = force_error(
                    fx, table_model_fp64(s, xyz), "FP64 evaluation of the same table model")["global_rel_l2"]
            errs[name] = e
        out["force_error"] = dict(errs["relaxed_start_of_timed_steps"],
                                  state="relaxed (the timed steps start here)")
        out["force_error_states"] = errs
        if checks:
            pos, sets = checks
            fp = fp64_reference(s, pos)
            for name, f in sets.items():
                out["precision"][name + "_vs_fp64_rel_l2"] = force_error(f, fp)["global_rel_l2"]
        if args.workload == 2 and tabular:
            out["analytic_mode"] = analytic_leg(s, relaxed, fd_relaxed, args.warmup,
                                                min(args.steps, 200) if args.steps > 200 else args.steps,
                                                base["value"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
