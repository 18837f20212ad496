#!/usr/bin/env python3
"""Time-step stability of the mixed-species LJ + Coulomb workloads, on the CPU.

    python oracle/stability_study.py [--replica 40] [--steps 400] [--dt 1.0] [--tables]

Test infrastructure (uses the oracle only).  Builds one replica of the batched set (same
species, charge and lattice recipe as the 92,224-atom and 1,048,576-atom systems),
relaxes it with the same capped steepest descent as bench.py (displacement <= 0.02 A,
step <= 1e-4 A^2 mol/kJ), then integrates velocity Verlet in FP64 with either the FP64
analytic forces (default) or the table forces of the force path (--tables), and prints
potential, kinetic and total energy every 25 steps.

Result for replica 40 (1,468 atoms) with the workload's masses (39.948 amu for every type):
at dt = 1 fs the trajectory stays bounded over 400 steps (total energy -7.25e4 ->
-8.26e4 kJ/mol while the strained relaxed lattice releases its energy).  The same replica
with carbon-like masses (12-16 amu) also stays bounded over 600 steps, but the
92,224-atom system with those masses diverges within 50 steps on the force path: among
92,224 atoms the N(0, 0.4^2) charges reach ~1.8 e, and the Coulomb-bound contacts of the
sigma = 1 A type become too stiff for velocity Verlet at 1 fs (omega*dt > 2).  The heavier
masses lower omega by sqrt(39.948/12.011) = 1.8.
"""
import argparse
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from sc2019_md_amd import mdrl  # noqa: E402

ACC = 1e-4   # (kJ/mol/A)/amu -> A/fs^2


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--replica", type=int, default=40)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--dt", type=float, default=1.0)
    ap.add_argument("--relax", type=int, default=300)
    ap.add_argument("--tables", action="store_true", help="table forces of the force path")
    a = ap.parse_args()

    s = mdrl.SyntheticSystem(4, a.replica)
    box = s.box.astype(np.float64)
    x = s.particle[:, :3].astype(np.float64).copy()
    v = s.particle[:, 3:6].astype(np.float64).copy()
    m = np.asarray(s.mass, np.float64)[s.type]
    ncell = [max(1, int(math.floor(float(s.box[d]) / s.rc))) for d in range(3)]
    tab6 = None
    if a.tables:
        t = mdrl.Table.load()
        tab6 = O.interleave([*t.get(mdrl.TERM_14), *t.get(mdrl.TERM_8), *t.get(mdrl.TERM_3)])

    def forces(x):
        x32 = x.astype(np.float32)
        f, e, _, _ = O.fp64_direct(x32, s.type, s.eps, s.sigma, s.box, s.rc,
                                   charge=s.charge, ke=s.coulomb_k)
        if a.tables:
            rl = O.RL(x32, s.box, ncell, s.rc, s.type, s.eps, s.sigma, tab6,
                      charge=s.charge, coulomb_k=s.coulomb_k)
            fix, _ = rl.forces_fix()
            f = fix.astype(np.float64) / 2 ** 32 * 0.1   # kJ/mol/nm -> kJ/mol/A
        return f, e

    for _ in range(a.relax):
        f, _ = forces(x)
        fm = np.linalg.norm(f, axis=1).max()
        x += f * min(0.02 / fm, 1e-4)
        x -= np.floor(x / box) * box
    f, e = forces(x)
    ke = lambda v: 0.5 * (m[:, None] * v ** 2).sum() * 1e4
    print("replica %d: %d atoms, box %.2f A, dt %.2f fs, %s forces" % (
        a.replica, s.n, box[0], a.dt, "table" if a.tables else "FP64 analytic"))
    print("step 0 PE %.6e KE %.6e E %.6e" % (e, ke(v), e + ke(v)), flush=True)
    for it in range(1, a.steps + 1):
        v += 0.5 * a.dt * ACC * f / m[:, None]
        x += a.dt * v
        x -= np.floor(x / box) * box
        f, e = forces(x)
        v += 0.5 * a.dt * ACC * f / m[:, None]
        if it % 25 == 0:
            print("step %d PE %.6e KE %.6e E %.6e" % (it, e, ke(v), e + ke(v)), flush=True)


if __name__ == "__main__":
    main()
