"""GPU: the half-shell force path against the reference C model's arithmetic, per atom, on
the 512-atom argon system (34 A box, 4 x 4 x 4 cells of rc = 8.5 A, sigma 3.4 A,
eps 0.996 kJ/mol).

The C model (R/C_Model_ForceEval/MD_Evaluation_Model.cpp:72-124) evaluates every ordered
pair with pow() in double, FP32 sums in input order and minimum-image PBC through (int)
casts; its oracle restatement is pinned against the reference binary.  Here its force loop
runs with this system's constants on the same positions as the GPU, at the start and after
100 velocity-Verlet steps of 2 fs on the GPU.  Per-atom relative difference
|F_gpu - F_cmodel| / |F_cmodel|: median, p99 and max, and the global relative L2.

Tolerances (stated here, from the distances measured on the CPU with the oracle, which the
GPU reproduces bit for bit, with about 2x margin):
- FP32 analytic pair function: global <= 1e-5 (the stated agreement tolerance), median
  <= 5e-6, p99 <= 5e-5, max <= 2e-4.  The C model itself sits 1.8e-6 (global) and 1.2e-5
  (p99) from an FP64 evaluation, so the per-atom tail is the C model's own FP32 rounding.
- the reference's first-order tables: global <= 1.5e-4, median <= 2e-4, p99 <= 1.5e-3,
  max <= 5e-3.  The tables carry up to 1e-4 relative error on the r^-14 term, which the
  cancellation of the r^-14 and r^-8 terms near the LJ minimum amplifies; this bounds
  agreement with any exact evaluation.
"""
import numpy as np
import pytest

from oracle import oracle as O
from sc2019_md_amd import mdrl

pytestmark = pytest.mark.gpu

TOL = {
    "analytic": {"global": 1e-5, "median": 5e-6, "p99": 5e-5, "max": 2e-4},
    "table": {"global": 1.5e-4, "median": 2e-4, "p99": 1.5e-3, "max": 5e-3},
}
FLAGS = {"analytic": mdrl.EVAL_ANALYTIC, "table": 0}


def say(*a):
    print(*a, flush=True)


def cmodel_forces(s, xyz):
    """The C model's force loop over all 512 reference atoms (kJ/mol/A)."""
    p = O.cmodel_default_params()
    p.eps, p.sigma, p.rc = float(s.eps[0]), float(s.sigma[0]), float(s.rc)
    p.box[:] = [float(b) for b in s.box]
    part = np.zeros((s.n, 11), np.float32)
    part[:, :3] = xyz
    cnt, f = O.cmodel_force_sample(part, p, s.n)
    return f.astype(np.float64), cnt


def agreement(f_fix, fc):
    f = f_fix.astype(np.float64) / 2 ** 32 * 0.1          # 2^-32 kJ/mol/nm -> kJ/mol/A
    per = np.linalg.norm(f - fc, axis=1) / np.linalg.norm(fc, axis=1)
    return {"global": float(np.linalg.norm(f - fc) / np.linalg.norm(fc)),
            "median": float(np.median(per)), "p99": float(np.percentile(per, 99)),
            "max": float(per.max())}


@pytest.mark.parametrize("mode", ["analytic", "table"])
def test_half_shell_agrees_with_the_c_model(mode):
    s = mdrl.SyntheticSystem(1)
    g = mdrl.System.from_synthetic(s, flags=FLAGS[mode])
    for when in ("start", "after 100 steps"):
        if when != "start":
            g.step(100)
        xyz = g.particles()[:, :3]
        fc, ordered = cmodel_forces(s, xyz)
        got = g.forces_fixed()
        # the same pair set up to FP32 ties at the cutoff (the C model rounds its own r2)
        assert abs(g.stats().pair_evals * 2 - ordered) <= 4, "pair sets differ"
        a = agreement(got, fc)
        say("%s, %s: global %.2e, per atom median %.2e, p99 %.2e, max %.2e"
            % (mode, when, a["global"], a["median"], a["p99"], a["max"]))
        for k, tol in TOL[mode].items():
            assert a[k] <= tol, "%s %s: %s %.3e > %.1e" % (mode, when, k, a[k], tol)
    g.close()
