"""GPU: the batched workloads at their full sizes against the oracle.

- 512 independent LJ + Coulomb replicas (log-uniform 256..16,384 atoms, seed 2019,
  1,909,304 atoms): one persistent launch of ~10k tasks, several queue passes per workgroup.
- 128 bonded replicas (64 chains x 48 beads each, 393,216 beads, 1.13 M bonded terms),
  with exclusion windows and the bonded blocks as tasks of the same persistent launch.

The bonded workload runs once more on its coarse grid (MDRL_COARSE_CELLS: 18 cells per
replica instead of 64), against the oracle on that grid.

Each batch runs the schedule its autotuner picks (the one the throughput run times).  Bars,
per system: int64 forces bit-exact against the oracle after one evaluation; positions (bits)
and forces bit-exact after five steps, the last four as two replays of a captured 2-step
hipGraph; the batch's pair count equal to the oracle's total.  The many-cell scheduling
that this exercises is the FPGA's, D/RL_LJ_Topest_Top_64_Cells.v:312-406 and
Bonded/Bonded_Top.v:202-380.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from sc2019_md_amd import mdrl

from tests.test_gpu_api import coarse_grid  # noqa: E402
from tests.test_oracle_golden import apoa1_tab6  # noqa: E402

pytestmark = pytest.mark.gpu

NREP_LJ, NREP_BONDED = 512, 128
STEPS = 5


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def apoa1():
    t = mdrl.Table.load()
    return {k: t.get(k) for k in range(6)}


def grid(s, coarse):
    """The system's cell grid: floor(L / rc) per axis, or the MDRL_COARSE_CELLS grid."""
    if coarse:
        return coarse_grid(s.n, s.box, s.rc)
    return [max(1, int(math.floor(float(s.box[d]) / s.rc))) for d in range(3)]


def oracle_fix(s, apoa1, pos4=None, coarse=False):
    """int64 forces and pair count of one system (range-limited with its exclusion windows,
    plus bonded terms), at the positions pos4 (default: the system's own)."""
    ncell = grid(s, coarse)
    x = s.particle[:, :3] if pos4 is None else pos4[:, :3]
    rl = O.RL(x, s.box, ncell, s.rc, s.type, s.eps, s.sigma, apoa1_tab6(apoa1),
              charge=s.charge if s.coulomb_k else None, coulomb_k=s.coulomb_k, excl=s.excl)
    fix, npairs = rl.forces_fix()
    if s.nterms:
        bfix, _ = O.bonded_fix(rl.pos4, s.box, s.term_atoms, s.term_params)
        fix = (fix.view(np.uint64) + bfix.view(np.uint64)).view(np.int64)
    return fix, npairs, rl.pos4


def oracle_trajectory(s, apoa1, nsteps, coarse=False):
    """Velocity Verlet in the oracle: positions (float4) and forces after nsteps."""
    fix, _, pos4 = oracle_fix(s, apoa1, coarse=coarse)
    pos4 = pos4.copy()
    vel = s.particle[:, 3:6].copy()
    hk = O.half_kick(s.mass, s.dt)
    for k in range(nsteps):
        pos4, vel, _ = O.motion_update(pos4, vel, fix, s.type, hk, s.mass, s.dt, s.box, k == 0)
        fix, _, _ = oracle_fix(s, apoa1, pos4, coarse)
    return pos4, fix


def check_batch(which, nrep, apoa1, coarse=False):
    systems = [mdrl.SyntheticSystem(which, k) for k in range(nrep)]
    natoms = sum(s.n for s in systems)
    say("batch %d: %d replicas, %d atoms, %d bonded terms, %s grid"
        % (which, nrep, natoms, sum(s.nterms for s in systems), "coarse" if coarse else "floor(L/rc)"))
    b = mdrl.Batch.from_synthetic(systems, flags=mdrl.COARSE_CELLS if coarse else 0)
    sched, ms = b.autotune(reps=1)
    say("  autotuned schedule %s (%.3f ms)" % (sched, ms))
    # one evaluation at the initial positions (the oracle runs the replicas side by side)
    ref = O.map_systems(lambda s: oracle_fix(s, apoa1, coarse=coarse)[:2], systems)
    total = 0
    for k, s in enumerate(systems):
        want, npairs = ref[k]
        got = b.system_forces_fixed(k)
        bad = np.nonzero(np.any(got != want, axis=1))[0]
        assert bad.size == 0, "replica %d (n=%d, L=%.2f): atoms %s differ" % (k, s.n, s.box[0], bad[:8])
        total += npairs
    assert b.stats().pair_evals == total
    say("  forces of %d replicas bit-exact, %d pair-evals" % (nrep, total))
    # five steps: one eager, then two replays of a captured 2-step graph
    b.set_graph(True)
    b.step(STEPS)
    traj = O.map_systems(lambda s: oracle_trajectory(s, apoa1, STEPS, coarse), systems)
    for k, s in enumerate(systems):
        pos4, fix = traj[k]
        got = b.system_particles(k)
        assert np.array_equal(got[:, :3].view(np.uint32), pos4[:, :3].view(np.uint32)), \
            "replica %d: positions after %d steps differ" % (k, STEPS)
        assert np.array_equal(b.system_forces_fixed(k), fix), "replica %d: forces after %d steps" % (k, STEPS)
    say("  %d graph-captured steps of %d replicas bit-exact" % (STEPS, nrep))
    b.close()


def test_512_replicas_full_size(apoa1):
    check_batch(4, NREP_LJ, apoa1)


def test_128_bonded_replicas_full_size(apoa1):
    check_batch(5, NREP_BONDED, apoa1)


def test_bonded_replicas_coarse_cells(apoa1):
    check_batch(5, 32, apoa1, coarse=True)
