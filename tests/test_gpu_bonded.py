"""GPU: bonded chains (64 chains x 48 beads per system) -- bonded terms and nonbonded
exclusions on the HIP path against the oracle.

Bars: int64 forces (range-limited with exclusion windows + bonded terms) and pair counts
bit-exact against the oracle, for one system and for a batch; velocity-Verlet trajectories
bit-exact; every term's fixed-point forces sum to zero, so the total force is exactly zero;
the FP64 potential energy within 2e-4 of the FP64 analytic oracle (which does not use the
tables); a relaxed system keeps its shifted total energy within 1% of the initial kinetic
energy over 2,000 steps of 2 fs.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from sc2019_md_amd import mdrl

from tests.test_oracle_golden import apoa1_tab6  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def apoa1():
    t = mdrl.Table.load()
    return {k: t.get(k) for k in range(6)}


def oracle_forces(s, apoa1, pos4=None, mode=0):
    ncell = [int(math.floor(float(s.box[d]) / s.rc)) for d in range(3)]
    x = s.particle[:, :3] if pos4 is None else pos4[:, :3]
    rl = O.RL(x, s.box, ncell, s.rc, s.type, s.eps, s.sigma, apoa1_tab6(apoa1),
              charge=s.charge, coulomb_k=s.coulomb_k, excl=s.excl, mode=mode)
    fix, npairs = rl.forces_fix()
    bfix, te = O.bonded_fix(rl.pos4, s.box, s.term_atoms, s.term_params)
    tot = (fix.view(np.uint64) + bfix.view(np.uint64)).view(np.int64)
    return tot, npairs, te, rl


@pytest.mark.parametrize("index", [0, 1])
def test_chain_forces_bit_exact(apoa1, index):
    s = mdrl.SyntheticSystem(5, index)
    say("chains: replica %d, %d beads, %d bonded terms" % (index, s.n, s.nterms))
    g = mdrl.System.from_synthetic(s)
    got = g.forces_fixed()
    npairs = g.stats().pair_evals
    g.close()
    want, wpairs, _, _ = oracle_forces(s, apoa1)
    assert npairs == wpairs
    bad = np.nonzero(np.any(got != want, axis=1))[0]
    assert bad.size == 0, "atoms differ: %s" % bad[:10]
    assert np.all(got.sum(axis=0) == 0)


def test_chain_forces_every_schedule(apoa1):
    """Exclusion windows and bonded blocks (tasks of the force kernel on this path) under
    every force-kernel schedule, persistent or one workgroup per task: int64 forces and pair
    counts equal the oracle's."""
    s = mdrl.SyntheticSystem(5, 0)
    want, wpairs, _, _ = oracle_forces(s, apoa1)
    g = mdrl.System.from_synthetic(s)
    for sc in [(a, t, gr, p) for a in (16, 32, 64) for t in (128, 256, 640) for gr in (1, 4) for p in (0, 1)
               if t != 640 or a == 64]:
        say("chains: schedule", sc)
        g.set_schedule(*sc)
        g.force_eval()
        got = g.forces_fixed()
        assert g.stats().pair_evals == wpairs, sc
        bad = np.nonzero(np.any(got != want, axis=1))[0]
        assert bad.size == 0, "schedule %s: atoms differ: %s" % (sc, bad[:10])
    g.close()


@pytest.mark.parametrize("mode", ["fast", "parity", "analytic"])
def test_chain_forces_without_exclusions(apoa1, mode):
    """Bonded terms on a system without exclusion windows, in each pair mode (the bonded
    blocks then run as a kernel of their own beside the force kernel): int64 forces and
    pair counts equal the oracle's, and the total force is exactly zero."""
    s = mdrl.SyntheticSystem(5, 1)
    s.excl = None
    flags = {"fast": 0, "parity": mdrl.EVAL_PARITY, "analytic": mdrl.EVAL_ANALYTIC}[mode]
    omode = {"fast": 0, "parity": O.MODE_PARITY, "analytic": O.MODE_ANALYTIC}[mode]
    say("chains without exclusions: mode %s" % mode)
    g = mdrl.System.from_synthetic(s, flags=flags)
    want, wpairs, _, _ = oracle_forces(s, apoa1, mode=omode)
    scheds = {"fast": [(64, 256, 1, 1), (64, 640, 1, 1)], "parity": [(64, 256, 1, 1), (64, 640, 1, 1)],
              "analytic": [(64, 256, 1, 1), (64, 640, 1, 1)]}[mode]
    for sc in scheds:
        g.set_schedule(*sc)
        g.force_eval()
        got = g.forces_fixed()
        assert g.stats().pair_evals == wpairs, sc
        bad = np.nonzero(np.any(got != want, axis=1))[0]
        assert bad.size == 0, "schedule %s: atoms differ: %s" % (sc, bad[:10])
        assert np.all(got.sum(axis=0) == 0)
    g.close()


def test_chain_forces_fp32_atomics(apoa1):
    """FP32-atomics accumulation with bonded terms (no exclusion windows): within 1e-5
    (relative L2) of the int64 forces."""
    s = mdrl.SyntheticSystem(5, 1)
    s.excl = None
    a = mdrl.System.from_synthetic(s)
    fa = a.forces_fixed().astype(np.float64)
    a.close()
    b = mdrl.System.from_synthetic(s, flags=mdrl.ACC_FP32)
    fb = b.forces_fixed().astype(np.float64)
    b.close()
    assert np.linalg.norm(fa - fb) / np.linalg.norm(fa) < 1e-5


def test_chain_steps_bit_exact(apoa1):
    s = mdrl.SyntheticSystem(5, 2)
    nsteps = 5
    say("chains: %d steps of %.1f fs" % (nsteps, s.dt))
    g = mdrl.System.from_synthetic(s)
    fix, _, _, rl = oracle_forces(s, apoa1)
    hk = O.half_kick(s.mass, s.dt)
    pos4, vel = rl.pos4.copy(), s.particle[:, 3:6].copy()
    for k in range(nsteps):
        pos4, vel, _ = O.motion_update(pos4, vel, fix, s.type, hk, s.mass, s.dt, s.box, k == 0)
        fix, _, _, _ = oracle_forces(s, apoa1, pos4)
    g.step(nsteps)
    got = g.particles()
    gfix = g.forces_fixed()
    g.close()
    assert np.array_equal(got[:, :3].view(np.uint32), pos4[:, :3].view(np.uint32))
    assert np.array_equal(gfix, fix)


def test_chain_batch_matches_single():
    systems = [mdrl.SyntheticSystem(5, k) for k in range(4)]
    b = mdrl.Batch.from_synthetic(systems)
    b.step(3)
    for k, s in enumerate(systems):
        g = mdrl.System.from_synthetic(s)
        g.step(3)
        assert np.array_equal(b.system_particles(k).view(np.uint32), g.particles().view(np.uint32)), k
        assert np.array_equal(b.system_forces_fixed(k), g.forces_fixed()), k
        g.close()
    b.close()


def test_chain_energy(apoa1):
    s = mdrl.SyntheticSystem(5, 3)
    g = mdrl.System.from_synthetic(s)
    pe, ke, _ = g.energy()
    pe2, _, _ = g.energy()
    g.close()
    assert pe == pe2
    _, e64, _, _ = O.fp64_direct(s.particle[:, :3], s.type, s.eps, s.sigma, s.box, s.rc,
                                 charge=s.charge, ke=s.coulomb_k, excl=s.excl)
    _, _, te, _ = oracle_forces(s, apoa1)
    _, eb = O.bonded_fp64(s.particle[:, :3].astype(np.float64), s.box, s.term_atoms, s.term_params)
    assert te.sum() == pytest.approx(eb, rel=1e-5)
    say("chains: PE %.3f (oracle RL %.3f + bonded %.3f)" % (pe, e64, eb))
    assert pe == pytest.approx(e64 + eb, rel=2e-4, abs=1e-3 * abs(eb))


def cutoff_shift(s, particle):
    """Sum over in-cutoff, non-excluded pairs of the pair energy at rc: the truncated
    potential jumps by that much whenever a pair crosses the cutoff, so PE minus it is the
    energy of the shifted potential, whose total with KE velocity Verlet conserves."""
    from scipy.spatial import cKDTree
    x = particle[:, :3].astype(np.float64)
    L = float(s.box[0])
    pairs = cKDTree(x % L, boxsize=L).query_pairs(s.rc, output_type="ndarray")
    i, j = pairs[:, 0], pairs[:, 1]
    d = j - i + 16
    inwin = (d >= 0) & (d < 32)
    excl = np.zeros(len(i), bool)
    excl[inwin] = ((s.excl[i[inwin]].astype(np.int64) >> d[inwin]) & 1) == 1
    i, j = i[~excl], j[~excl]
    sig, eps = float(s.sigma[0]), float(s.eps[0])
    u_lj = 4 * eps * ((sig / s.rc) ** 12 - (sig / s.rc) ** 6)
    qq = s.charge[i].astype(np.float64) * s.charge[j]
    return len(i) * u_lj + s.coulomb_k * qq.sum() / s.rc


def test_chain_nve():
    """2,000 steps of 2 fs after relaxation: the shifted total energy stays within 1% of
    the kinetic energy (the raw total moves by k q q / rc ~ 8.7 kJ/mol per pair crossing
    the cutoff, with no drift)."""
    s = mdrl.SyntheticSystem(5, 4)
    g = mdrl.System.from_synthetic(s)
    g.minimize(2000, 0.02, 1e-4)
    pe0, ke0, _ = g.energy()
    e0 = pe0 + ke0 - cutoff_shift(s, g.particles())
    g.set_graph(True)
    dev, raw = 0.0, 0.0
    for _ in range(10):
        g.step(200)
        pe, ke, _ = g.energy()
        e = pe + ke - cutoff_shift(s, g.particles())
        dev = max(dev, abs(e - e0))
        raw = max(raw, abs(pe + ke - pe0 - ke0))
    g.close()
    say("chains NVE: KE0 %.1f, max |E - E0| over 2,000 steps: shifted %.2f, raw %.2f" % (ke0, dev, raw))
    assert dev < 0.01 * ke0
