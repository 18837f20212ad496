"""GPU: the 1,048,576-atom precision configuration (219.0 A box, rc 12 A, 18^3 cells).

Bars: int64 forces bit-exact against the oracle and bit-identical over 10 repeat
evaluations (hashes of the raw int64 arrays), Newton's third law exact (the int64 sum of
all forces is zero); FP32-atomics forces within 1e-5 relative L2 of the int64 forces; the
FP64-reduced potential energy within 2e-4 of the FP64 analytic energy.  The 92,224-atom
system: 34 steps bit-identical between two contexts, eager and graph-captured, and the
shifted energy conserved over 2,000 steps of NVE.
"""
import hashlib

import numpy as np
import pytest

from oracle import oracle as O
from sc2019_md_amd import mdrl

from tests.test_oracle_golden import rl_from_synthetic  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def big():
    return mdrl.SyntheticSystem(3)


@pytest.fixture(scope="module")
def apoa1():
    t = mdrl.Table.load()
    return {k: t.get(k) for k in range(6)}


@pytest.fixture(scope="module")
def int64_forces(big):
    g = mdrl.System.from_synthetic(big)
    f = g.forces_fixed()
    npairs = g.stats().pair_evals
    g.close()
    return f, npairs


def test_1m_forces_repeatable(big):
    """Ten force evaluations of the same positions give bit-identical int64 forces."""
    say("1M atoms: 10 repeat force evaluations")
    g = mdrl.System.from_synthetic(big)
    digests = set()
    for _ in range(10):
        g.force_eval()
        digests.add(hashlib.sha256(g.forces_fixed().tobytes()).hexdigest())
    g.close()
    assert len(digests) == 1


def test_1m_forces_match_oracle(big, apoa1, int64_forces):
    say("1M atoms: GPU int64 forces vs the oracle")
    got, npairs = int64_forces
    want, want_pairs = rl_from_synthetic(big, apoa1).forces_fix()
    assert npairs == want_pairs
    bad = np.nonzero(np.any(got != want, axis=1))[0]
    assert bad.size == 0, "atoms differ: %s" % bad[:10]
    # Newton's third law holds exactly in fixed point (wrap-around sum)
    assert np.all(got.sum(axis=0, dtype=np.int64) == 0)


def test_1m_fp32_atomics_close(big, int64_forces):
    """FP32 atomics sum in arrival order: close to the int64 forces, not bit-stable."""
    a = int64_forces[0].astype(np.float64)
    for _ in range(2):
        g = mdrl.System.from_synthetic(big, flags=mdrl.ACC_FP32)
        b = g.forces_fixed().astype(np.float64)
        g.close()
        rel = np.linalg.norm(a - b) / np.linalg.norm(a)
        say("1M atoms: FP32 atomics vs int64, relative L2 %.3e" % rel)
        assert rel < 1e-5


def test_1m_energy_vs_fp64(big):
    say("1M atoms: FP64-reduced energy vs FP64 analytic")
    g = mdrl.System.from_synthetic(big)
    pe, ke, _ = g.energy()
    pe2, ke2, _ = g.energy()
    g.close()
    assert pe == pe2 and ke == ke2
    _, e64, _, _ = O.fp64_direct(big.particle[:, :3], big.type, big.eps, big.sigma, big.box,
                                 big.rc, charge=big.charge, ke=big.coulomb_k)
    assert pe == pytest.approx(e64, rel=2e-4)


def test_92k_steps_deterministic_and_graph_equal():
    """The 92,224-atom system, relaxed and then stepped 34 times at 1 fs: two contexts give
    bit-identical positions, velocities and forces, one stepping eagerly and one through
    captured graphs (one 16-step graph, 8 two-step graphs, an eager first step and two
    eager steps after energy reads); Newton's third law holds exactly at the end."""
    s = mdrl.SyntheticSystem(2)
    runs = []
    for graph in (False, True):
        g = mdrl.System.from_synthetic(s)
        g.minimize(20, 0.02, 1e-4)
        g.set_graph(graph)
        g.step(17)
        g.energy()
        g.step(16)
        g.energy()
        g.step(1)
        runs.append((g.particles(), g.forces_fixed(), g.stats().pair_evals))
        g.close()
    (p0, f0, n0), (p1, f1, n1) = runs
    say("92k steps: pairs %d" % n0)
    assert np.all(np.isfinite(p0[:, :6]))
    assert np.array_equal(p0.view(np.uint32), p1.view(np.uint32))
    assert np.array_equal(f0, f1) and n0 == n1
    assert np.all(f0.sum(axis=0, dtype=np.int64) == 0)


def test_92k_nve_shifted_energy_conserved():
    """The 92,224-atom LJ + Coulomb system (four species, rc 12 A), relaxed, then 2,000
    velocity-Verlet steps of 1 fs through captured graphs, sampled every 200 steps.  The reference potential is
    truncated without a shift, so KE + PE jumps by U(rc) at every cutoff crossing; the
    shifted energy KE + PE - cutoff_shift (one U(rc) per in-cutoff pair) is what the
    integrator conserves.  Bars, relative to the initial kinetic energy KE0: the shifted
    energy stays within 1% at every sample, and the raw energy's largest excursion is at
    least 10x the shifted energy's (the shift accounts for the jumps).  The kinetic energy
    roughly doubles over the run as the relaxed lattice equilibrates."""
    s = mdrl.SyntheticSystem(2)
    g = mdrl.System.from_synthetic(s)
    g.minimize(200, 0.02, 1e-4)
    g.set_graph(True)
    t = g.energy_terms()
    k0, p0, c0 = t["kinetic"], t["potential"], t["cutoff_shift"]
    raw, shifted = [], []
    for k in range(10):
        g.step(200)
        t = g.energy_terms()
        raw.append((t["kinetic"] + t["potential"] - k0 - p0) / k0)
        shifted.append((t["kinetic"] + t["potential"] - t["cutoff_shift"] - (k0 + p0 - c0)) / k0)
        say("NVE 92k step %d: raw %+.3e shifted %+.3e KE/KE0 %.3f"
            % ((k + 1) * 200, raw[-1], shifted[-1], t["kinetic"] / k0))
    p = g.particles()
    g.close()
    assert np.all(np.isfinite(p[:, :6]))
    assert max(abs(x) for x in shifted) < 1e-2
    assert max(abs(x) for x in raw) > 10 * max(abs(x) for x in shifted)
