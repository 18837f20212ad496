"""GPU: batches of independent systems, force-kernel schedules and graph-captured steps.

Bars: int64 forces and trajectories bit-exact against the oracle per system and against
single-system runs; every schedule and the graph-captured step give identical bits.
"""
import numpy as np
import pytest

from oracle import oracle as O
from sc2019_md_amd import mdrl

from tests.test_oracle_golden import rl_from_synthetic  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def apoa1():
    t = mdrl.Table.load()
    return {k: t.get(k) for k in range(6)}


def say(*a):
    print(*a, flush=True)


def replicas(idx):
    return [mdrl.SyntheticSystem(4, k) for k in idx]


IDX = [0, 1, 2, 3, 5, 8, 13, 21, 34, 55]   # sizes from 1 to 5 cells per axis


def test_batch_forces_match_oracle(apoa1):
    systems = replicas(IDX)
    say("batch forces: %d systems, %d atoms" % (len(systems), sum(s.n for s in systems)))
    b = mdrl.Batch.from_synthetic(systems)
    total = 0
    for k, s in enumerate(systems):
        want, npairs = rl_from_synthetic(s, apoa1).forces_fix()
        got = b.system_forces_fixed(k)
        assert np.array_equal(got, want), "system %d (n=%d, L=%.1f)" % (k, s.n, s.box[0])
        total += npairs
    assert b.stats().pair_evals == total
    b.close()


def test_batch_with_an_anisotropic_system(apoa1):
    """A 3 x 5 x 7-cell system between two cubic replicas of one batch: its forces equal the
    oracle's, and six batched steps equal the same system stepped alone."""
    from tests.test_oracle_golden import anisotropic_system
    systems = [mdrl.SyntheticSystem(4, 0), anisotropic_system(), mdrl.SyntheticSystem(4, 3)]
    say("batch with a 3 x 5 x 7-cell system: %s atoms" % [s.n for s in systems])
    b = mdrl.Batch.from_synthetic(systems)
    for k, s in enumerate(systems):
        want, _ = rl_from_synthetic(s, apoa1).forces_fix()
        assert np.array_equal(b.system_forces_fixed(k), want), "system %d" % k
    b.step(6)
    for k, s in enumerate(systems):
        g = mdrl.System.from_synthetic(s)
        g.step(6)
        want = g.particles()
        g.close()
        got = b.system_particles(k)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), "system %d" % k
    b.close()


def test_batch_steps_match_single_systems():
    systems = replicas(IDX[:6])
    b = mdrl.Batch.from_synthetic(systems)
    b.step(6)
    for k, s in enumerate(systems):
        g = mdrl.System.from_synthetic(s)
        g.step(6)
        want = g.particles()
        g.close()
        got = b.system_particles(k)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), "system %d" % k
    b.close()


@pytest.mark.parametrize("which", ["batch", "single"])
def test_schedules_bit_identical(which):
    if which == "batch":
        g = mdrl.Batch.from_synthetic(replicas(IDX))
        get = lambda: np.concatenate([g.system_forces_fixed(k) for k in range(len(IDX))])
    else:
        g = mdrl.System.from_synthetic(mdrl.SyntheticSystem(2))
        get = g.forces_fixed
    ref = get()
    npairs = g.stats().pair_evals
    for apw in (16, 32, 64):
        for tile in (128, 256, 640):
            if tile == 640 and apw != 64:
                continue
            for groups in (1, 2, 4):
                for pers in (0, 1):
                    g.set_schedule(apw, tile, groups, pers)
                    g.force_eval()
                    assert np.array_equal(get(), ref), (apw, tile, groups, pers)
                    assert g.stats().pair_evals == npairs
    g.close()


def test_autotune_picks_a_schedule():
    g = mdrl.Batch.from_synthetic(replicas(IDX))
    ref = np.concatenate([g.system_forces_fixed(k) for k in range(len(IDX))])
    sc, ms = g.autotune(reps=2)
    say("autotune:", sc, "%.3f ms" % ms)
    assert ms > 0 and sc.atoms_per_wave in (16, 32, 64)
    assert np.array_equal(np.concatenate([g.system_forces_fixed(k) for k in range(len(IDX))]), ref)
    g.close()


CACHE_CHILD = """
import sys
sys.path.insert(0, %r)
from sc2019_md_amd import mdrl
g = mdrl.System.from_synthetic(mdrl.SyntheticSystem(1))
sc, ms, hit = g.autotune(reps=1, cache=True)
print(int(hit), sc.atoms_per_wave, sc.tile_depth, sc.offset_groups, sc.persistent, flush=True)
g.close()
"""


def test_autotune_cache(tmp_path, monkeypatch):
    """md_rl_autotune_cached: a miss sweeps and appends one JSON line to $MDRL_TUNE_CACHE
    (lines that do not parse are skipped); a second context of the same shape hits with the
    same schedule and bit-identical forces; a fresh process hits from the file alone; a
    batch of another shape misses and appends its own line."""
    import json
    import os
    import subprocess
    import sys
    path = tmp_path / "schedules.jsonl"
    path.write_text("not json\n{\"key\": \"truncated\n")
    monkeypatch.setenv("MDRL_TUNE_CACHE", str(path))
    s = mdrl.SyntheticSystem(1)
    g = mdrl.System.from_synthetic(s)
    sc, ms, hit = g.autotune(reps=1, cache=True)
    f0 = g.forces_fixed()
    g.close()
    say("autotune cache: miss ->", sc.as_dict(), "%.4f ms" % ms)
    assert not hit and ms > 0
    rows = [json.loads(l) for l in path.read_text().splitlines()[2:]]
    assert len(rows) == 1 and {k: rows[0][k] for k in sc.as_dict()} == sc.as_dict()
    assert "gfx950" in rows[0]["key"] and "atoms512 " in rows[0]["key"]

    h = mdrl.System.from_synthetic(s)
    sc2, ms2, hit2 = h.autotune(reps=1, cache=True)
    assert hit2 and sc2.as_dict() == sc.as_dict() and ms2 == ms
    assert np.array_equal(h.forces_fixed(), f0)
    h.close()

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", CACHE_CHILD % root], capture_output=True, text=True, timeout=120,
                       env=dict(os.environ, MDRL_TUNE_CACHE=str(path)))
    assert r.returncode == 0, r.stderr[-2000:]
    got = [int(v) for v in r.stdout.split()]
    assert got == [1, sc.atoms_per_wave, sc.tile_depth, sc.offset_groups, sc.persistent]

    b = mdrl.Batch.from_synthetic(replicas(IDX))
    _, _, hit3 = b.autotune(reps=1, cache=True)
    b.close()
    assert not hit3 and len(path.read_text().splitlines()) == 4


@pytest.mark.parametrize("which", ["batch", "single"])
def test_graph_steps_bit_identical(which):
    """hipGraph-captured steps (replayed many times, with eager work between replays)
    give the same bits as eager steps."""
    if which == "batch":
        mk = lambda: mdrl.Batch.from_synthetic(replicas(IDX[:6]))
        read = lambda g: np.concatenate([g.system_particles(k) for k in range(6)])
    else:
        mk = lambda: mdrl.System.from_synthetic(mdrl.SyntheticSystem(1))
        read = lambda g: g.particles()
    a = mk()
    a.step(11)
    a.energy()
    a.step(10)
    a.step(37)
    want = read(a)
    a.close()
    b = mk()
    b.set_graph(True)
    b.step(11)
    b.energy()
    b.step(10)
    b.step(37)          # two 16-step graphs, two 2-step graphs, one eager step
    got = read(b)
    b.close()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_batch_energies_deterministic():
    systems = replicas(IDX[:4])
    b = mdrl.Batch.from_synthetic(systems)
    e1 = [b.system_energy(k) for k in range(4)]
    e2 = [b.system_energy(k) for k in range(4)]
    for (p1, k1, v1), (p2, k2, v2) in zip(e1, e2):
        assert p1 == p2 and k1 == k2 and np.array_equal(v1, v2)
    pe, ke, vir = b.energy()
    assert pe == pytest.approx(sum(e[0] for e in e1), rel=1e-12)
    vsum = sum(e[2] for e in e1)
    assert np.abs(vir - vsum).max() <= 1e-12 * np.abs(vsum).max()
    for k, s in enumerate(systems):
        _, e64, v64, _ = O.fp64_direct(s.particle[:, :3], s.type, s.eps, s.sigma, s.box, s.rc,
                                       charge=s.charge, ke=s.coulomb_k)
        assert e1[k][0] == pytest.approx(e64, rel=2e-4, abs=1e-3 * abs(e64) + 1.0)
        assert np.linalg.norm(e1[k][2] - v64) <= 1e-3 * np.linalg.norm(v64) + 1.0, "system %d" % k
    b.close()


def test_batch_with_empty_and_tiny_systems(apoa1):
    """A batch may hold systems with no atom, one atom or two atoms beside ordinary ones:
    the ordinary systems keep the bits of their single-system runs, the lone atom feels no
    force from its own images (box > rc), the two atoms feel equal and opposite forces, and
    the empty system reports no atoms and zero energy."""
    big = mdrl.SyntheticSystem(4, 3)
    s0 = big
    eps, sig, mass = s0.eps, s0.sigma, s0.mass
    sizes = [s0.n, 0, 1, 2]
    boxes = [s0.box, [22.0, 22.0, 22.0], [25.0, 25.0, 25.0], [21.0, 21.0, 21.0]]
    b = mdrl.Batch(sizes, boxes, s0.rc, eps, sig, mass, dt=s0.dt, coulomb_k=s0.coulomb_k)
    b.set_system(0, s0.particle, s0.type, s0.charge)
    one = np.zeros((1, 11), np.float32)
    one[0, :3] = [3.0, 4.0, 5.0]
    b.set_system(2, one, np.zeros(1, np.int32), np.zeros(1, np.float32))
    two = np.zeros((2, 11), np.float32)
    two[0, :3] = [20.5, 10.0, 10.0]
    two[1, :3] = [2.0, 10.5, 10.0]                   # 2.5 A away through the x faces
    b.set_system(3, two, np.array([1, 2], np.int32), np.array([0.3, -0.3], np.float32))
    b.commit()
    b.step(4)
    g = mdrl.System.from_synthetic(s0)
    g.step(4)
    assert np.array_equal(b.system_particles(0).view(np.uint32), g.particles().view(np.uint32))
    assert np.array_equal(b.system_forces_fixed(0), g.forces_fixed())
    g.close()
    assert b.system_particles(1).shape == (0, 11)
    pe, ke, _ = b.system_energy(1)
    assert pe == 0.0 and ke == 0.0
    assert np.all(b.system_forces_fixed(2) == 0)
    f2 = b.system_forces_fixed(3)
    assert np.all(f2[0] == -f2[1]) and np.any(f2 != 0)
    b.close()


def test_prepared_graphs_bit_identical():
    """md_rl_graph_prepare captures the graphs of the next 37 steps without running them
    (the state is unchanged), and the steps that then replay them give the eager bits."""
    s = mdrl.SyntheticSystem(1)
    a = mdrl.System.from_synthetic(s)
    a.step(38)
    want = a.particles()
    a.close()
    b = mdrl.System.from_synthetic(s)
    b.set_graph(True)
    b.step(1)
    before = b.particles()
    b.prepare_graphs(37)          # one 16-step graph and one 2-step graph
    assert np.array_equal(b.particles().view(np.uint32), before.view(np.uint32))
    b.step(37)                    # two 16-step replays, two 2-step replays, one eager step
    got = b.particles()
    b.close()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
