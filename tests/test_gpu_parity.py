"""GPU parity: the HIP path through the C ABI against the oracle (MI355X only).

Bars: bit-exact for int64 fixed-point forces, positions, velocities, pair counts and the
golden pair records; FP32-atomics forces within 1e-5 relative L2 of the int64 forces
(their order of summation is not fixed); the C model bit-exact in its printed energies.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from sc2019_md_amd import mdrl

pytestmark = pytest.mark.gpu

from tests.test_oracle_golden import (  # noqa: E402
    anisotropic_system, apoa1_tab6, cmodel_inputs, cutoff_edge_pairs, face_and_interior_pairs,
    hex2f, parse_pdb_float32,
    parse_reference_output, pdb_text, rl_from_synthetic)


@pytest.fixture(scope="module")
def table():
    return mdrl.Table.load()


@pytest.fixture(scope="module")
def apoa1(table):
    return {k: table.get(k) for k in range(6)}


def say(*a):
    print(*a, flush=True)


# ------------------------------------------------------------------------ pair batch
def test_eval_pairs_golden_bit_exact(table):
    """md_rl_eval_pairs (the OpenCL LJ kernel's interface) reproduces all 10,000 records."""
    say("pairs: golden 10,000 records, parity+strict")
    recs = parse_reference_output()
    ref = hex2f(recs[:, 0:3].ravel()).reshape(-1, 3)
    nbr = hex2f(recs[:, 3:6].ravel()).reshape(-1, 3)
    out = mdrl.eval_pairs(table, ref, nbr, 196.0, mdrl.PAIR_PARITY | mdrl.PAIR_STRICT,
                          ref_id=np.arange(10000) // 100, nbr_id=np.arange(10000) % 100)
    want = hex2f(recs[:, 7:10].ravel()).reshape(-1, 3)
    assert np.array_equal(out.view(np.uint32), want.view(np.uint32))
    fs, fx = mdrl.eval_pairs_ex(table, ref, nbr, 196.0, mdrl.PAIR_PARITY | mdrl.PAIR_STRICT)
    assert np.array_equal(fs.view(np.uint32), hex2f(recs[:, 6]).view(np.uint32))


def test_eval_pairs_testbench_vectors(table, apoa1):
    """The known answers of the reference's r2 and pair-evaluation testbenches through the
    pair API, against the oracle (whose r2 and d reproduce the testbench comments)."""
    from tests.test_oracle_golden import TB_NBR, TB_REF
    want = O.pair_golden(*apoa1[mdrl.TERM_14], *apoa1[mdrl.TERM_8], TB_REF, TB_NBR, 196.0)
    fs, fx = mdrl.eval_pairs_ex(table, TB_REF, TB_NBR, 196.0, mdrl.PAIR_PARITY | mdrl.PAIR_STRICT)
    assert np.array_equal(fs.view(np.uint32), want[:, 0].view(np.uint32))
    assert np.array_equal(fx.view(np.uint32), want[:, 1:4].view(np.uint32))


def test_testbench_periodic_vector(apoa1):
    """r2_compute_with_pbc_tb.v:129-140: ref (59, 50.5, 50.5) and nbr (0, 0, 0) in the
    59.5 x 51 x 51 A box of the LJArgon-20k data give d = (-0.5, -0.5, -0.5), r2 = 0.75:
    as a two-atom system (rc 8.5 A, 7 x 6 x 6 cells) the pair is found through the periodic
    image, once, with forces equal to the oracle's and opposite on the two atoms."""
    box = np.array([59.5, 51.0, 51.0], np.float32)
    p = np.zeros((2, 11), np.float32)
    p[0, :3] = [59.0, 50.5, 50.5]
    eps, sig, mass = [0.996], [1.0], [39.948]     # r = 0.87 A < 2^(1/6) sigma: repulsive
    g = mdrl.System(2, box, 8.5, eps, sig, mass, dt=2.0)
    g.set_types(np.zeros(2, np.int32))
    g.set_particles(p)
    got = g.forces_fixed()
    npairs = g.stats().pair_evals
    g.close()
    rl = O.RL(p[:, :3], box, [7, 6, 6], 8.5, np.zeros(2, np.int32), eps, sig, apoa1_tab6(apoa1))
    want, wpairs = rl.forces_fix()
    assert npairs == wpairs == 1
    assert np.array_equal(got, want)
    # repulsion pushes the atom at the upper faces away from the image across them
    assert np.all(got[0] == -got[1]) and got[0, 0] == got[0, 1] == got[0, 2] and got[0, 0] < 0


def test_eval_pairs_fast_mode_close(table):
    """Fused arithmetic differs from the golden records by rounding only."""
    recs = parse_reference_output()
    ref = hex2f(recs[:, 0:3].ravel()).reshape(-1, 3)
    nbr = hex2f(recs[:, 3:6].ravel()).reshape(-1, 3)
    fast = mdrl.eval_pairs(table, ref, nbr, 196.0, mdrl.PAIR_STRICT)
    want = hex2f(recs[:, 7:10].ravel()).reshape(-1, 3)
    assert np.linalg.norm(fast - want) / np.linalg.norm(want) < 1e-5


def test_eval_pairs_edge_cases(table):
    """r2 == 0 and r2 > rc2 give 0; r2 == rc2 is accepted unless strict; empty batch."""
    ref = np.array([[1, 1, 1], [0, 0, 0], [0, 0, 0], [0, 0, 0]], np.float32)
    nbr = np.array([[1, 1, 1], [12, 0, 0], [12.5, 0, 0], [3, 0, 0]], np.float32)
    out = mdrl.eval_pairs(table, ref, nbr, 144.0)
    assert np.all(out[0] == 0) and out[1, 0] != 0 and np.all(out[2] == 0) and out[3, 0] != 0
    out = mdrl.eval_pairs(table, ref, nbr, 144.0, mdrl.PAIR_STRICT)
    assert out[1, 0] == 0
    assert mdrl.eval_pairs(table, ref[:0], nbr[:0], 144.0).shape == (0, 3)


# ------------------------------------------------------------------------ C model
@pytest.mark.parametrize("case", ["uniform", "corner_cluster"])
def test_cmodel_gpu_matches_oracle(case):
    """The C model on the GPU (md_rl_cmodel_run) against its CPU restatement, 100 iterations."""
    say("cmodel:", case)
    text = pdb_text(cmodel_inputs()[case])
    p = np.zeros((500, 11), np.float32)
    p[:, :3] = parse_pdb_float32(text)
    params = O.cmodel_default_params()
    want, we, _ = O.cmodel_run(p, params)
    got, ge = mdrl.cmodel_run(p)
    assert np.array_equal(ge.view(np.uint32), we.view(np.uint32))
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


# ------------------------------------------------------------------------ half-shell forces
def gpu_system(s, flags=0):
    return mdrl.System.from_synthetic(s, flags=flags)


CASES = [(1, 0), (4, 0), (4, 3), (4, 11), (4, 40), (2, 0)]


@pytest.mark.parametrize("which,index", CASES)
@pytest.mark.parametrize("mode", ["fast", "parity", "analytic"])
def test_halfshell_int64_bit_exact(apoa1, which, index, mode):
    """int64 forces of the HIP half-shell kernel equal the oracle's bit for bit."""
    s = mdrl.SyntheticSystem(which, index)
    say("halfshell int64: workload %d replica %d n=%d mode=%s" % (which, index, s.n, mode))
    flags = {"fast": 0, "parity": mdrl.EVAL_PARITY, "analytic": mdrl.EVAL_ANALYTIC}[mode]
    omode = {"fast": 0, "parity": O.MODE_PARITY, "analytic": O.MODE_ANALYTIC}[mode]
    g = gpu_system(s, flags)
    got = g.forces_fixed()
    st = g.stats()
    rl = rl_from_synthetic(s, apoa1, mode=omode)
    want, npairs = rl.forces_fix()
    assert st.pair_evals == npairs
    bad = np.nonzero(np.any(got != want, axis=1))[0]
    assert bad.size == 0, "atoms differ: %s" % bad[:10]
    if mode != "fast":
        # the 640-atom neighbour tile of the other modes, with and without the task queue
        for sc in [(64, 640, 1, 1), (64, 640, 4, 0)]:
            g.set_schedule(*sc)
            g.force_eval()
            assert g.stats().pair_evals == npairs, sc
            bad = np.nonzero(np.any(g.forces_fixed() != want, axis=1))[0]
            assert bad.size == 0, "schedule %s: atoms differ: %s" % (sc, bad[:10])
    g.close()


def test_schedules_per_mode():
    """Which schedules each pair mode takes: the table mode every listed one, the analytic
    and parity modes and FP32 accumulation tiles 256 and 640 at 64 atoms per wave; the
    others raise and leave the schedule as it was."""
    s = mdrl.SyntheticSystem(4, 3)
    other = {(64, 256), (64, 640)}
    for flags, ok in [(0, {(a, t) for a in (16, 32, 64) for t in (128, 256)} | {(64, 640)}),
                      (mdrl.EVAL_ANALYTIC, other), (mdrl.EVAL_PARITY, other), (mdrl.ACC_FP32, other),
                      (mdrl.EVAL_ANALYTIC | mdrl.ACC_FP32, other), (mdrl.EVAL_PARITY | mdrl.ACC_FP32, other)]:
        g = gpu_system(s, flags)
        for a in (16, 32, 64):
            for t in (128, 256, 512, 640):
                before = g.schedule().as_dict()
                if (a, t) in ok:
                    g.set_schedule(a, t, 1, 1)
                    assert (g.schedule().atoms_per_wave, g.schedule().tile_depth) == (a, t)
                else:
                    with pytest.raises(mdrl.MdrlError):
                        g.set_schedule(a, t, 1, 1)
                    assert g.schedule().as_dict() == before, (flags, a, t)
        g.close()


@pytest.mark.parametrize("which,index", [(1, 0), (4, 3), (2, 0)])
def test_halfshell_fp32_atomics_close(which, index):
    s = mdrl.SyntheticSystem(which, index)
    say("halfshell fp32: workload %d replica %d" % (which, index))
    a = gpu_system(s).forces_fixed().astype(np.float64)
    g = gpu_system(s, mdrl.ACC_FP32)
    for tile in (256, 640):
        g.set_schedule(64, tile, 1, 1)
        g.force_eval()
        b = g.forces_fixed().astype(np.float64)
        assert np.linalg.norm(a - b) / np.linalg.norm(a) < 1e-5, tile
    g.close()


def test_small_boxes_all_images(apoa1):
    """Replicas with one or two cells per axis: every periodic image within rc counted once."""
    for index in range(64):
        s = mdrl.SyntheticSystem(4, index)
        if s.box[0] >= 30.0:
            continue
        g = gpu_system(s)
        st = g.stats()
        g.close()
        _, _, _, nb = O.fp64_direct(s.particle[:, :3], s.type, s.eps, s.sigma, s.box, s.rc,
                                    use_cells=False)
        assert st.pair_evals == nb, "replica %d (L=%.2f)" % (index, s.box[0])


@pytest.mark.parametrize("nt", [3, 5, 7])
@pytest.mark.parametrize("mode", ["fast", "parity", "analytic"])
def test_species_counts_not_powers_of_two(apoa1, nt, mode):
    """Type-pair multipliers for 3, 5 and 7 species (the queue entry adds type_i * ntypes
    and type_j, which only an OR would confuse for non-powers of two)."""
    s = mdrl.SyntheticSystem(4, 3)
    rng = np.random.default_rng(nt)
    types = rng.integers(0, nt, size=s.n).astype(np.int32)
    eps = np.linspace(0.2, 0.7, nt).astype(np.float32)
    sig = np.linspace(1.0, 3.5, nt).astype(np.float32)
    mass = np.full(nt, 12.011, np.float32)
    say("species: %d types, n=%d, mode %s" % (nt, s.n, mode))
    flags = {"fast": 0, "parity": mdrl.EVAL_PARITY, "analytic": mdrl.EVAL_ANALYTIC}[mode]
    omode = {"fast": 0, "parity": O.MODE_PARITY, "analytic": O.MODE_ANALYTIC}[mode]
    g = mdrl.System(s.n, s.box, s.rc, eps, sig, mass, dt=s.dt, coulomb_k=s.coulomb_k, flags=flags)
    g.set_types(types, s.charge)
    g.set_particles(s.particle)
    got = g.forces_fixed()
    npairs = g.stats().pair_evals
    g.close()
    ncell = [int(math.floor(float(s.box[d]) / s.rc)) for d in range(3)]
    rl = O.RL(s.particle[:, :3], s.box, ncell, s.rc, types, eps, sig, apoa1_tab6(apoa1),
              charge=s.charge, coulomb_k=s.coulomb_k, mode=omode)
    want, wpairs = rl.forces_fix()
    assert npairs == wpairs
    bad = np.nonzero(np.any(got != want, axis=1))[0]
    assert bad.size == 0, "atoms differ: %s" % bad[:10]


@pytest.mark.parametrize("mode", ["fast", "parity", "analytic"])
def test_anisotropic_box_forces(apoa1, mode):
    """A different cell count, cell length and periodic wrap on each axis (3 x 5 x 7 cells,
    sides not multiples of rc): int64 forces and the pair count equal the oracle's."""
    s = anisotropic_system()
    say("anisotropic box %s: n=%d mode=%s" % (s.box.tolist(), s.n, mode))
    flags = {"fast": 0, "parity": mdrl.EVAL_PARITY, "analytic": mdrl.EVAL_ANALYTIC}[mode]
    omode = {"fast": 0, "parity": O.MODE_PARITY, "analytic": O.MODE_ANALYTIC}[mode]
    g = gpu_system(s, flags)
    got = g.forces_fixed()
    npairs = g.stats().pair_evals
    g.close()
    rl = rl_from_synthetic(s, apoa1, omode)
    assert list(rl.s.ncell) == [3, 5, 7]
    want, wpairs = rl.forces_fix()
    assert npairs == wpairs
    bad = np.nonzero(np.any(got != want, axis=1))[0]
    assert bad.size == 0, "atoms differ: %s" % bad[:10]


def test_anisotropic_box_steps(apoa1):
    """Ten steps on the 3 x 5 x 7-cell box (atoms cross cells and the periodic faces of
    each axis): positions and forces bit-exact, step by step in the oracle."""
    s = anisotropic_system()
    say("anisotropic box steps: n=%d" % s.n)
    g = gpu_system(s)
    rl = rl_from_synthetic(s, apoa1)
    hk = O.half_kick(s.mass, s.dt)
    pos4 = rl.pos4.copy()
    vel = s.particle[:, 3:6].copy()
    fix, _ = rl.forces_fix()
    for k in range(10):
        pos4, vel, _ = O.motion_update(pos4, vel, fix, s.type, hk, s.mass, s.dt, s.box, k == 0)
        rl.pos4[:] = pos4
        fix, _ = rl.forces_fix()
    wrapped = np.abs(pos4[:, :3] - s.particle[:, :3]) > s.box / 2
    assert np.all(wrapped.sum(axis=0) > 0), "an axis without a periodic crossing"
    g.step(10)
    got = g.particles()
    gfix = g.forces_fixed()
    g.close()
    assert np.array_equal(got[:, :3].view(np.uint32), pos4[:, :3].view(np.uint32))
    assert np.array_equal(gfix, fix)


# ------------------------------------------------------------------------ steps
@pytest.mark.parametrize("which,index,nsteps", [(1, 0, 20), (4, 5, 10), (2, 0, 3)])
def test_steps_bit_exact(apoa1, which, index, nsteps):
    """Velocity-Verlet steps (update + rebin + forces) equal the oracle's step by step."""
    s = mdrl.SyntheticSystem(which, index)
    say("steps: workload %d replica %d n=%d steps=%d" % (which, index, s.n, nsteps))
    g = gpu_system(s)
    rl = rl_from_synthetic(s, apoa1)
    hk = O.half_kick(s.mass, s.dt)
    pos4 = rl.pos4.copy()
    vel = s.particle[:, 3:6].copy()
    fix, _ = rl.forces_fix()
    for k in range(nsteps):
        pos4, vel, _ = O.motion_update(pos4, vel, fix, s.type, hk, s.mass, s.dt, s.box, k == 0)
        rl.pos4[:] = pos4
        fix, _ = rl.forces_fix()
    g.step(nsteps)
    got = g.particles()
    gfix = g.forces_fixed()
    g.close()
    assert np.array_equal(got[:, :3].view(np.uint32), pos4[:, :3].view(np.uint32))
    assert np.array_equal(gfix, fix)
    # full-step velocity = vh + dv(F)
    dv = (fix.astype(np.float64) * hk[s.type][:, None]).astype(np.float32)
    vfull = (vel + dv).astype(np.float32)
    assert np.array_equal(got[:, 3:6].view(np.uint32), vfull.view(np.uint32))


def test_kick_drift_mode(apoa1):
    s = mdrl.SyntheticSystem(1)
    g = gpu_system(s, mdrl.KICK_DRIFT)
    rl = rl_from_synthetic(s, apoa1)
    hk = O.half_kick(s.mass, s.dt)
    pos4, vel = rl.pos4.copy(), s.particle[:, 3:6].copy()
    for k in range(5):
        fix, _ = rl.forces_fix()
        pos4, vel, _ = O.motion_update(pos4, vel, fix, s.type, hk, s.mass, s.dt, s.box, False,
                                       kick_drift=True)
        rl.pos4[:] = pos4
    g.step(5)
    got = g.particles()
    g.close()
    assert np.array_equal(got[:, :3].view(np.uint32), pos4[:, :3].view(np.uint32))


def test_deterministic_repeats():
    """int64 accumulation: two runs of the same steps give identical bits."""
    s = mdrl.SyntheticSystem(2)
    out = []
    for _ in range(2):
        g = gpu_system(s)
        g.step(5)
        out.append(g.particles())
        g.close()
    assert np.array_equal(out[0].view(np.uint32), out[1].view(np.uint32))


def test_energy_vs_fp64(apoa1):
    """FP64-reduced table energy vs the FP64 analytic energy; kinetic energy closed form."""
    s = mdrl.SyntheticSystem(1)
    g = gpu_system(s)
    pe, ke, vir = g.energy()
    g.close()
    _, e64, _, _ = O.fp64_direct(s.particle[:, :3], s.type, s.eps, s.sigma, s.box, s.rc)
    assert pe == pytest.approx(e64, rel=1e-4)
    v = s.particle[:, 3:6].astype(np.float64)
    ke_want = 0.5 * float(s.mass[0]) * float((v * v).sum()) * 1e4
    assert ke == pytest.approx(ke_want, rel=1e-12)


@pytest.mark.parametrize("which,index", [(1, 0), (4, 3)])
def test_virial_vs_fp64(which, index):
    """The pair virial sum_ij F_ij (x) r_ij from the force tables, FP64-reduced on the GPU,
    against the FP64 analytic oracle's (argon; a mixed-species replica with Coulomb). Its
    trace is sum_ij F_ij . r_ij; the table model carries up to 1.1e-4 relative error."""
    s = mdrl.SyntheticSystem(which, index)
    g = gpu_system(s)
    pe, _, vir = g.energy()
    g.close()
    q = s.charge if s.coulomb_k else None
    _, e64, vir64, _ = O.fp64_direct(s.particle[:, :3], s.type, s.eps, s.sigma, s.box, s.rc,
                                     charge=q, ke=s.coulomb_k)
    say("virial: workload %d replica %d trace %.6e (FP64 %.6e)" % (which, index, np.trace(vir),
                                                                 np.trace(vir64)))
    assert pe == pytest.approx(e64, rel=2e-4)
    assert np.linalg.norm(vir - vir64) / np.linalg.norm(vir64) < 2e-4
    assert np.trace(vir) == pytest.approx(np.trace(vir64), rel=2e-4)


def test_energy_terms_consistent():
    """md_rl_get_energy_terms against md_rl_energy on the same state (bit-identical kinetic,
    potential and virial), the virial-free pass's potential bit-identical to the full pass's,
    and the cutoff shift: one U(rc) = 4 eps ((sigma/rc)^12 - (sigma/rc)^6) per in-cutoff
    pair (argon, one species), within the tables' 1e-4."""
    s = mdrl.SyntheticSystem(1)
    g = gpu_system(s)
    g.step(10)
    pe, ke, vir = g.energy()
    pe2, ke2, none = g.energy(virial=False)
    t = g.energy_terms()
    npairs = g.stats().pair_evals
    g.close()
    assert none is None and pe2 == pe and ke2 == ke
    assert t["potential"] == pe and t["kinetic"] == ke and np.array_equal(t["virial"], vir)
    assert t["range_limited"] == pe and t["bonded"] == 0.0 and t["long_range"] == 0.0
    u_rc = 4 * 0.996 * ((3.4 / 8.5) ** 12 - (3.4 / 8.5) ** 6)
    say("cutoff shift %.6f over %d pairs: %.6e per pair (U(rc) %.6e)"
        % (t["cutoff_shift"], npairs, t["cutoff_shift"] / npairs, u_rc))
    assert t["cutoff_shift"] / npairs == pytest.approx(u_rc, rel=1e-4)


def test_nve_energy_drift():
    """Argon, 2,000 steps of 2 fs.  The reference potential is truncated without a shift
    (the *_no_smooth tables), so the potential jumps by U(rc) whenever a pair crosses the
    cutoff; the shifted energy PE - npairs * U(rc) removes the jumps and is conserved."""
    s = mdrl.SyntheticSystem(1)
    g = gpu_system(s)
    u_rc = 4 * 0.996 * ((3.4 / 8.5) ** 12 - (3.4 / 8.5) ** 6)
    pe0, ke0, _ = g.energy()
    n0 = g.stats().pair_evals
    g.step(2000)
    pe1, ke1, _ = g.energy()
    n1 = g.stats().pair_evals
    g.close()
    e0 = pe0 - n0 * u_rc + ke0
    e1 = pe1 - n1 * u_rc + ke1
    say("NVE: E0 %.4f E1 %.4f (shifted), raw %.4f -> %.4f" % (e0, e1, pe0 + ke0, pe1 + ke1))
    assert abs(e1 - e0) < 2e-3 * abs(pe0)


@pytest.mark.parametrize("which", [1, 2])
def test_minimize_bit_exact(apoa1, which):
    """Capped steepest descent on the GPU equals the oracle's iteration by iteration."""
    s = mdrl.SyntheticSystem(which)
    say("minimize: workload %d n=%d" % (which, s.n))
    g = gpu_system(s)
    rl = rl_from_synthetic(s, apoa1)
    fix, _ = rl.forces_fix()
    gamma, dmax = float(np.float32(1e-4)), float(np.float32(0.05))   # the ABI takes floats
    for _ in range(4):
        rl.pos4[:] = O.minimize_update(rl.pos4, fix, gamma, dmax, s.box)
        fix, _ = rl.forces_fix()
    g.minimize(4, 0.05, 1e-4)
    got = g.particles()
    gfix = g.forces_fixed()
    g.close()
    assert np.array_equal(got[:, :3].view(np.uint32), rl.pos4[:, :3].view(np.uint32))
    assert np.array_equal(gfix, fix)


def test_invalid_inputs_raise():
    s = mdrl.SyntheticSystem(1)
    with pytest.raises(mdrl.MdrlError):
        mdrl.System(s.n, s.box, 20.0, s.eps, s.sigma, s.mass)   # rc^2 = 400 beyond the table
    g = gpu_system(s)
    bad = s.particle.copy()
    bad[3, 0] = np.nan
    with pytest.raises(mdrl.MdrlError):
        g.set_particles(bad)
    g.close()


def ragged(index=20, width=2.5, seed=7):
    """A replica (4x4x4 cells) whose atoms with x < 0.3 L are moved into a dense blob:
    empty cells next to cells of ~1,600 atoms (split into several home and neighbour
    tiles) and ordinary cells of ~180."""
    s = mdrl.SyntheticSystem(4, index)
    L = np.asarray(s.box, np.float64)
    x = s.particle[:, :3].astype(np.float64)
    sel = x[:, 0] < 0.3 * L[0]
    rng = np.random.default_rng(seed)
    c = L * np.array([0.62, 0.55, 0.5])
    x[sel] = (c + width * rng.normal(size=(int(sel.sum()), 3))) % L
    s.particle[:, :3] = x.astype(np.float32)
    return s


@pytest.mark.parametrize("mode", ["fast", "parity", "analytic"])
def test_ragged_cells_bit_exact(apoa1, mode):
    s = ragged()
    say("ragged cells: n=%d box %.1f mode=%s" % (s.n, s.box[0], mode))
    flags = {"fast": 0, "parity": mdrl.EVAL_PARITY, "analytic": mdrl.EVAL_ANALYTIC}[mode]
    omode = {"fast": 0, "parity": O.MODE_PARITY, "analytic": O.MODE_ANALYTIC}[mode]
    g = gpu_system(s, flags)
    st = g.stats()
    assert st.max_cell_count > 512
    want, npairs = rl_from_synthetic(s, apoa1, mode=omode).forces_fix()
    scheds = [(64, 256, 1, 1)]
    if mode != "fast":
        scheds = [(64, t, gr, p) for t in (256, 640) for gr in (1, 4) for p in (0, 1)]
    if mode == "fast":
        scheds = [(a, t, gr, p) for a in (16, 32, 64) for t in (128, 256, 640) for gr in (1, 2, 4) for p in (0, 1)
                  if t != 640 or a == 64]
    for sc in scheds:
        g.set_schedule(*sc)
        g.force_eval()
        got = g.forces_fixed()
        assert g.stats().pair_evals == npairs, sc
        bad = np.nonzero(np.any(got != want, axis=1))[0]
        assert bad.size == 0, "schedule %s: atoms differ: %s" % (sc, bad[:10])
    g.close()


def test_cell_over_sort_capacity(apoa1):
    """7,951 atoms held as one cell (a 43 A box with an explicit 1 x 1 x 1 grid, above the
    2,048 atoms the in-cell sort holds in LDS): the gather sorts the cell in global memory
    and raises MDRL_F_CELL_OVERFLOW, and the int64 forces equal the oracle's before and
    after two steps (a lost or repeated atom in the gather would change them)."""
    s = mdrl.SyntheticSystem(4, 24)
    say("one-cell system: n=%d box %.1f" % (s.n, s.box[0]))
    g = mdrl.System(s.n, s.box, s.rc, s.eps, s.sigma, s.mass, dt=s.dt, coulomb_k=s.coulomb_k,
                    ncell=(1, 1, 1))
    g.set_types(s.type, s.charge)
    g.set_particles(s.particle)
    got = g.forces_fixed()
    st = g.stats()
    assert st.max_cell_count == s.n > 2048
    assert st.flags & mdrl.F_CELL_OVERFLOW
    tab = apoa1_tab6(apoa1)
    rl = O.RL(s.particle[:, :3], s.box, (1, 1, 1), s.rc, s.type, s.eps, s.sigma, tab,
              charge=s.charge, coulomb_k=s.coulomb_k)
    want, npairs = rl.forces_fix()
    assert st.pair_evals == npairs
    assert np.array_equal(got, want)
    g.step(2)
    p = g.particles()
    rl = O.RL(p[:, :3], s.box, (1, 1, 1), s.rc, s.type, s.eps, s.sigma, tab,
              charge=s.charge, coulomb_k=s.coulomb_k)
    want, npairs = rl.forces_fix()
    assert g.stats().pair_evals == npairs
    assert np.array_equal(g.forces_fixed(), want)
    g.close()


@pytest.mark.parametrize("mode", ["fast", "analytic"])
def test_saturated_pair_force(apoa1, mode):
    """Two atoms 0.2 A apart: F * dx * 2^32 exceeds 2^62, so the x components are clamped to
    -+2^62 exactly as the oracle clamps them, y and z stay 0, and MDRL_F_SATURATED is
    raised."""
    s0 = mdrl.SyntheticSystem(1)
    box = np.array([30.0, 30.0, 30.0], np.float32)
    part = np.zeros((2, 11), np.float32)
    part[0, :3] = [10.0, 10.0, 10.0]
    part[1, :3] = [10.2, 10.0, 10.0]
    types = np.zeros(2, np.int32)
    flags = 0 if mode == "fast" else mdrl.EVAL_ANALYTIC
    g = mdrl.System(2, box, s0.rc, s0.eps, s0.sigma, s0.mass, dt=s0.dt, flags=flags)
    g.set_types(types)
    g.set_particles(part)
    got = g.forces_fixed()
    st = g.stats()
    g.close()
    rl = O.RL(part[:, :3], box, (3, 3, 3), s0.rc, types, s0.eps, s0.sigma, apoa1_tab6(apoa1),
              mode=0 if mode == "fast" else O.MODE_ANALYTIC)
    want, npairs = rl.forces_fix()
    assert npairs == 1 and st.pair_evals == 1
    assert np.array_equal(got, want)
    assert abs(int(got[0, 0])) == 1 << 62 and got[0, 0] == -got[1, 0]
    assert np.all(got[:, 1:] == 0)
    assert st.flags & mdrl.F_SATURATED


def test_non_finite_states_reported(apoa1):
    """A NaN position is refused at input, and a velocity that carries an atom to infinity
    makes the step report MDRL_ENONFINITE (-6) instead of returning silently."""
    s = mdrl.SyntheticSystem(1)
    g = gpu_system(s)
    bad = s.particle.copy()
    bad[5, 1] = np.nan
    with pytest.raises(mdrl.MdrlError) as e:
        g.set_particles(bad)
    assert e.value.code == -6
    fast = s.particle.copy()
    fast[7, 3] = 3e38
    g.set_particles(fast)
    with pytest.raises(mdrl.MdrlError) as e:
        g.step(1)
    assert e.value.code == -6
    assert g.stats().flags & mdrl.F_NONFINITE
    g.close()


def test_analytic_cutoff_decided_exactly(apoa1):
    """400 two-atom systems within 2e-7 of rc across a periodic face, as one analytic-mode
    batch: each pair is kept iff its exact FP64 r^2 <= rc^2, so the per-system int64 forces
    equal the oracle's analytic mode and the pair count is the exact one.  The FP32 filter
    of the table path would decide some of them the other way."""
    base = mdrl.SyntheticSystem(4, 3)
    L, rc = 60.0, 10.0
    cases = cutoff_edge_pairs(400, L=L, rc=rc)
    t = np.array([0, 1], np.int32)
    q = np.array([0.3, -0.3], np.float32)
    b = mdrl.Batch([2] * len(cases), [[L] * 3] * len(cases), rc, base.eps, base.sigma, base.mass,
                   coulomb_k=base.coulomb_k, flags=mdrl.EVAL_ANALYTIC)
    for k, (x, _) in enumerate(cases):
        p = np.zeros((2, 11), np.float32)
        p[:, :3] = x
        b.set_system(k, p, t, q)
    b.commit()
    inside = 0
    for k, (x, exact) in enumerate(cases):
        want, n = O.RL(x, np.full(3, L, np.float32), [int(L // rc)] * 3, rc, t, base.eps,
                       base.sigma, apoa1_tab6(apoa1), charge=q, coulomb_k=base.coulomb_k,
                       mode=O.MODE_ANALYTIC).forces_fix()
        assert n == int(exact)
        assert np.array_equal(b.system_forces_fixed(k), want), "system %d" % k
        inside += n
    say("edge pairs: %d of %d within rc exactly" % (inside, len(cases)))
    assert b.stats().pair_evals == inside
    b.close()


def test_analytic_separation_exact_across_the_periodic_face(apoa1):
    """Pairs across a periodic face (where the staged image fl(x_j + L) rounds) and the same
    exact separations inside the box, as one analytic-mode batch: every system's int64
    forces equal the oracle's, and each face pair's forces equal its interior copy's."""
    base = mdrl.SyntheticSystem(4, 3)
    L, rc = 60.0, 12.0
    cases = face_and_interior_pairs(100, L=L)
    systems = [x for pair in cases for x in pair]
    t = np.array([1, 2], np.int32)
    q = np.array([0.3, -0.3], np.float32)
    b = mdrl.Batch([2] * len(systems), [[L] * 3] * len(systems), rc, base.eps, base.sigma,
                   base.mass, coulomb_k=base.coulomb_k, flags=mdrl.EVAL_ANALYTIC)
    for k, x in enumerate(systems):
        p = np.zeros((2, 11), np.float32)
        p[:, :3] = x
        b.set_system(k, p, t, q)
    b.commit()
    got = [b.system_forces_fixed(k) for k in range(len(systems))]
    assert b.stats().pair_evals == len(systems)
    b.close()
    for k, x in enumerate(systems):
        want, n = O.RL(x, np.full(3, L, np.float32), [5] * 3, rc, t, base.eps, base.sigma,
                       apoa1_tab6(apoa1), charge=q, coulomb_k=base.coulomb_k,
                       mode=O.MODE_ANALYTIC).forces_fix()
        assert n == 1 and np.array_equal(got[k], want), "system %d" % k
    for k in range(len(cases)):
        assert np.array_equal(got[2 * k], got[2 * k + 1]), "pair %d: face != interior" % k
