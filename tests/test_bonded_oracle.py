"""CPU: the bonded-term and exclusion parts of the oracle, and the chain workload generator.

Pins: the FP64 bonded potentials against central finite differences of their own energy
and against closed forms (two-atom bond, an angle at its rest value, torsion extrema);
the FP32 fixed-point restatement against the FP64 one; arctan against libm; exclusion
windows against brute-force pair counts and an independent FP64 evaluation.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from sc2019_md_amd import mdrl

from tests.test_oracle_golden import apoa1_tab6  # noqa: E402

BOX = np.array([30.0, 30.0, 30.0])


def chain(n, rng, noise=0.08):
    x = [np.array([5.0, 6.0, 7.0])]
    d = rng.normal(size=3)
    x.append(x[0] + 1.53 * d / np.linalg.norm(d))
    for _ in range(2, n):
        b = x[-1] - x[-2]
        b /= np.linalg.norm(b)
        v = rng.normal(size=3)
        perp = v - b * (v @ b)
        perp /= np.linalg.norm(perp)
        th = math.radians(111.0 + 6.0 * rng.normal())
        x.append(x[-1] + 1.53 * (-b * math.cos(th) + perp * math.sin(th)))
    x = np.array(x)
    return x + noise * rng.normal(size=x.shape)


def chain_terms(n, k_tors=5.9, phase=0.0, nper=3):
    atoms, params = [], []
    for k in range(n - 1):
        atoms.append([k, k + 1, -1, -1]); params.append([1.53, 1100.0, 0, 0])
    for k in range(n - 2):
        atoms.append([k, k + 1, k + 2, -1]); params.append([math.radians(111.0), 260.0, 0, 0])
    for k in range(n - 3):
        atoms.append([k, k + 1, k + 2, k + 3]); params.append([k_tors, nper, phase, 0])
    return np.array(atoms, np.int32), np.array(params, np.float64)


def test_atan2f_against_libm():
    rng = np.random.default_rng(3)
    y = np.abs(rng.normal(size=4000)).astype(np.float32)
    x = rng.normal(size=4000).astype(np.float32)
    got = np.array([O.atan2f(a, b) for a, b in zip(y, x)])
    want = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    assert np.max(np.abs(got - want)) < 5e-7
    assert O.atan2f(0.0, -1.0) == pytest.approx(math.pi, abs=1e-6)
    assert O.atan2f(1.0, 0.0) == pytest.approx(math.pi / 2, abs=1e-6)
    assert O.atan2f(0.0, 0.0) == 0.0


@pytest.mark.parametrize("phase,nper", [(0.0, 3), (0.7, 2), (-1.3, 1)])
def test_bonded_fp64_is_the_gradient(phase, nper):
    rng = np.random.default_rng(11)
    x = chain(7, rng)
    atoms, params = chain_terms(7, phase=phase, nper=nper)
    f, _ = O.bonded_fp64(x, BOX, atoms, params)
    h = 1e-6
    fd = np.zeros_like(x)
    for i in range(x.shape[0]):
        for d in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, d] += h
            xm[i, d] -= h
            fd[i, d] = -(O.bonded_fp64(xp, BOX, atoms, params)[1] -
                         O.bonded_fp64(xm, BOX, atoms, params)[1]) / (2 * h)
    assert np.max(np.abs(f - fd)) < 1e-5 * np.max(np.abs(f))


def test_bonded_closed_forms():
    # bond: U = k (r - r0)^2, F1 = -2 k (r - r0) r12 / r
    x = np.array([[1.0, 2.0, 3.0], [1.0, 2.0, 4.8]])
    f, e = O.bonded_fp64(x, BOX, [[0, 1, -1, -1]], [[1.53, 1100.0, 0, 0]])
    assert e == pytest.approx(1100.0 * 0.27 ** 2)
    assert f[0] == pytest.approx([0.0, 0.0, 2 * 1100.0 * 0.27])
    assert np.allclose(f[0] + f[1], 0.0)
    # angle at its rest value: no force; at 90 degrees: U = k (pi/2 - theta0)^2
    th0 = math.radians(111.0)
    x = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 0.0], [math.cos(th0), math.sin(th0), 0.0]]) * 1.53 + 5.0
    f, e = O.bonded_fp64(x, BOX, [[0, 1, 2, -1]], [[th0, 260.0, 0, 0]])
    assert e == pytest.approx(0.0, abs=1e-18) and np.max(np.abs(f)) < 1e-9
    x = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, 1.0, 0.0]]) * 1.53 + 5.0
    f, e = O.bonded_fp64(x, BOX, [[0, 1, 2, -1]], [[th0, 260.0, 0, 0]])
    assert e == pytest.approx(260.0 * (math.pi / 2 - th0) ** 2)
    assert np.allclose(f.sum(axis=0), 0.0, atol=1e-9)
    # torsion about the x axis: trans (phi = 180) is a minimum of k (1 + cos 3 phi), cis
    # (phi = 0) the maximum 2 k; both extrema carry no force; phi = 60 is a minimum again
    for phi, u in [(math.pi, 0.0), (0.0, 2 * 5.9), (math.pi / 3, 0.0), (math.pi / 2, 5.9)]:
        x = np.array([[0.0, 1.0, 0.0], [0.0, 0.0, 0.0], [1.53, 0.0, 0.0],
                      [1.53, math.cos(phi), math.sin(phi)]]) + 5.0
        f, e = O.bonded_fp64(x, BOX, [[0, 1, 2, 3]], [[5.9, 3, 0.0, 0]])
        assert e == pytest.approx(u, abs=1e-9)
        if u in (0.0, 2 * 5.9):
            assert np.max(np.abs(f)) < 1e-8
        else:
            assert np.max(np.abs(f)) > 1.0 and np.allclose(f.sum(axis=0), 0.0, atol=1e-9)


def test_bonded_fix_matches_fp64():
    rng = np.random.default_rng(5)
    x = chain(20, rng)
    atoms, params = chain_terms(20)
    f64, e64 = O.bonded_fp64(x, BOX, atoms, params)
    pos4 = np.zeros((20, 4), np.float32)
    pos4[:, :3] = x
    fix, te = O.bonded_fix(pos4, BOX, atoms, params)
    f = fix.astype(np.float64) / 2 ** 32 * 0.1
    # the FP32 arithmetic rounds; the FP64 reference sees the same FP32 positions
    f64b, e64b = O.bonded_fp64(pos4[:, :3], BOX, atoms, params)
    assert np.linalg.norm(f - f64b) / np.linalg.norm(f64b) < 2e-5
    assert te.sum() == pytest.approx(e64b, rel=1e-5)
    assert np.all(fix.sum(axis=0) == 0)    # each term's fixed-point forces sum to zero


def test_bonded_minimum_image():
    """A bond across the periodic boundary gives the force of the same bond inside the box."""
    a = np.array([[29.5, 10.0, 10.0, 0.0], [0.8, 10.4, 10.0, 0.0]], np.float32)   # across x = 0
    b = np.array([[19.5, 10.0, 10.0, 0.0], [20.8, 10.4, 10.0, 0.0]], np.float32)  # inside
    fa, _ = O.bonded_fix(a, BOX, [[0, 1, -1, -1]], [[1.53, 1100, 0, 0]])
    fb, _ = O.bonded_fix(b, BOX, [[0, 1, -1, -1]], [[1.53, 1100, 0, 0]])
    assert np.max(np.abs(fa - fb)) <= 2 ** 20 * 1e-3 * np.max(np.abs(fb))   # FP32 rounding of x
    assert np.sign(fa[0, 0]) == np.sign(fb[0, 0]) != 0


@pytest.fixture(scope="module")
def apoa1():
    t = mdrl.Table.load()
    return {k: t.get(k) for k in range(6)}


@pytest.fixture(scope="module")
def chains():
    return mdrl.SyntheticSystem(5, 0)


def test_chain_generator(chains):
    s = chains
    assert s.n == 3072 and s.nterms == 64 * (47 + 46 + 45)
    x = s.particle[:, :3].astype(np.float64)
    L = float(s.box[0])
    assert L == pytest.approx((3072 / 0.033) ** (1 / 3), rel=1e-6)
    bonds = s.term_atoms[s.term_atoms[:, 2] < 0]
    d = x[bonds[:, 0]] - x[bonds[:, 1]]
    d -= L * np.round(d / L)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.53, atol=2e-5)
    ang = s.term_atoms[(s.term_atoms[:, 2] >= 0) & (s.term_atoms[:, 3] < 0)]
    u = x[ang[:, 0]] - x[ang[:, 1]]
    v = x[ang[:, 2]] - x[ang[:, 1]]
    u -= L * np.round(u / L)
    v -= L * np.round(v / L)
    th = np.degrees(np.arccos(np.sum(u * v, 1) / np.linalg.norm(u, axis=1) / np.linalg.norm(v, axis=1)))
    assert np.allclose(th, 111.0, atol=1e-3)
    # exclusion windows: symmetric, the 1-2, 1-3 and 1-4 neighbours of each chain
    m = s.excl
    for i in range(s.n):
        for dbit in range(32):
            if (int(m[i]) >> dbit) & 1:
                j = i + dbit - 16
                assert (int(m[j]) >> (32 - dbit)) & 1 and i // 48 == j // 48 and 1 <= abs(i - j) <= 3
    assert bin(int(m[0])).count("1") == 3 and bin(int(m[10])).count("1") == 6
    # no non-excluded pair closer than 2.3 A; alternating charges, neutral chains
    from scipy.spatial import cKDTree
    pairs = cKDTree(x, boxsize=L).query_pairs(2.3, output_type="ndarray")
    close = [(i, j) for i, j in pairs if not (i // 48 == j // 48 and abs(i - j) <= 3)]
    assert not close
    assert np.all(s.charge[0::2] == np.float32(0.25)) and np.all(s.charge[1::2] == np.float32(-0.25))
    again = mdrl.SyntheticSystem(5, 0)
    assert np.array_equal(again.particle, s.particle)


def test_exclusions_in_the_oracle(chains, apoa1):
    """Pair count with exclusion windows = brute-force count of non-excluded pairs; forces
    agree with an independent FP64 evaluation that skips the same pairs."""
    s = chains
    ncell = [int(math.floor(float(s.box[d]) / s.rc)) for d in range(3)]
    rl = O.RL(s.particle[:, :3], s.box, ncell, s.rc, s.type, s.eps, s.sigma, apoa1_tab6(apoa1),
              charge=s.charge, coulomb_k=s.coulomb_k, excl=s.excl)
    fix, npairs = rl.forces_fix()
    rl_all = O.RL(s.particle[:, :3], s.box, ncell, s.rc, s.type, s.eps, s.sigma, apoa1_tab6(apoa1),
                  charge=s.charge, coulomb_k=s.coulomb_k)
    _, nall = rl_all.forces_fix()
    f64, _, _, n64 = O.fp64_direct(s.particle[:, :3], s.type, s.eps, s.sigma, s.box, s.rc,
                                   charge=s.charge, ke=s.coulomb_k, excl=s.excl)
    x = s.particle[:, :3].astype(np.float64)
    L = float(s.box[0])
    nex = 0
    for i in range(s.n):
        for j in range(i + 1, min(s.n, i + 4)):
            if i // 48 == j // 48:
                d = x[i] - x[j]
                d -= L * np.round(d / L)
                nex += int(d @ d <= s.rc ** 2)
    assert npairs == nall - nex
    assert abs(n64 - npairs) <= 2          # FP64 vs FP32 r2 at the cutoff
    f = fix.astype(np.float64) / 2 ** 32 * 0.1
    assert np.linalg.norm(f - f64) / np.linalg.norm(f64) < 2e-4
    assert np.all(fix.sum(axis=0) == 0)
