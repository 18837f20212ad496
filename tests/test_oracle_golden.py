"""The oracle pinned against the reference's own golden vectors and known answers (CPU).

Fixtures (tests/golden/PROVENANCE.json lists source paths and SHA-256):
  VERIFICATION_REFERENCE_OUTPUT.txt.gz   10,000 pair records from
      MatlabScripts/Simulation_Verification_Raw_ApoA1_LJ_no_smooth.m:113-200
  VERIFICATION_PARTICLE_PAIR_INPUT.sha256  the half-shell pair stream of
      MatlabScripts/Simulation_Verification_LJArgon.m:518-572
  input_positions_ljargon_20000_box_58_49_49.txt.gz   its input
"""
import gzip
import hashlib
import math
import os
import tempfile

import numpy as np
import pytest

from oracle import oracle as O
from sc2019_md_amd import mdrl

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TABLES = os.path.join(os.path.dirname(GOLD), "..", "sc2019_md_amd", "tables", "apoa1")


def hex2f(h):
    return np.array([int(x, 16) for x in h], dtype=np.uint32).view(np.float32)


@pytest.fixture(scope="module")
def apoa1():
    t = mdrl.Table.load()
    return {k: t.get(k) for k in range(6)}


def parse_reference_output():
    with gzip.open(os.path.join(GOLD, "VERIFICATION_REFERENCE_OUTPUT.txt.gz"), "rt") as f:
        lines = f.read().splitlines()[2:]
    recs = []
    for a, b in zip(lines[0::2], lines[1::2]):
        head, forces = a.split(": ")
        pa, pb = head[1:-1].split("),(")
        dist = b.split(": ")[1]
        recs.append(pa.split(",") + pb.split(",") + forces.split("\t") + dist.split(", "))
    return np.array(recs)


def test_reference_output_bit_exact(apoa1):
    """All 10,000 records: F, Fx, Fy, Fz, r2, dx, dy, dz bit-identical."""
    recs = parse_reference_output()
    assert recs.shape == (10000, 14)
    ref = hex2f(recs[:, 0:3].ravel()).reshape(-1, 3)
    nbr = hex2f(recs[:, 3:6].ravel()).reshape(-1, 3)
    want = recs[:, 6:14]
    c0_14, c1_14 = apoa1[mdrl.TERM_14]
    c0_8, c1_8 = apoa1[mdrl.TERM_8]
    out = O.pair_golden(c0_14, c1_14, c0_8, c1_8, ref, nbr, 196.0, strict=True)
    got = np.array([["%08X" % v for v in row] for row in out.view(np.uint32)])
    mism = np.nonzero(np.any(got != want, axis=1))[0]
    assert mism.size == 0, "records differ: %s" % mism[:10]
    # 9,894 of the records lie inside the cutoff
    assert int(np.count_nonzero(out[:, 0])) == 9894


def test_reference_output_positions_are_apoa1_rows():
    recs = parse_reference_output()
    rows = np.loadtxt(os.path.join(GOLD, "input_positions_ApoA1_first100.txt"))
    ref = hex2f(recs[:, 0:3].ravel()).reshape(100, 100, 3)
    assert np.array_equal(ref[:, 0, :], rows.astype(np.float32))


def test_fused_rounding_breaks_parity(apoa1):
    """Planted mistake on the CPU: with fused c1*r2 + c0 most records no longer match, so the
    bit-exact test above really pins the rounding order."""
    recs = parse_reference_output()
    ref = hex2f(recs[:, 0:3].ravel()).reshape(-1, 3)
    nbr = hex2f(recs[:, 3:6].ravel()).reshape(-1, 3)
    c0_14, c1_14 = apoa1[mdrl.TERM_14]
    c0_8, c1_8 = apoa1[mdrl.TERM_8]
    out = O.pair_golden(c0_14, c1_14, c0_8, c1_8, ref, nbr, 196.0)
    r2 = out[:, 4].astype(np.float64)
    idx = np.array([O.lut_index(x) for x in out[:, 4]])
    t14 = (c1_14[idx].astype(np.float64) * r2 + c0_14[idx]).astype(np.float32)
    t8 = (c1_8[idx].astype(np.float64) * r2 + c0_8[idx]).astype(np.float32)
    fused = (t14 - t8) * (out[:, 0] != 0)
    assert np.count_nonzero(fused.view(np.uint32) != out[:, 0].view(np.uint32)) > 1000


def test_pair_stream_reproduces_reference_file():
    """Binning (1-based ceil, stable order) + 8-filter half-shell stream == the 3.8 MB file."""
    with gzip.open(os.path.join(GOLD, "input_positions_ljargon_20000_box_58_49_49.txt.gz"), "rt") as f:
        raw = np.array([[float(x) for x in l.split()[1:4]] for l in f if l.strip()])
    assert raw.shape == (20000, 3)
    text = O.pair_stream_text(raw, 8.5, (7, 6, 6), (2, 2, 2), 100)
    want_sha, want_lines = open(os.path.join(GOLD, "VERIFICATION_PARTICLE_PAIR_INPUT.sha256")).read().split()
    got_lines = text.split(b"\n")
    assert len(got_lines) == int(want_lines)
    with gzip.open(os.path.join(GOLD, "VERIFICATION_PARTICLE_PAIR_INPUT.head.txt.gz"), "rb") as f:
        head = f.read().split(b"\n")
    for k, (g, w) in enumerate(zip(got_lines[:720], head[:720])):
        assert g == w, "line %d differs" % k
    assert hashlib.sha256(text).hexdigest() == want_sha


def test_table_formats_agree():
    """Intel-HEX (OpenCL_HDL_Lib/HDL_Source/c0_14.hex) and .mif copies hold the same words."""
    a = mdrl.read_table_file(os.path.join(GOLD, "c0_14.hex"))
    b = mdrl.read_table_file(os.path.join(mdrl.default_table_dir(), "c0_14.mif"))
    assert a.shape == (3584,) and np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("term,k,scale,tol", [
    (mdrl.TERM_14, 14, 48.0, 1.1e-4), (mdrl.TERM_8, 8, 24.0, 4e-5),
    (mdrl.TERM_3, 3, 1.0 / (4 * 3.14), 1e-5), (mdrl.TERM_12, 12, 4.0, 1e-4),
    (mdrl.TERM_6, 6, 4.0, 3e-5), (mdrl.TERM_1, 1, 1.0 / (4 * 3.14), 5e-6)])
def test_table_matches_closed_form(apoa1, term, k, scale, tol):
    """Indexing by {exp - 121, mantissa[22:15]} with t = c1 r2 + c0 reproduces scale * r^-k
    over r2 in [2^-6, 2^8): pins the offset, the bin bits and the c0/c1 roles."""
    rng = np.random.default_rng(1)
    r2 = np.float32(2.0) ** rng.uniform(-6, 8, 20000).astype(np.float32)
    c0, c1 = apoa1[term]
    idx = np.array([O.lut_index(x) for x in r2])
    t = c1[idx].astype(np.float64) * r2 + c0[idx]
    exact = scale * r2.astype(np.float64) ** (-k / 2.0)
    rel = np.abs(t / exact - 1.0)
    assert rel.max() < tol


def test_lut_index_out_of_range_is_segment_zero():
    # r2 below 2^-6 and above 2^8 fall back to segment 0 with the mantissa bin (HDL rule)
    for r2 in (2.0 ** -7 * 1.5, 300.0):
        u = np.array([r2], np.float32).view(np.uint32)[0]
        assert O.lut_index(r2) == (u >> 15) & 0xFF
    assert O.lut_index(1.0) == 6 * 256
    assert O.lut_index(144.0) == 13 * 256 + ((np.array([144.0], np.float32).view(np.uint32)[0] >> 15) & 0xFF)


# ------------------------------------------------------------------------------ C model
def pdb_text(xyz):
    lines = ["REMARK    generated for the C-model parity check"] * 5
    for i, (x, y, z) in enumerate(xyz):
        lines.append("ATOM  %5d  AR   AR A   1    %8.3f%8.3f%8.3f  1.00  0.00" % (i + 1, x, y, z))
    return "\n".join(lines) + "\n"


def parse_pdb_float32(text):
    """The reference's sscanf("%f") reading, via libc strtof."""
    import ctypes
    libc = ctypes.CDLL(None)
    libc.strtof.restype = ctypes.c_float
    libc.strtof.argtypes = [ctypes.c_char_p, ctypes.c_void_p]
    out = []
    for l in text.splitlines()[5:]:
        tok = l.split()
        out.append([libc.strtof(t.encode(), None) for t in tok[6:9]])
    return np.array(out, np.float32)


def cmodel_inputs():
    rng = np.random.default_rng(7)
    uniform = rng.uniform(0, 24, (500, 3))
    g = np.arange(8) * 0.40
    cl = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)[:500] - 1.4
    cluster = np.mod(cl + rng.uniform(-0.02, 0.02, cl.shape), 24.0)
    return {"uniform": uniform, "corner_cluster": cluster}


@pytest.mark.skipif(not os.path.exists(O.REF_BIN), reason="oracle/_ref not built (needs the reference tree)")
@pytest.mark.parametrize("case", ["uniform", "corner_cluster"])
def test_cmodel_matches_reference_binary(case):
    """The restated C model prints exactly what the reference C model, compiled from
    Range_Limited/C_Model_ForceEval/MD_Evaluation_Model.cpp, prints for the same PDB."""
    text = pdb_text(cmodel_inputs()[case])
    with tempfile.TemporaryDirectory() as d:
        want = O.run_reference_cmodel(text, d)
    xyz = parse_pdb_float32(text)
    p = np.zeros((500, 11), np.float32)
    p[:, :3] = xyz
    params = O.cmodel_default_params()
    _, _, lines = O.cmodel_run(p, params, with_lines=True)
    assert lines == want
    assert len(want.splitlines()) == 100


def test_cmodel_two_atom_closed_form():
    """Two atoms on the x axis: F = 48 eps s^12 r^-13 - 24 eps s^6 r^-7 (C-model units)."""
    params = O.cmodel_default_params()
    params.iterations = 1
    p = np.zeros((2, 11), np.float32)
    p[0, :3] = [10.0, 12.0, 12.0]
    p[1, :3] = [10.45, 12.0, 12.0]
    _, e, _ = O.cmodel_run(p, params)
    cnt, f = O.cmodel_force_sample(p, params, 2)
    assert cnt == 2
    r = float(np.float32(10.45) - np.float32(10.0))
    eps, s = float(params.eps), float(params.sigma)
    fx = -(48 * eps * s ** 12 / r ** 13 - 24 * eps * s ** 6 / r ** 7)
    assert f[0, 0] == pytest.approx(fx, rel=1e-5)
    assert f[1, 0] == pytest.approx(-fx, rel=1e-5)
    en = 4 * eps * ((s / r) ** 12 - (s / r) ** 6)
    assert e[0, 0] == pytest.approx(en, rel=1e-5)


# ------------------------------------------------------------------------------ half-shell model
def apoa1_tab6(apoa1):
    return O.interleave([*apoa1[mdrl.TERM_14], *apoa1[mdrl.TERM_8], *apoa1[mdrl.TERM_3]])


def rl_from_synthetic(s, apoa1, mode=0):
    ncell = [max(1, int(math.floor(float(s.box[d]) / s.rc))) for d in range(3)]
    return O.RL(s.particle[:, :3], s.box, ncell, s.rc, s.type, s.eps, s.sigma, apoa1_tab6(apoa1),
                charge=s.charge if s.coulomb_k else None, coulomb_k=s.coulomb_k, mode=mode)


def anisotropic_system():
    """The mixed-species parameters of a 5,710-atom replica on a 34 x 52 x 79 A box with
    rc = 10 A: 3, 5 and 7 cells per axis, each longer than rc by a different amount.
    The atoms sit on a 2.6 A lattice jittered by up to 0.2 A, with seeded types and
    charges."""
    s = mdrl.SyntheticSystem(4, 3)
    s.box = np.array([34.0, 52.0, 79.0], np.float32)
    rng = np.random.default_rng(357)
    axes = [np.arange(int(s.box[d] // 2.6)) * 2.6 + 0.5 for d in range(3)]
    grid = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 3)
    s.n = grid.shape[0]
    s.particle = np.zeros((s.n, 11), np.float32)
    s.particle[:, :3] = grid + rng.uniform(-0.2, 0.2, grid.shape)
    s.particle[:, 3:6] = rng.normal(0.0, 0.04, (s.n, 3))   # A/fs
    s.type = rng.integers(0, s.ntypes, s.n).astype(np.int32)
    s.charge = rng.normal(0.0, 0.4, s.n).astype(np.float32)
    return s


def cutoff_edge_pairs(count, seed=11, L=60.0, rc=10.0):
    """Two-atom systems whose separation lies within 2e-7 relative of rc, across a periodic
    face (the staged image shift rounds there): positions as float32 rows in [0, L) and
    whether each pair lies within rc exactly (FP64 of the float32 positions, minimum
    image)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        a = k % 3
        xi = rng.uniform(0.3 * L, 0.7 * L, 3)
        xi[a] = L - rng.uniform(0.0, 0.5 * rc)
        u = rng.normal(size=3)
        u[a] = abs(u[a]) + 1.0
        u /= np.linalg.norm(u)
        xj = xi + rc * (1.0 + rng.uniform(-2e-7, 2e-7)) * u
        x = np.stack([xi, xj]).astype(np.float32) % np.float32(L)
        x = np.where(x >= np.float32(L), np.float32(0.0), x).astype(np.float32)
        d = x[0].astype(np.float64) - x[1].astype(np.float64)
        d -= L * np.round(d / L)
        out.append((x, bool((d * d).sum() <= np.float64(np.float32(rc)) ** 2)))
    return out


def test_analytic_mode_decides_the_cutoff_exactly(apoa1):
    """The analytic mode keeps a pair iff its exact FP64 r^2 <= rc^2, as the FP64 direct
    evaluation does; the table path keeps the reference's FP32 test r2 <= rc2
    (D/Filter_Logic.v:234), which differs on some of these pairs across the periodic face."""
    base = mdrl.SyntheticSystem(4, 3)
    L, rc = 60.0, 10.0
    flips = 0
    for x, inside in cutoff_edge_pairs(400, L=L, rc=rc):
        t = np.array([0, 1], np.int32)
        q = np.array([0.3, -0.3], np.float32)
        box = np.full(3, L, np.float32)
        ncell = [int(L // rc)] * 3
        kw = dict(charge=q, coulomb_k=base.coulomb_k)
        tab = apoa1_tab6(apoa1)
        _, n_ana = O.RL(x, box, ncell, rc, t, base.eps, base.sigma, tab, mode=O.MODE_ANALYTIC,
                        **kw).forces_fix()
        _, n_tab = O.RL(x, box, ncell, rc, t, base.eps, base.sigma, tab, **kw).forces_fix()
        _, _, _, n64 = O.fp64_direct(x, t, base.eps, base.sigma, box, rc, charge=q,
                                     ke=base.coulomb_k, use_cells=False)
        assert n_ana == int(inside) == n64
        flips += n_tab != n_ana
    assert flips >= 8, "the sample should hold pairs the FP32 test decides the other way"


def face_and_interior_pairs(count, seed=5, L=60.0):
    """Pairs of two-atom systems (box L = 60 A, rc 12 A) with the same exact separation D:
    one across the periodic face of an axis, where x_j + L lands in [64, 72) and rounds (x_j
    an odd multiple of 2^-18), one inside the box (S = 0, every difference exact).  All
    coordinates and D are multiples of 2^-18, so D is exact in FP32."""
    rng = np.random.default_rng(seed)
    q = 2.0 ** -18
    out = []
    for k in range(count):
        a = k % 3
        xj_face = 4.0 + (2 * int(rng.integers(0, 2 ** 17)) + 1) * q      # in [4, 5)
        xi_face = 56.0 + int(rng.integers(0, 3 * 2 ** 18)) * q              # in [56, 59)
        D = np.zeros(3)
        D[a] = xi_face - xj_face - L                                        # in (-9, -5)
        if D[a] < -10.0 or D[a] > -2.0:
            continue
        for b in range(3):
            if b != a:
                D[b] = int(rng.integers(-3 * 2 ** 18, 3 * 2 ** 18)) * q
        face = np.full((2, 3), 30.0)
        face[0] = 30.0 + D
        face[0, a], face[1, a] = xi_face, xj_face
        inner = np.full((2, 3), 30.0)
        inner[0] = 30.0 + D
        out.append((face.astype(np.float32), inner.astype(np.float32)))
    return out


def test_analytic_separation_exact_across_the_periodic_face(apoa1):
    """The analytic mode evaluates a pair across a periodic face on its exact separation:
    the int64 forces equal those of the same separation inside the box, bit for bit.  The
    table path's pre-shifted image fl(x_j + L) carries the rounding (its forces differ on
    some of these pairs, which shows the comparison can fail)."""
    base = mdrl.SyntheticSystem(4, 3)
    L, rc = 60.0, 12.0
    box = np.full(3, L, np.float32)
    t = np.array([1, 2], np.int32)
    q = np.array([0.3, -0.3], np.float32)
    tab = apoa1_tab6(apoa1)
    cases = face_and_interior_pairs(100, L=L)
    assert len(cases) > 50
    differ_tab = 0
    for face, inner in cases:
        got = []
        for mode in (O.MODE_ANALYTIC, 0):
            ff, nf = O.RL(face, box, [5] * 3, rc, t, base.eps, base.sigma, tab, charge=q,
                          coulomb_k=base.coulomb_k, mode=mode).forces_fix()
            fi, ni = O.RL(inner, box, [5] * 3, rc, t, base.eps, base.sigma, tab, charge=q,
                          coulomb_k=base.coulomb_k, mode=mode).forces_fix()
            assert nf == ni == 1
            got.append((ff, fi))
        (fa, fai), (ft, fti) = got
        assert np.array_equal(fa, fai), (face, inner)
        differ_tab += not np.array_equal(ft, fti)
    assert differ_tab >= 5


def test_half_shell_newton_third_law_exact(apoa1):
    s = mdrl.SyntheticSystem(1)
    rl = rl_from_synthetic(s, apoa1)
    fix, npairs = rl.forces_fix()
    assert np.all(fix.sum(axis=0) == 0)
    assert 6000 < npairs < 9000


def test_half_shell_pairs_equal_brute_force(apoa1):
    """Every unordered in-cutoff pair (all images) is evaluated exactly once."""
    s = mdrl.SyntheticSystem(1)
    rl = rl_from_synthetic(s, apoa1)
    _, npairs = rl.forces_fix()
    _, _, _, nbrute = O.fp64_direct(s.particle[:, :3], s.type, s.eps, s.sigma, s.box, s.rc,
                                    use_cells=False)
    assert npairs == nbrute


def test_anisotropic_box_pairs_equal_brute_force(apoa1):
    """3 x 5 x 7 cells on sides that are not multiples of rc: the half-shell pairs are the
    brute-force pairs over all images, the forces sum to zero exactly, and they agree with
    FP64 direct evaluation to the table model's error."""
    s = anisotropic_system()
    rl = rl_from_synthetic(s, apoa1)
    assert list(rl.s.ncell) == [3, 5, 7]
    fix, npairs = rl.forces_fix()
    assert np.all(fix.sum(axis=0) == 0)
    fd, _, _, nbrute = O.fp64_direct(s.particle[:, :3], s.type, s.eps, s.sigma, s.box, s.rc,
                                     charge=s.charge, ke=s.coulomb_k, use_cells=False)
    assert npairs == nbrute
    f = fix.astype(np.float64) / 2 ** 32 * 0.1
    assert np.linalg.norm(f - fd) / np.linalg.norm(fd) < 1e-4


@pytest.mark.parametrize("which,index", [(1, 0), (4, 3)])
def test_fp64_virial(which, index):
    """The FP64 oracle's pair virial W = sum_ij F_ij (x) r_ij: the cell-list path equals the
    brute-force path over all images, and (argon) its trace equals -dU/dl for positions and
    box scaled by l, a central difference of the oracle's own energy."""
    s = mdrl.SyntheticSystem(which, index)
    q = s.charge if s.coulomb_k else None
    args = (s.type, s.eps, s.sigma)
    _, e, w, _ = O.fp64_direct(s.particle[:, :3], *args, s.box, s.rc, charge=q, ke=s.coulomb_k)
    _, eb, wb, _ = O.fp64_direct(s.particle[:, :3], *args, s.box, s.rc, charge=q,
                                 ke=s.coulomb_k, use_cells=False)
    assert np.linalg.norm(w) > 0
    assert eb == pytest.approx(e, rel=1e-12)
    assert np.abs(w - wb).max() <= 1e-12 * np.abs(wb).max()
    assert np.abs(w - w.T).max() <= 1e-12 * np.abs(w).max()
    if which == 1:
        # no pair lies within rc * 1e-7 of the cutoff, so none crosses it under the scaling
        h = 1e-7
        x = s.particle[:, :3].astype(np.float64)
        up = O.fp64_direct(x * (1 + h), *args, s.box.astype(np.float64) * (1 + h), s.rc)[1]
        dn = O.fp64_direct(x * (1 - h), *args, s.box.astype(np.float64) * (1 - h), s.rc)[1]
        assert np.trace(w) == pytest.approx(-(up - dn) / (2 * h), rel=1e-6)


@pytest.mark.parametrize("which,index", [(1, 0), (4, 3), (4, 11)])
def test_half_shell_vs_fp64(apoa1, which, index):
    """Kernel arithmetic vs FP64 of the same table model (error (i)), and the table model vs
    the FP64 analytic oracle (error (ii)); analytic FP32 mode vs FP64 analytic."""
    s = mdrl.SyntheticSystem(which, index)
    rl = rl_from_synthetic(s, apoa1)
    fix, _ = rl.forces_fix()
    f = fix.astype(np.float64) / 2 ** 32 * 0.1
    ft, _ = rl.forces_table_fp64()
    fd, _, _, _ = O.fp64_direct(s.particle[:, :3], s.type, s.eps, s.sigma, s.box, s.rc,
                                charge=s.charge if s.coulomb_k else None, ke=s.coulomb_k)
    err_i = np.linalg.norm(f - ft) / np.linalg.norm(ft)
    err_ii = np.linalg.norm(ft - fd) / np.linalg.norm(fd)
    assert err_i < 1e-5
    assert err_ii < 1e-4
    rla = rl_from_synthetic(s, apoa1, mode=O.MODE_ANALYTIC)
    fa, _ = rla.forces_fix()
    err_a = np.linalg.norm(fa.astype(np.float64) / 2 ** 32 * 0.1 - fd) / np.linalg.norm(fd)
    assert err_a < 1e-5


def test_two_atom_table_force(apoa1):
    """One pair along x across the periodic boundary: F = A*t14 - B*t8 of the table."""
    box = np.array([40.0, 40.0, 40.0], np.float32)
    xyz = np.array([[0.5, 20.0, 20.0], [36.5, 20.0, 20.0]], np.float32)   # r = 4 across x = 0
    rl = O.RL(xyz, box, (4, 4, 4), 10.0, [0, 0], [0.996], [3.4], apoa1_tab6(apoa1))
    fix, npairs = rl.forces_fix()
    assert npairs == 1
    f = fix.astype(np.float64) / 2 ** 32 * 0.1
    r = 4.0
    exact = 48 * 0.996 * 3.4 ** 12 / r ** 13 - 24 * 0.996 * 3.4 ** 6 / r ** 7
    assert f[0, 0] == pytest.approx(exact, rel=2e-4)
    assert f[1, 0] == -f[0, 0]


# Known answers of the reference's testbenches: r2_compute_with_pbc_tb.v:70-129 and
# RL_LJ_Evaluate_Pairs_1st_Order_tb.v:87-104 (d = ref - nbr and r2 in the comments).
TB_REF = np.array([[2, 4, 8], [2, 2, 2], [1, 1, 1], [1, 2, 4], [2, 2, 2]], np.float32)
TB_NBR = np.array([[1, 1, 1], [1, 1, 1], [2, 4, 8], [0, 0, 0], [0, 0, 0]], np.float32)
TB_R2 = ["426C0000", "40400000", "426C0000", "41A80000", "41400000"]     # 59, 3, 59, 21, 12
TB_D = [["3F800000", "40400000", "40E00000"], ["3F800000"] * 3, ["BF800000", "C0400000", "C0E00000"],
        ["3F800000", "40000000", "40800000"], ["40000000"] * 3]


def test_testbench_r2_vectors(apoa1):
    out = O.pair_golden(*apoa1[mdrl.TERM_14], *apoa1[mdrl.TERM_8], TB_REF, TB_NBR, 196.0)
    assert np.array_equal(out[:, 4].view(np.uint32), hex2f(TB_R2).view(np.uint32))
    assert np.array_equal(out[:, 5:8].view(np.uint32), hex2f(sum(TB_D, [])).reshape(5, 3).view(np.uint32))


def test_fix_rounding_and_saturation():
    assert O.fix(1.0) == 2 ** 32
    assert O.fix(-0.25) == -(2 ** 30)
    assert O.fix(2.0 ** -33) == 0          # ties to even
    assert O.fix(3 * 2.0 ** -33) == 2      # 1.5 -> 2
    assert O.fix(1e30) == 2 ** 62
    assert O.fix(-1e30) == -(2 ** 62)


def test_lr_fix_rounding_and_saturation():
    """Outside forces (kJ/mol/A) in the accumulator unit: 10 * 2^32 per kJ/mol/A."""
    f = np.array([1.0, -0.1, 2.0 ** -34, 3 * 2.0 ** -34, -(2.0 ** -34), 5 * 2.0 ** -35,
                  1e15, -1e15, np.nan, 0.0], np.float32)
    got = O.lr_fix(f)
    assert got[0] == 10 * 2 ** 32
    assert got[1] == round(float(np.float32(-0.1)) * 10 * 2 ** 32)
    assert got[2] == 2 and got[3] == 8 and got[4] == -2     # 2.5 -> 2, 7.5 -> 8, -2.5 -> -2
    assert got[5] == round(5 * 10 / 8)                      # 6.25 -> 6
    assert got[6] == 2 ** 62 and got[7] == -(2 ** 62) and got[8] == 0 and got[9] == 0
    # where f * 10 is exact in FP32 it is the pair-force conversion of f * 10 (kJ/mol/nm)
    for v in (0.75, -3.125, 1024.5, 2.0 ** -20):
        assert O.lr_fix(np.float32(v)) == O.fix(np.float32(v * 10))


def test_motion_update_closed_forms():
    """Free flight: x' = x + v dt.  Constant force: kick-drift and velocity Verlet both give
    v_n = v_0 + n a dt; VV positions follow x0 + v0 t + a t^2 / 2."""
    box = np.array([100, 100, 100], np.float32)
    pos = np.array([[10, 20, 30, 0]], np.float32)
    vel = np.array([[0.01, -0.02, 0.005]], np.float32)
    hk = O.half_kick([40.0], 2.0)
    p1, v1, _ = O.motion_update(pos, vel, np.zeros((1, 3), np.int64), [0], hk, [40.0], 2.0, box, True)
    assert np.array_equal(v1, vel)
    assert np.array_equal(p1[0, :3], (pos[0, :3] + vel[0] * np.float32(2.0)).astype(np.float32))
    F = 2.0 ** 32 * 10.0                   # 10 kJ/mol/nm
    fix = np.array([[int(F), 0, 0]], np.int64)
    a = 10.0 * 1e-5 / 40.0                 # A/fs^2
    p, v = pos.copy(), vel.copy()
    for k in range(50):
        p, v, _ = O.motion_update(p, v, fix, [0], hk, [40.0], 2.0, box, k == 0)
    # vh after 50 steps = v0 + (50 - 0.5) * a * dt (half-step velocity)
    assert float(v[0, 0]) == pytest.approx(0.01 + 49.5 * a * 2.0, rel=1e-5)
    x_half = 10 + sum((0.01 + (k + 0.5) * a * 2.0) * 2.0 for k in range(50))
    assert float(p[0, 0]) == pytest.approx(x_half, rel=1e-6)
