"""ctypes wrapper of liboracle.so, the CPU restatement of the reference's range-limited path.

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by bench.py's CPU
baseline leg, never by the product package.  See oracle.h for the reference file:line each
function restates and for what pins it.
"""
import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "md_eval_model_ref")

MODE_PARITY, MODE_COULOMB, MODE_ANALYTIC = 1, 2, 4

_lib = None

F = ctypes.POINTER(ctypes.c_float)
D = ctypes.POINTER(ctypes.c_double)
I = ctypes.POINTER(ctypes.c_int)
I64 = ctypes.POINTER(ctypes.c_int64)


class CModelParams(ctypes.Structure):
    _fields_ = [("eps", ctypes.c_float), ("sigma", ctypes.c_float), ("mass", ctypes.c_float),
                ("dt", ctypes.c_float), ("rc", ctypes.c_float), ("box", ctypes.c_float * 3),
                ("iterations", ctypes.c_int)]


class RLSystem(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("box", ctypes.c_float * 3), ("ncell", ctypes.c_int * 3),
                ("rc2", ctypes.c_float), ("pos4", F), ("type", I), ("ntypes", ctypes.c_int),
                ("pair_a", F), ("pair_b", F), ("tab", F), ("exp_offset", ctypes.c_int),
                ("nseg", ctypes.c_int), ("mode", ctypes.c_int), ("excl", ctypes.c_void_p),
                ("rc", ctypes.c_float)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.or_lut_index.argtypes = [ctypes.c_float, ctypes.c_int, ctypes.c_int]
        L.or_pair_golden.argtypes = [F, F, F, F, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                     ctypes.c_int, F, F, ctypes.c_int64, F]
        L.or_pair_golden.restype = None
        L.or_cmodel_default_params.argtypes = [ctypes.POINTER(CModelParams)]
        L.or_cmodel_default_params.restype = None
        L.or_cmodel_run.argtypes = [F, ctypes.c_int, ctypes.POINTER(CModelParams), F,
                                    ctypes.c_char_p, ctypes.c_int]
        L.or_cmodel_force_sample.argtypes = [F, ctypes.c_int, ctypes.POINTER(CModelParams),
                                             ctypes.c_int, F]
        L.or_cmodel_force_sample.restype = ctypes.c_int64
        L.or_bin_ceil.argtypes = [F, ctypes.c_int, ctypes.c_float, I, I]
        L.or_pair_stream_text.argtypes = [D, ctypes.c_int, ctypes.c_float, I, I, ctypes.c_int,
                                          ctypes.c_char_p, ctypes.c_int64]
        L.or_pair_stream_text.restype = ctypes.c_int64
        L.or_cell_of.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_int]
        L.or_set_threads.argtypes = [ctypes.c_int]
        L.or_set_threads.restype = None
        L.or_fix.argtypes = [ctypes.c_float]
        L.or_fix.restype = ctypes.c_int64
        L.or_rl_forces_fix.argtypes = [ctypes.POINTER(RLSystem), I64]
        L.or_rl_forces_fix.restype = ctypes.c_int64
        L.or_table_fp64.argtypes = [ctypes.POINTER(RLSystem), D]
        L.or_table_fp64.restype = ctypes.c_int64
        L.or_fp64_direct.argtypes = [ctypes.c_int, D, I, D, ctypes.c_int, D, D, ctypes.c_double,
                                     ctypes.c_int, D, ctypes.c_double, ctypes.c_int, D, D, D,
                                     ctypes.c_void_p]
        L.or_fp64_direct.restype = ctypes.c_int64
        L.or_atan2f.argtypes = [ctypes.c_float, ctypes.c_float]
        L.or_atan2f.restype = ctypes.c_float
        L.or_bonded_fix.argtypes = [ctypes.c_int, F, F, ctypes.c_int, I, F, I64, D]
        L.or_bonded_fix.restype = None
        L.or_bonded_fp64.argtypes = [ctypes.c_int, D, D, ctypes.c_int, I, D, D]
        L.or_bonded_fp64.restype = ctypes.c_double
        L.or_motion_update.argtypes = [ctypes.c_int, F, F, I64, I, D, D, ctypes.c_float, F,
                                       ctypes.c_int, ctypes.c_int, D]
        L.or_motion_update.restype = None
        L.or_minimize_update.argtypes = [ctypes.c_int, F, I64, ctypes.c_double, ctypes.c_double, F]
        L.or_minimize_update.restype = None
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


# ---------------------------------------------------------------------------- tables
def lut_index(r2, exp_offset=121, nseg=14):
    return lib().or_lut_index(float(np.float32(r2)), exp_offset, nseg)


def pair_golden(c0_14, c1_14, c0_8, c1_8, ref, nbr, rc2, exp_offset=121, nseg=14, strict=True):
    ref = np.ascontiguousarray(ref, np.float32).reshape(-1, 3)
    nbr = np.ascontiguousarray(nbr, np.float32).reshape(-1, 3)
    tabs = [np.ascontiguousarray(t, np.float32) for t in (c0_14, c1_14, c0_8, c1_8)]
    out = np.zeros((ref.shape[0], 8), np.float32)
    lib().or_pair_golden(*[_p(t, F) for t in tabs], exp_offset, nseg, float(rc2), int(strict),
                         _p(ref, F), _p(nbr, F), ref.shape[0], _p(out, F))
    return out


# ---------------------------------------------------------------------------- C model
def cmodel_default_params():
    p = CModelParams()
    lib().or_cmodel_default_params(ctypes.byref(p))
    return p


def cmodel_run(particle, params, with_lines=False):
    a = np.ascontiguousarray(particle, np.float32).copy()
    n = a.shape[0]
    e = np.zeros(2 * params.iterations, np.float32)
    buf = ctypes.create_string_buffer(128 * params.iterations + 16) if with_lines else None
    lib().or_cmodel_run(_p(a, F), n, ctypes.byref(params), _p(e, F), buf, 0)
    lines = buf.value.decode() if with_lines else None
    return a, e.reshape(-1, 2), lines


def cmodel_force_sample(particle, params, nref):
    a = np.ascontiguousarray(particle, np.float32)
    f = np.zeros((nref, 3), np.float32)
    cnt = lib().or_cmodel_force_sample(_p(a, F), a.shape[0], ctypes.byref(params), nref, _p(f, F))
    return cnt, f


# ---------------------------------------------------------------------------- binning
def bin_ceil(pos, rc, ncell):
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    nc = np.ascontiguousarray(ncell, np.int32)
    out = np.zeros((pos.shape[0], 3), np.int32)
    nout = lib().or_bin_ceil(_p(pos, F), pos.shape[0], float(rc), _p(nc, I), _p(out, I))
    return out, nout


def pair_stream_text(raw_xyz, rc, ncell, home, cell_particle_max=100):
    raw = np.ascontiguousarray(raw_xyz, np.float64).reshape(-1, 3)
    nc = np.ascontiguousarray(ncell, np.int32)
    hm = np.ascontiguousarray(home, np.int32)
    cap = 1 << 20
    while True:
        buf = ctypes.create_string_buffer(cap)
        n = lib().or_pair_stream_text(_p(raw, D), raw.shape[0], float(rc), _p(nc, I), _p(hm, I),
                                      cell_particle_max, buf, cap)
        if n >= 0:
            return buf.raw[:n]
        cap = -n + 16


def cell_of(x, inv_edge, nc):
    return lib().or_cell_of(float(x), float(inv_edge), nc)


def fix(f):
    return lib().or_fix(float(np.float32(f)))


def set_threads(n):
    """OpenMP threads of this (Python) thread's later oracle calls."""
    lib().or_set_threads(int(n))


def map_systems(fn, items, workers=None, cost=None):
    """fn over items on a pool of threads, each running the oracle on one OpenMP thread
    (many small systems side by side instead of one at a time over all cores), the most
    costly items (cost(item), default its atom count n) submitted first; results in input
    order."""
    from concurrent.futures import ThreadPoolExecutor
    workers = workers or min(16, os.cpu_count() or 1)
    cost = cost or (lambda it: getattr(it, "n", 0))
    order = sorted(range(len(items)), key=lambda k: -cost(items[k]))
    out = [None] * len(items)
    with ThreadPoolExecutor(max_workers=workers, initializer=set_threads, initargs=(1,)) as ex:
        futs = {k: ex.submit(fn, items[k]) for k in order}
        for k, f in futs.items():
            out[k] = f.result()
    return out


# ---------------------------------------------------------------------------- host-side setup
# The same host arithmetic as libmdrl's md_rl_create (Lorentz-Berthelot in double, rounded
# once to float; forces in kJ/mol/nm, hence the factor 10).
def pair_multipliers(eps, sigma, analytic=False):
    nt = len(eps)
    a = np.zeros(nt * nt, np.float32)
    b = np.zeros(nt * nt, np.float32)
    for i in range(nt):
        for j in range(nt):
            sg = 0.5 * (float(np.float32(sigma[i])) + float(np.float32(sigma[j])))
            ep = math.sqrt(float(np.float32(eps[i])) * float(np.float32(eps[j])))
            s6 = math.pow(sg, 6.0)
            s12 = s6 * s6
            if analytic:
                a[i * nt + j] = np.float32(480.0 * ep * s12)
                b[i * nt + j] = np.float32(240.0 * ep * s6)
            else:
                a[i * nt + j] = np.float32(10.0 * ep * s12)
                b[i * nt + j] = np.float32(10.0 * ep * s6)
    return a, b


def charge_scale(coulomb_k):
    return math.sqrt(10.0 * float(np.float32(coulomb_k)) * 4.0 * 3.14) if coulomb_k else 0.0


def half_kick(mass, dt):
    return np.array([math.ldexp(1.0e-5 / float(np.float32(m)) * (0.5 * float(np.float32(dt))), -32)
                     for m in mass], np.float64)


def interleave(tabs):
    """tabs: list of 6 arrays c0_a, c1_a, c0_b, c1_b, c0_c, c1_c (None -> zeros)."""
    n = len(tabs[0])
    out = np.zeros((n, 6), np.float32)
    for k, t in enumerate(tabs):
        if t is not None:
            out[:, k] = t
    return out


class RL:
    """An or_rl_system with owned arrays."""

    def __init__(self, pos_xyz, box, ncell, rc, types, eps, sigma, tab6, charge=None,
                 coulomb_k=0.0, mode=0, exp_offset=121, nseg=14, excl=None):
        n = pos_xyz.shape[0]
        self.pos4 = np.zeros((n, 4), np.float32)
        self.pos4[:, :3] = pos_xyz
        qs = charge_scale(coulomb_k)
        if charge is not None and coulomb_k:
            self.pos4[:, 3] = (np.asarray(charge, np.float32).astype(np.float64) * qs).astype(np.float32)
            mode |= MODE_COULOMB
        self.type = np.ascontiguousarray(types, np.int32)
        self.pa, self.pb = pair_multipliers(eps, sigma, analytic=bool(mode & MODE_ANALYTIC))
        self.tab = np.ascontiguousarray(tab6, np.float32)
        s = RLSystem()
        s.n = n
        s.box[:] = [float(np.float32(b)) for b in box]
        s.ncell[:] = list(ncell)
        s.rc2 = float(np.float32(rc) * np.float32(rc))
        s.rc = float(np.float32(rc))
        s.pos4, s.type = _p(self.pos4, F), _p(self.type, I)
        s.ntypes = len(eps)
        s.pair_a, s.pair_b, s.tab = _p(self.pa, F), _p(self.pb, F), _p(self.tab, F)
        s.exp_offset, s.nseg, s.mode = exp_offset, nseg, mode
        self.excl = None if excl is None else np.ascontiguousarray(excl, np.uint32)
        s.excl = None if self.excl is None else self.excl.ctypes.data_as(ctypes.c_void_p)
        self.s = s
        self.n = n

    def forces_fix(self):
        out = np.zeros((self.n, 3), np.int64)
        npairs = lib().or_rl_forces_fix(ctypes.byref(self.s), _p(out, I64))
        return out, npairs

    def forces_table_fp64(self):
        out = np.zeros((self.n, 3), np.float64)
        npairs = lib().or_table_fp64(ctypes.byref(self.s), _p(out, D))
        return out, npairs


def fp64_direct(xyz, types, eps, sigma, box, rc, charge=None, ke=0.0, use_cells=True, excl=None):
    """Analytic FP64 forces (kJ/mol/A), LJ with Lorentz-Berthelot + cutoff Coulomb;
    excl: optional exclusion windows (bit d of excl[i] excludes the pair (i, i + d - 16))."""
    xyz = np.ascontiguousarray(xyz, np.float64)
    n = xyz.shape[0]
    nt = len(eps)
    e = np.zeros(nt * nt)
    s = np.zeros(nt * nt)
    for i in range(nt):
        for j in range(nt):
            s[i * nt + j] = 0.5 * (float(sigma[i]) + float(sigma[j]))
            e[i * nt + j] = math.sqrt(float(eps[i]) * float(eps[j]))
    t = np.ascontiguousarray(types, np.int32)
    q = None if charge is None else np.ascontiguousarray(charge, np.float64)
    bx = np.ascontiguousarray(box, np.float64)
    f = np.zeros((n, 3))
    en = ctypes.c_double()
    vir = np.zeros(9)
    npairs = lib().or_fp64_direct(n, _p(xyz, D), _p(t, I), None if q is None else _p(q, D), nt,
                                  _p(e, D), _p(s, D), float(ke), 1 if (q is not None and ke) else 0,
                                  _p(bx, D), float(rc), 1 if use_cells else 0, _p(f, D),
                                  ctypes.byref(en), _p(vir, D),
                                  None if excl is None else
                                  np.ascontiguousarray(excl, np.uint32).ctypes.data_as(ctypes.c_void_p))
    return f, en.value, vir.reshape(3, 3), npairs


def motion_update(pos4, vel3, fixf, types, hk, mass, dt, box, first, kick_drift=False):
    pos4 = np.ascontiguousarray(pos4, np.float32).copy()
    vel3 = np.ascontiguousarray(vel3, np.float32).copy()
    fixf = np.ascontiguousarray(fixf, np.int64)
    t = np.ascontiguousarray(types, np.int32)
    hk = np.ascontiguousarray(hk, np.float64)
    ms = np.ascontiguousarray(mass, np.float64)
    bx = np.ascontiguousarray(box, np.float32)
    ke = ctypes.c_double()
    lib().or_motion_update(pos4.shape[0], _p(pos4, F), _p(vel3, F), _p(fixf, I64), _p(t, I),
                           _p(hk, D), _p(ms, D), float(dt), _p(bx, F), int(first),
                           int(kick_drift), ctypes.byref(ke))
    return pos4, vel3, ke.value


# ---------------------------------------------------------------------------- summation hook
def lr_fix(f):
    """Forces of an outside provider (kJ/mol/A, FP32) in the fixed-point unit of the
    accumulators: rint(f * 10 * 2^32), ties to even, saturated at +-2^62, NaN to 0.  The
    reference adds its short-range, long-range and bonded forces in FP32 per particle
    (Summation_Logic/Force_Summation_Unit.v:62-126, long-range + short-range first, then
    bonded); here the three meet in int64, where the sum does not depend on the order.
    Pinned by exact values (tests/test_oracle_golden.py::test_lr_fix_rounding_and_saturation):
    ties to even, saturation, NaN, and equality with the pair-force conversion of f * 10
    wherever that product is exact in FP32."""
    x = np.asarray(f, np.float32).astype(np.float64) * 42949672960.0
    x = np.where(np.isnan(x), 0.0, np.clip(x, -2.0 ** 62, 2.0 ** 62))
    return np.rint(x).astype(np.int64)


# ---------------------------------------------------------------------------- cell init files
def generator_positions(raw_xyz):
    """The cell-file generators' preprocessing (M/GenInputCellList_LJArgon.m:86-99,
    M/GenInputCellList_ApoA1.m:84-97): raw positions minus the per-axis minimum, in FP64,
    stored as FP32."""
    raw = np.asarray(raw_xyz, np.float64).reshape(-1, 3)
    return (raw - raw.min(axis=0)).astype(np.float32)


def generator_cell_lists(pos, edge, ncell, rule="floor"):
    """The particle-to-cell loop of M/GenInputCellList_LJArgon.m:105-135 (rule "floor":
    cell = floor(x / edge) + 1) and M/GenInputCellList_ApoA1.m:103-133 (rule "ceil":
    ceil(x / edge)), per coordinate, FP32 division, a coordinate of exactly 0 in cell 1;
    particles outside 1..ncell are dropped.  Returns ({(X, Y, Z): k x 3 FP32 positions in
    input order}, number dropped).  Pinned by the reference's own output of this script: all
    92 LJArgon cell files shipped in D/CellMemoryModules (tests/test_cell_files.py)."""
    pos = np.asarray(pos, np.float32).reshape(-1, 3)
    e = np.float32(edge)
    lists, dropped = {}, 0
    for p in pos:
        c = []
        for d in range(3):
            if p[d] == 0:
                c.append(1)
            elif rule == "floor":
                c.append(int(math.floor(float(np.float32(p[d] / e)))) + 1)
            else:
                c.append(int(math.ceil(float(np.float32(p[d] / e)))))
        if all(1 <= c[d] <= ncell[d] for d in range(3)):
            lists.setdefault(tuple(c), []).append(p)
        else:
            dropped += 1
    return {k: np.array(v, np.float32) for k, v in lists.items()}, dropped


def run_reference_cmodel(pdb_text, workdir):
    """Runs the reference C model (built under _ref/ from the reference sources) on a PDB.

    The model opens the relative path below (R/C_Model_ForceEval/MD_Evaluation_Model.cpp:30-32);
    the PDB is written there inside workdir.  Returns the printed lines."""
    rel = ("C:/Users/Ethan/Desktop/WorkingFolder/Ethan_RL_Pipeline_1st_Order_SingleFloat_18.0/"
           "SourceCode/MatlabScripts")
    d = os.path.join(workdir, rel)
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "ar_gas.pdb"), "w") as f:
        f.write(pdb_text)
    r = subprocess.run([REF_BIN], cwd=workdir, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError("reference C model failed: %s" % r.stderr)
    return r.stdout


def minimize_update(pos4, fixf, gamma, dmax, box):
    pos4 = np.ascontiguousarray(pos4, np.float32).copy()
    fixf = np.ascontiguousarray(fixf, np.int64)
    bx = np.ascontiguousarray(box, np.float32)
    lib().or_minimize_update(pos4.shape[0], _p(pos4, F), _p(fixf, I64), float(gamma),
                             float(dmax), _p(bx, F))
    return pos4


# ---------------------------------------------------------------------------- bonded
def atan2f(y, x):
    return lib().or_atan2f(float(np.float32(y)), float(np.float32(x)))


def device_bonded_params(atoms, params):
    """User-level term parameters -> the FP32 constants the kernels use, with the same
    host arithmetic as the library: bond/angle (x0, 10*k), torsion (10*k, n, cos d, sin d)."""
    atoms = np.ascontiguousarray(atoms, np.int32).reshape(-1, 4)
    params = np.ascontiguousarray(params, np.float32).reshape(-1, 4)
    out = np.zeros_like(params)
    tors = atoms[:, 3] >= 0
    out[~tors, 0] = params[~tors, 0]
    out[~tors, 1] = np.float32(10.0) * params[~tors, 1]
    out[tors, 0] = np.float32(10.0) * params[tors, 0]
    out[tors, 1] = params[tors, 1]
    ph = params[tors, 2].astype(np.float64)
    out[tors, 2] = np.cos(ph).astype(np.float32)
    out[tors, 3] = np.sin(ph).astype(np.float32)
    return out


def bonded_fix(pos4, box, atoms, params):
    """int64 bonded forces (2^-32 kJ/mol/nm) and FP64 per-term energies (kJ/mol), from
    user-level parameters (converted by device_bonded_params)."""
    pos4 = np.ascontiguousarray(pos4, np.float32)
    n = pos4.shape[0]
    atoms = np.ascontiguousarray(atoms, np.int32).reshape(-1, 4)
    dp = device_bonded_params(atoms, params)
    bx = np.ascontiguousarray(box, np.float32)
    fix = np.zeros((n, 3), np.int64)
    e = np.zeros(atoms.shape[0])
    lib().or_bonded_fix(n, _p(pos4, F), _p(bx, F), atoms.shape[0], _p(atoms, I), _p(dp, F),
                        _p(fix, I64), _p(e, D))
    return fix, e


def bonded_fp64(xyz, box, atoms, params):
    """FP64 bonded forces (kJ/mol/A) and total energy (kJ/mol) from user-level parameters."""
    xyz = np.ascontiguousarray(xyz, np.float64)
    atoms = np.ascontiguousarray(atoms, np.int32).reshape(-1, 4)
    pp = np.ascontiguousarray(params, np.float64).reshape(-1, 4)
    bx = np.ascontiguousarray(box, np.float64)
    f = np.zeros((xyz.shape[0], 3))
    e = lib().or_bonded_fp64(xyz.shape[0], _p(xyz, D), _p(bx, D), atoms.shape[0], _p(atoms, I),
                             _p(pp, D), _p(f, D))
    return f, e
