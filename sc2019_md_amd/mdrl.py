"""ctypes bindings of libmdrl.so (the C ABI declared in include/mdrl.h).

The Python side is a thin host mirror of the reference's interfaces for this path:
``eval_pairs`` is the OpenCL ``LJ`` kernel / ``RL_LJ_Evaluation`` function
(Range_Limited_demo/SourceCode/OpenCL_HDL_Lib/test_hdl_lib/device/LJ.cl:31-80),
``System`` drives the half-shell force path on the C model's ``particle[N][11]`` layout
(Range_Limited/C_Model_ForceEval/MD_Evaluation_Model.cpp:35), and ``cmodel_run`` /
``MD_Evaluation_Model`` are the C model itself (:9-205).

There is no fallback: if libmdrl.so is missing or no gfx950 device is visible, the calls
raise.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmdrl.so")

MDRL_OK = 0
ERRORS = {
    -1: "EINVAL", -2: "ENOMEM", -3: "EIO", -4: "EHIP", -5: "ECELL", -6: "ENONFINITE",
    -7: "ESATURATED", -8: "ENODEV",
}

# pair flags
PAIR_PARITY, PAIR_STRICT, PAIR_COULOMB = 1, 2, 4
# system flags
ACC_FP32, EVAL_ANALYTIC, EVAL_PARITY, KICK_DRIFT, NO_COULOMB, COARSE_CELLS = 1, 2, 4, 8, 16, 32
# sticky kernel flags (Stats.flags)
F_SATURATED, F_NONFINITE, F_CELL_OVERFLOW = 1, 2, 4
TERM_14, TERM_8, TERM_3, TERM_12, TERM_6, TERM_1 = range(6)
NTERMS = 6

_lib = None


class MdrlError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "?"), code, msg))
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int), ("box", ctypes.c_float * 3), ("rc", ctypes.c_float),
        ("ncell", ctypes.c_int * 3), ("dt", ctypes.c_float), ("ntypes", ctypes.c_int),
        ("eps", ctypes.POINTER(ctypes.c_float)), ("sigma", ctypes.POINTER(ctypes.c_float)),
        ("mass", ctypes.POINTER(ctypes.c_float)), ("coulomb_k", ctypes.c_float),
        ("flags", ctypes.c_int), ("table", ctypes.c_void_p), ("cell_capacity", ctypes.c_int),
        ("reserved", ctypes.c_int * 7),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("pair_evals", ctypes.c_int64), ("candidates", ctypes.c_int64),
        ("max_cell_count", ctypes.c_int), ("flags", ctypes.c_int),
        ("force_ms", ctypes.c_double), ("step_ms", ctypes.c_double),
        ("timed_steps", ctypes.c_int),
    ]


class Schedule(ctypes.Structure):
    _fields_ = [("atoms_per_wave", ctypes.c_int), ("tile_depth", ctypes.c_int),
                ("offset_groups", ctypes.c_int), ("persistent", ctypes.c_int)]

    def __repr__(self):
        return "Schedule(apw=%d, tile=%d, groups=%d, persistent=%d)" % (
            self.atoms_per_wave, self.tile_depth, self.offset_groups, self.persistent)

    def as_dict(self):
        return {"atoms_per_wave": self.atoms_per_wave, "tile_depth": self.tile_depth,
                "offset_groups": self.offset_groups, "persistent": self.persistent}


class CModelCfg(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int), ("iterations", ctypes.c_int), ("eps", ctypes.c_float),
        ("sigma", ctypes.c_float), ("mass", ctypes.c_float), ("dt", ctypes.c_float),
        ("rc", ctypes.c_float), ("box", ctypes.c_float * 3), ("print", ctypes.c_int),
    ]


class EnergyTerms(ctypes.Structure):
    _fields_ = [("kinetic", ctypes.c_double), ("potential", ctypes.c_double),
                ("range_limited", ctypes.c_double), ("bonded", ctypes.c_double),
                ("long_range", ctypes.c_double), ("virial", ctypes.c_double * 9),
                ("cutoff_shift", ctypes.c_double)]


class SystemDesc(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int), ("ntypes", ctypes.c_int), ("box", ctypes.c_float * 3),
        ("rc", ctypes.c_float), ("dt", ctypes.c_float), ("coulomb_k", ctypes.c_float),
        ("particle", ctypes.POINTER(ctypes.c_float)), ("type", ctypes.POINTER(ctypes.c_int32)),
        ("charge", ctypes.POINTER(ctypes.c_float)), ("eps", ctypes.c_float * 8),
        ("sigma", ctypes.c_float * 8), ("mass", ctypes.c_float * 8),
        ("nterms", ctypes.c_int), ("term_atoms", ctypes.POINTER(ctypes.c_int32)),
        ("term_params", ctypes.POINTER(ctypes.c_float)), ("excl", ctypes.POINTER(ctypes.c_uint32)),
    ]


class LRView(ctypes.Structure):
    """md_rl_lr_view: what a summation-hook provider is handed (device addresses as ints)."""
    _fields_ = [
        ("nsys", ctypes.c_int), ("n", ctypes.c_int),
        ("atom_off", ctypes.POINTER(ctypes.c_int32)), ("box", ctypes.POINTER(ctypes.c_float)),
        ("xq", ctypes.c_void_p), ("force", ctypes.c_void_p), ("energy", ctypes.c_void_p),
        ("want_energy", ctypes.c_int),
    ]


LR_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(LRView), ctypes.c_void_p)


def lib():
    """Loads libmdrl.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libmdrl.so not built: run python sc2019_md_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        F = ctypes.POINTER(ctypes.c_float)
        I = ctypes.c_int
        L.md_rl_last_error.restype = ctypes.c_char_p
        L.md_rl_default_table_dir.restype = ctypes.c_char_p
        L.md_rl_table_read_file.argtypes = [ctypes.c_char_p, F, I]
        L.md_rl_table_load.argtypes = [ctypes.POINTER(P), ctypes.c_char_p, I, I]
        L.md_rl_table_load_default.argtypes = [ctypes.POINTER(P)]
        L.md_rl_table_get.argtypes = [P, I, ctypes.POINTER(F), ctypes.POINTER(F),
                                      ctypes.POINTER(I), ctypes.POINTER(I), ctypes.POINTER(I)]
        L.md_rl_table_free.argtypes = [P]
        L.md_rl_table_from_arrays.argtypes = [ctypes.POINTER(P), I, I, ctypes.POINTER(F), ctypes.POINTER(F)]
        L.md_rl_table_free.restype = None
        L.md_rl_eval_pairs.argtypes = [P, P, P, F, F, F, F, F, F, F, F, F, ctypes.c_int64,
                                       ctypes.c_float, I]
        L.md_rl_eval_pairs_ex.argtypes = [P, F, F, F, F, F, F, F, ctypes.c_int64,
                                          ctypes.c_float, I]
        L.md_rl_eval_pairs_ex_noalloc.argtypes = [P, P, P, P, P, P, P, P, ctypes.c_int64,
                                                  ctypes.c_float, I, P]
        L.md_rl_create.argtypes = [ctypes.POINTER(P), ctypes.POINTER(Params)]
        for name in ("md_rl_set_particles", "md_rl_get_particles"):
            getattr(L, name).argtypes = [P, F, I]
        L.md_rl_set_types.argtypes = [P, ctypes.POINTER(ctypes.c_int32), F, I]
        L.md_rl_force_eval.argtypes = [P]
        L.md_rl_step.argtypes = [P, I]
        L.md_rl_step_async.argtypes = [P, I, P]
        L.md_rl_sync.argtypes = [P]
        L.md_rl_minimize.argtypes = [P, I, ctypes.c_float, ctypes.c_float]
        L.md_rl_get_forces_fixed.argtypes = [P, ctypes.POINTER(ctypes.c_int64), I]
        L.md_rl_energy.argtypes = [P, ctypes.POINTER(ctypes.c_double),
                                   ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.md_rl_get_stats.argtypes = [P, ctypes.POINTER(Stats)]
        L.md_rl_set_timing.argtypes = [P, I]
        L.md_rl_destroy.argtypes = [P]
        L.md_rl_create_batch.argtypes = [ctypes.POINTER(P), ctypes.POINTER(Params), I,
                                         ctypes.POINTER(ctypes.c_int32), F]
        L.md_rl_num_systems.argtypes = [P]
        L.md_rl_system_size.argtypes = [P, I, ctypes.POINTER(ctypes.c_int32), F,
                                        ctypes.POINTER(I)]
        L.md_rl_set_system.argtypes = [P, I, F, ctypes.POINTER(ctypes.c_int32), F]
        L.md_rl_set_bonded.argtypes = [P, I, I, ctypes.POINTER(ctypes.c_int32), F]
        L.md_rl_set_exclusions.argtypes = [P, I, ctypes.POINTER(ctypes.c_uint32)]
        L.md_rl_check_exclusions.argtypes = [ctypes.POINTER(ctypes.c_uint32), I]
        L.md_rl_commit.argtypes = [P]
        L.md_rl_get_system.argtypes = [P, I, F]
        L.md_rl_get_system_forces_fixed.argtypes = [P, I, ctypes.POINTER(ctypes.c_int64)]
        L.md_rl_system_energy.argtypes = [P, I, ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_double)]
        L.md_rl_set_graph.argtypes = [P, I]
        L.md_rl_get_energy_terms.argtypes = [P, I, ctypes.POINTER(EnergyTerms)]
        L.md_rl_graph_prepare.argtypes = [P, I]
        L.md_rl_set_schedule.argtypes = [P, ctypes.POINTER(Schedule)]
        L.md_rl_get_schedule.argtypes = [P, ctypes.POINTER(Schedule)]
        L.md_rl_autotune.argtypes = [P, I, ctypes.POINTER(Schedule), ctypes.POINTER(ctypes.c_double)]
        L.md_rl_autotune_cached.argtypes = [P, I, ctypes.POINTER(Schedule), ctypes.POINTER(ctypes.c_double),
                                            ctypes.POINTER(I)]
        L.md_rl_cmodel_default.argtypes = [ctypes.POINTER(CModelCfg)]
        L.md_rl_cmodel_default.restype = None
        L.md_rl_cmodel_run.argtypes = [ctypes.POINTER(CModelCfg), F, F]
        L.md_rl_make_system.argtypes = [ctypes.POINTER(SystemDesc), I, I, ctypes.c_uint64]
        L.md_rl_free_system.argtypes = [ctypes.POINTER(SystemDesc)]
        L.md_rl_free_system.restype = None
        L.md_rl_set_long_range.argtypes = [P, LR_FN, P]
        L.md_rl_cell_file_read.argtypes = [ctypes.c_char_p, F, I]
        L.md_rl_cell_file_write.argtypes = [ctypes.c_char_p, F, I, I, I]
        L.md_rl_boundary_file_read.argtypes = [ctypes.c_char_p, F, I]
        L.md_rl_boundary_file_write.argtypes = [ctypes.c_char_p, ctypes.c_float, I, I]
        L.md_rl_cell_dir_read.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(ctypes.c_int), F, ctypes.POINTER(ctypes.c_int32), I]
        _lib = L
    return _lib


def _check(code):
    if code != MDRL_OK:
        raise MdrlError(code, lib().md_rl_last_error().decode())
    return code


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        a = a.reshape(shape)
    return a


# ---------------------------------------------------------------------------------- tables
class Table:
    """First-order interpolation tables (the reference's lut*_* ROM contents)."""

    def __init__(self, handle):
        self.handle = handle

    @classmethod
    def load(cls, directory=None, exp_offset=121, nseg=14):
        h = ctypes.c_void_p()
        if directory is None:
            _check(lib().md_rl_table_load_default(ctypes.byref(h)))
        else:
            _check(lib().md_rl_table_load(ctypes.byref(h), directory.encode(), exp_offset, nseg))
        return cls(h)

    @classmethod
    def from_arrays(cls, c0, c1, exp_offset=121, nseg=14):
        """A table from per-term coefficient arrays (c0[k], c1[k]: nseg*256 floats each, or
        None where the term is absent), in MDRL_TERM_* order (md_rl_table_from_arrays)."""
        F = ctypes.POINTER(ctypes.c_float)
        p0, p1, keep = (F * NTERMS)(), (F * NTERMS)(), []
        for k in range(NTERMS):
            if c0[k] is not None and c1[k] is not None:
                a0, a1 = _f32(c0[k], (nseg * 256,)), _f32(c1[k], (nseg * 256,))
                keep += [a0, a1]
                p0[k], p1[k] = _fp(a0), _fp(a1)
        h = ctypes.c_void_p()
        _check(lib().md_rl_table_from_arrays(ctypes.byref(h), exp_offset, nseg, p0, p1))
        return cls(h)

    def get(self, term):
        c0 = ctypes.POINTER(ctypes.c_float)()
        c1 = ctypes.POINTER(ctypes.c_float)()
        n, off, nseg = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib().md_rl_table_get(self.handle, term, ctypes.byref(c0), ctypes.byref(c1),
                                     ctypes.byref(n), ctypes.byref(off), ctypes.byref(nseg)))
        if not c0:
            return None, None
        return (np.ctypeslib.as_array(c0, shape=(n.value,)).copy(),
                np.ctypeslib.as_array(c1, shape=(n.value,)).copy())

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.md_rl_table_free(self.handle)
            self.handle = None


def read_table_file(path, cap=1 << 16):
    out = np.zeros(cap, dtype=np.float32)
    n = lib().md_rl_table_read_file(path.encode(), _fp(out), cap)
    if n < 0:
        _check(n)
    return out[:n]


def default_table_dir():
    return lib().md_rl_default_table_dir().decode()


# ---------------------------------------------------------------------------------- pairs
def eval_pairs(table, ref, nbr, rc2, flags=0, ref_id=None, nbr_id=None):
    """Pair batch as the reference's LJ kernel: ref/nbr (n, 3) -> forces (n, 3)."""
    ref = _f32(ref, (-1, 3))
    nbr = _f32(nbr, (-1, 3))
    n = ref.shape[0]
    cols = [np.ascontiguousarray(ref[:, k]) for k in range(3)] + \
           [np.ascontiguousarray(nbr[:, k]) for k in range(3)]
    out = [np.zeros(n, np.float32) for _ in range(3)]
    rid = np.ascontiguousarray(ref_id, np.int32) if ref_id is not None else None
    nid = np.ascontiguousarray(nbr_id, np.int32) if nbr_id is not None else None
    _check(lib().md_rl_eval_pairs(table.handle,
                                  rid.ctypes.data if rid is not None else None,
                                  nid.ctypes.data if nid is not None else None,
                                  *[_fp(c) for c in cols], *[_fp(o) for o in out],
                                  n, rc2, flags))
    return np.stack(out, axis=1)


def eval_pairs_ex(table, ref, nbr, rc2, flags=0, p_a=None, p_b=None, p_qq=None):
    """Pair batch with per-pair multipliers; returns (F, Fxyz)."""
    ref = _f32(ref, (-1, 3))
    nbr = _f32(nbr, (-1, 3))
    n = ref.shape[0]
    fs = np.zeros(n, np.float32)
    fx = np.zeros((n, 3), np.float32)
    arrs = [None if a is None else _f32(a) for a in (p_a, p_b, p_qq)]
    _check(lib().md_rl_eval_pairs_ex(table.handle, _fp(ref), _fp(nbr),
                                     *[None if a is None else _fp(a) for a in arrs],
                                     _fp(fs), _fp(fx), n, rc2, flags))
    return fs, fx


def eval_pairs_ex_noalloc(table, ref_ptr, nbr_ptr, npairs, rc2, fxyz_ptr, fs_ptr=None, flags=0,
                          p_a_ptr=None, p_b_ptr=None, p_qq_ptr=None, stream=None):
    """md_rl_eval_pairs_ex_noalloc on device pointers (integers, e.g. torch data_ptr());
    enqueued on stream (a hipStream_t as an integer; None: the null stream)."""
    _check(lib().md_rl_eval_pairs_ex_noalloc(table.handle, ref_ptr, nbr_ptr, p_a_ptr, p_b_ptr,
                                             p_qq_ptr, fs_ptr, fxyz_ptr, npairs, rc2, flags, stream))


# ---------------------------------------------------------------------------------- systems
class SyntheticSystem:
    """A seeded benchmark system from md_rl_make_system (owned numpy copies)."""

    def __init__(self, which, index=0, seed=2019):
        d = SystemDesc()
        _check(lib().md_rl_make_system(ctypes.byref(d), which, index, seed))
        try:
            n = d.n
            self.n = n
            self.ntypes = d.ntypes
            self.box = np.array(d.box[:], np.float32)
            self.rc = float(d.rc)
            self.dt = float(d.dt)
            self.coulomb_k = float(d.coulomb_k)
            self.particle = np.ctypeslib.as_array(d.particle, shape=(n, 11)).copy()
            self.type = np.ctypeslib.as_array(d.type, shape=(n,)).copy()
            self.charge = np.ctypeslib.as_array(d.charge, shape=(n,)).copy()
            self.eps = np.array(d.eps[:self.ntypes], np.float32)
            self.sigma = np.array(d.sigma[:self.ntypes], np.float32)
            self.mass = np.array(d.mass[:self.ntypes], np.float32)
            self.nterms = int(d.nterms)
            if self.nterms:
                self.term_atoms = np.ctypeslib.as_array(d.term_atoms, shape=(self.nterms, 4)).copy()
                self.term_params = np.ctypeslib.as_array(d.term_params, shape=(self.nterms, 4)).copy()
            else:
                self.term_atoms = np.zeros((0, 4), np.int32)
                self.term_params = np.zeros((0, 4), np.float32)
            self.excl = np.ctypeslib.as_array(d.excl, shape=(n,)).copy() if d.excl else None
        finally:
            lib().md_rl_free_system(ctypes.byref(d))


class System:
    """One periodic system on the GPU (md_rl_create ... md_rl_destroy)."""

    def __init__(self, n, box, rc, eps, sigma, mass, dt=1.0, coulomb_k=0.0, flags=0,
                 ncell=(0, 0, 0), table=None):
        self._eps = _f32(eps)
        self._sig = _f32(sigma)
        self._mass = _f32(mass)
        p = Params()
        p.n = n
        p.box[:] = [float(b) for b in box]
        p.rc = rc
        p.ncell[:] = list(ncell)
        p.dt = dt
        p.ntypes = len(self._eps)
        p.eps, p.sigma, p.mass = _fp(self._eps), _fp(self._sig), _fp(self._mass)
        p.coulomb_k = coulomb_k
        p.flags = flags
        p.table = table.handle if table is not None else None
        self._table = table
        self.n = n
        self.flags = flags
        h = ctypes.c_void_p()
        _check(lib().md_rl_create(ctypes.byref(h), ctypes.byref(p)))
        self.handle = h

    @classmethod
    def from_synthetic(cls, s, flags=0, table=None):
        sysm = cls(s.n, s.box, s.rc, s.eps, s.sigma, s.mass, dt=s.dt, coulomb_k=s.coulomb_k,
                   flags=flags, table=table)
        sysm.set_types(s.type, s.charge if s.coulomb_k else None)
        if getattr(s, "nterms", 0):
            sysm.set_bonded(s.term_atoms, s.term_params)
        if getattr(s, "excl", None) is not None:
            sysm.set_exclusions(s.excl)
        sysm.set_particles(s.particle)
        return sysm

    def set_bonded(self, atoms, params, sys=0):
        """Bonded terms (rows of 4 input indices, -1 padded; rows of 4 parameters)."""
        a = np.ascontiguousarray(atoms, np.int32).reshape(-1, 4)
        p = _f32(params, (a.shape[0], 4))
        _check(lib().md_rl_set_bonded(self.handle, sys, a.shape[0],
                                      a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _fp(p)))

    def set_exclusions(self, mask, sys=0):
        """Exclusion windows (bit d of mask[i] excludes the pair (i, i + d - 16)); None
        removes them."""
        if mask is None:
            _check(lib().md_rl_set_exclusions(self.handle, sys, None))
            return
        m = np.ascontiguousarray(mask, np.uint32)
        _check(lib().md_rl_set_exclusions(self.handle, sys, m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))))

    def set_long_range(self, fn):
        """Summation hook (md_rl_set_long_range): fn(view, stream) -> int is called with an
        LRView after the range-limited and bonded forces of every evaluation and adds
        forces into view.force on that stream (see host_provider for a numpy one); None
        removes it."""
        old = getattr(self, "_lr", None)
        if fn is None:
            _check(lib().md_rl_set_long_range(self.handle, LR_FN(), None))
            self._lr = self._lr_old = None
            return

        def trampoline(user, view, stream):
            try:
                return int(fn(view.contents, stream) or 0)
            except Exception:
                import traceback
                traceback.print_exc()
                return -1

        # the library keeps the hook even when its first evaluation fails, so the callback
        # object stays referenced from here until it is replaced
        self._lr_old = old
        self._lr = LR_FN(trampoline)
        _check(lib().md_rl_set_long_range(self.handle, self._lr, None))

    def set_particles(self, particle):
        a = _f32(particle, (self.n, 11))
        _check(lib().md_rl_set_particles(self.handle, _fp(a), self.n))

    def set_types(self, types, charge=None):
        t = np.ascontiguousarray(types, np.int32)
        q = None if charge is None else _f32(charge)
        _check(lib().md_rl_set_types(self.handle, t.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                     None if q is None else _fp(q), self.n))

    def force_eval(self):
        _check(lib().md_rl_force_eval(self.handle))

    def step(self, nsteps=1):
        _check(lib().md_rl_step(self.handle, nsteps))

    def step_async(self, nsteps=1):
        _check(lib().md_rl_step_async(self.handle, nsteps, None))

    def sync(self):
        _check(lib().md_rl_sync(self.handle))

    def minimize(self, iterations, max_disp=0.05, gamma=1e-4):
        _check(lib().md_rl_minimize(self.handle, iterations, max_disp, gamma))

    def particles(self):
        out = np.zeros((self.n, 11), np.float32)
        _check(lib().md_rl_get_particles(self.handle, _fp(out), self.n))
        return out

    def forces_fixed(self):
        out = np.zeros((self.n, 3), np.int64)
        _check(lib().md_rl_get_forces_fixed(
            self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), self.n))
        return out

    def energy(self, virial=True):
        """md_rl_energy: (potential, kinetic, virial 3x3 or None); virial=False skips the
        pair virial (a cheaper energy pass)."""
        pe, ke = ctypes.c_double(), ctypes.c_double()
        vir = (ctypes.c_double * 9)() if virial else None
        _check(lib().md_rl_energy(self.handle, ctypes.byref(pe), ctypes.byref(ke), vir))
        return pe.value, ke.value, (np.array(vir[:]).reshape(3, 3) if virial else None)

    def energy_terms(self, sys=-1):
        """md_rl_get_energy_terms: kinetic, potential and its parts, virial, and the sum of
        the in-cutoff pairs' energies at rc (potential - cutoff_shift is the shifted
        potential that velocity Verlet conserves with the kinetic energy)."""
        t = EnergyTerms()
        _check(lib().md_rl_get_energy_terms(self.handle, sys, ctypes.byref(t)))
        return {"kinetic": t.kinetic, "potential": t.potential, "range_limited": t.range_limited,
                "bonded": t.bonded, "long_range": t.long_range,
                "virial": np.array(t.virial[:]).reshape(3, 3), "cutoff_shift": t.cutoff_shift}

    def stats(self):
        s = Stats()
        _check(lib().md_rl_get_stats(self.handle, ctypes.byref(s)))
        return s

    def set_timing(self, on=True):
        _check(lib().md_rl_set_timing(self.handle, 1 if on else 0))

    def set_graph(self, on=True):
        _check(lib().md_rl_set_graph(self.handle, 1 if on else 0))

    def prepare_graphs(self, nsteps):
        """Capture (without running) the graphs that step(nsteps) would replay."""
        _check(lib().md_rl_graph_prepare(self.handle, nsteps))

    def set_schedule(self, apw=64, tile=640, groups=1, persistent=1):
        sc = Schedule(apw, tile, groups, persistent)
        _check(lib().md_rl_set_schedule(self.handle, ctypes.byref(sc)))

    def schedule(self):
        sc = Schedule()
        _check(lib().md_rl_get_schedule(self.handle, ctypes.byref(sc)))
        return sc

    def autotune(self, reps=3, cache=False):
        """Sweeps the schedules (md_rl_autotune) and keeps the fastest: (schedule, ms).
        cache=True goes through md_rl_autotune_cached and returns (schedule, ms, hit)."""
        sc = Schedule()
        ms = ctypes.c_double()
        if cache:
            hit = ctypes.c_int()
            _check(lib().md_rl_autotune_cached(self.handle, reps, ctypes.byref(sc), ctypes.byref(ms),
                                               ctypes.byref(hit)))
            return sc, ms.value, bool(hit.value)
        _check(lib().md_rl_autotune(self.handle, reps, ctypes.byref(sc), ctypes.byref(ms)))
        return sc, ms.value

    def close(self):
        if getattr(self, "handle", None):
            lib().md_rl_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch(System):
    """Independent periodic systems stepped together (md_rl_create_batch)."""

    def __init__(self, sizes, boxes, rc, eps, sigma, mass, dt=1.0, coulomb_k=0.0, flags=0,
                 table=None):
        self._eps = _f32(eps)
        self._sig = _f32(sigma)
        self._mass = _f32(mass)
        p = Params()
        p.rc = rc
        p.dt = dt
        p.ntypes = len(self._eps)
        p.eps, p.sigma, p.mass = _fp(self._eps), _fp(self._sig), _fp(self._mass)
        p.coulomb_k = coulomb_k
        p.flags = flags
        p.table = table.handle if table is not None else None
        self._table = table
        self.sizes = np.ascontiguousarray(sizes, np.int32)
        b = _f32(boxes, (-1, 3))
        self.nsys = len(self.sizes)
        self.n = int(self.sizes.sum())
        self.flags = flags
        h = ctypes.c_void_p()
        _check(lib().md_rl_create_batch(ctypes.byref(h), ctypes.byref(p), self.nsys,
                                        self.sizes.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                        _fp(b)))
        self.handle = h

    @classmethod
    def from_synthetic(cls, systems, flags=0, table=None):
        s0 = systems[0]
        b = cls([s.n for s in systems], [s.box for s in systems], s0.rc, s0.eps, s0.sigma,
                s0.mass, dt=s0.dt, coulomb_k=s0.coulomb_k, flags=flags, table=table)
        for k, s in enumerate(systems):
            b.set_system(k, s.particle, s.type, s.charge if s.coulomb_k else None)
            if getattr(s, "nterms", 0):
                b.set_bonded(s.term_atoms, s.term_params, sys=k)
            if getattr(s, "excl", None) is not None:
                b.set_exclusions(s.excl, sys=k)
        b.commit()
        return b

    def set_system(self, sys, particle=None, types=None, charge=None):
        n = int(self.sizes[sys])
        a = None if particle is None else _f32(particle, (n, 11))
        t = None if types is None else np.ascontiguousarray(types, np.int32)
        q = None if charge is None else _f32(charge)
        _check(lib().md_rl_set_system(self.handle, sys, None if a is None else _fp(a),
                                      None if t is None else t.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                      None if q is None else _fp(q)))

    def commit(self):
        _check(lib().md_rl_commit(self.handle))

    def system_particles(self, sys):
        out = np.zeros((int(self.sizes[sys]), 11), np.float32)
        _check(lib().md_rl_get_system(self.handle, sys, _fp(out)))
        return out

    def system_forces_fixed(self, sys):
        out = np.zeros((int(self.sizes[sys]), 3), np.int64)
        _check(lib().md_rl_get_system_forces_fixed(self.handle, sys, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return out

    def system_energy(self, sys):
        pe, ke = ctypes.c_double(), ctypes.c_double()
        vir = (ctypes.c_double * 9)()
        _check(lib().md_rl_system_energy(self.handle, sys, ctypes.byref(pe), ctypes.byref(ke), vir))
        return pe.value, ke.value, np.array(vir[:]).reshape(3, 3)


def check_exclusions(mask):
    """Host-only check of exclusion windows (md_rl_check_exclusions); raises MdrlError."""
    m = np.ascontiguousarray(mask, np.uint32)
    _check(lib().md_rl_check_exclusions(m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), m.shape[0]))


# ------------------------------------------------------------ cell-memory initialisation files
FILE_CRLF = 1


def _path(p):
    return os.fsencode(os.fspath(p))


def _count(code):
    if code < 0:
        raise MdrlError(code, lib().md_rl_last_error().decode())
    return code


def read_cell_file(path, cap=4096):
    """One cell_ini_file_X_Y_Z.{mif,hex}: the particles' x, y, z (count x 3, float32)."""
    xyz = np.zeros((cap, 3), np.float32)
    n = _count(lib().md_rl_cell_file_read(_path(path), _fp(xyz), cap))
    return xyz[:n].copy()


def write_cell_file(path, xyz, depth, crlf=True):
    """Writes particles (count x 3: x, y, z) as a cell file of depth words (.mif or .hex
    by extension), as the reference's generator and Quartus' conversion write it."""
    a = _f32(xyz, (-1, 3))
    _check(lib().md_rl_cell_file_write(_path(path), _fp(a), a.shape[0], depth, FILE_CRLF if crlf else 0))


def read_boundary_file(path, cap=4096):
    b = np.zeros(cap, np.float32)
    n = _count(lib().md_rl_boundary_file_read(_path(path), _fp(b), cap))
    return b[:n].copy()


def write_boundary_file(path, edge, depth, crlf=True):
    _check(lib().md_rl_boundary_file_write(_path(path), edge, depth, FILE_CRLF if crlf else 0))


def read_cell_dir(directory, lo, hi, prefix=None, cap=1 << 22):
    """Every cell file <prefix>X_Y_Z of a directory for lo <= (X, Y, Z) <= hi (from 1):
    particle rows (n x 11, positions in cols 0-2) in the generator's order and each
    particle's (X, Y, Z)."""
    lo_ = (ctypes.c_int * 3)(*lo)
    hi_ = (ctypes.c_int * 3)(*hi)
    part = np.zeros((cap, 11), np.float32)
    cell = np.zeros((cap, 3), np.int32)
    n = _count(lib().md_rl_cell_dir_read(_path(directory), None if prefix is None else prefix.encode(),
                                         lo_, hi_, _fp(part), cell.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                         cap))
    return part[:n].copy(), cell[:n].copy()


def bin_cells(xyz, edge, rule="floor"):
    """Cells (from 1) of the reference's generators: floor(x / edge) + 1
    (M/GenInputCellList_LJArgon.m:101-136) or ceil(x / edge) (M/GenInputCellList_ApoA1.m:
    99-134), FP32 division, and cell 1 for a coordinate of exactly 0."""
    x = _f32(xyz, (-1, 3))
    q = x / np.float32(edge)
    c = (np.floor(q) + 1) if rule == "floor" else np.ceil(q)
    return np.where(x != 0, c, 1).astype(np.int64)


def write_cell_dir(directory, xyz, edge, lo, hi, depth, rule="floor", ext=".hex", crlf=True,
                   prefix="cell_ini_file_"):
    """Bins particles (n x 3, already in the cell frame) as the reference's generators do and
    writes one cell file per cell lo <= (X, Y, Z) <= hi, particles in input order; returns
    the number of particles that fell in no written cell."""
    x = _f32(xyz, (-1, 3))
    cells = bin_cells(x, edge, rule)
    written = 0
    for cx in range(lo[0], hi[0] + 1):
        for cy in range(lo[1], hi[1] + 1):
            for cz in range(lo[2], hi[2] + 1):
                sel = np.all(cells == (cx, cy, cz), axis=1)
                write_cell_file(os.path.join(directory, "%s%d_%d_%d%s" % (prefix, cx, cy, cz, ext)),
                                x[sel], depth, crlf)
                written += int(sel.sum())
    return x.shape[0] - written


# ------------------------------------------------------------------ summation-hook helpers
_hip = None


def hip():
    """The HIP runtime (libamdhip64) through ctypes, for providers written in Python."""
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                H = ctypes.CDLL(name)
                break
            except OSError:
                H = None
        if H is None:
            raise ImportError("libamdhip64.so not found")
        P, Z = ctypes.c_void_p, ctypes.c_size_t
        H.hipMalloc.argtypes = [ctypes.POINTER(P), Z]
        H.hipFree.argtypes = [P]
        H.hipMemcpy.argtypes = [P, P, Z, ctypes.c_int]
        H.hipMemcpyAsync.argtypes = [P, P, Z, ctypes.c_int, P]
        H.hipStreamSynchronize.argtypes = [P]
        H.hipStreamIsCapturing.argtypes = [P, ctypes.POINTER(ctypes.c_int)]
        _hip = H
    return _hip


H2D, D2H, D2D = 1, 2, 3


def _hipcheck(e, what):
    if e != 0:
        raise RuntimeError("%s failed: HIP error %d" % (what, e))


class DeviceBuffer:
    """hipMalloc'd bytes holding a numpy array (for device-only providers)."""

    def __init__(self, array):
        a = np.ascontiguousarray(array)
        self.nbytes = a.nbytes
        self.ptr = ctypes.c_void_p()
        _hipcheck(hip().hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 1)), "hipMalloc")
        _hipcheck(hip().hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), self.nbytes, H2D), "hipMemcpy")

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def host_provider(fn):
    """Wraps fn(xq, box, atom_off, want_energy) -> (force, energy) computed in numpy
    (xq: n x 4 float32 {x, y, z, q} in input order; box: nsys x 3; atom_off: nsys + 1;
    force: n x 3 kJ/mol/A; energy: nsys kJ/mol or None) as a summation-hook provider.  It
    waits for the stream and copies through host memory, so it cannot run inside a
    captured graph: with graphs on it returns an error."""

    def provider(view, stream):
        H = hip()
        cap = ctypes.c_int(0)
        _hipcheck(H.hipStreamIsCapturing(stream, ctypes.byref(cap)), "hipStreamIsCapturing")
        if cap.value:
            return -2
        n, nsys = view.n, view.nsys
        _hipcheck(H.hipStreamSynchronize(stream), "hipStreamSynchronize")
        xq = np.zeros((n, 4), np.float32)
        if n:
            _hipcheck(H.hipMemcpy(xq.ctypes.data_as(ctypes.c_void_p), view.xq, xq.nbytes, D2H), "hipMemcpy")
        box = np.ctypeslib.as_array(view.box, (nsys, 3)).copy()
        off = np.ctypeslib.as_array(view.atom_off, (nsys + 1,)).copy()
        force, energy = fn(xq, box, off, bool(view.want_energy))
        f = np.ascontiguousarray(force, np.float32).reshape(n, 3)
        if n:
            _hipcheck(H.hipMemcpy(view.force, f.ctypes.data_as(ctypes.c_void_p), f.nbytes, H2D), "hipMemcpy")
        if energy is not None:
            e = np.ascontiguousarray(energy, np.float64).reshape(nsys)
            _hipcheck(H.hipMemcpy(view.energy, e.ctypes.data_as(ctypes.c_void_p), e.nbytes, H2D), "hipMemcpy")
        return 0

    return provider


def device_provider(force_buf, energy_buf=None):
    """A provider that copies fixed device buffers (n x 3 float32 forces, nsys float64
    energies) into the view on the stream: capturable, so it also runs under graphs."""

    def provider(view, stream):
        H = hip()
        if force_buf.nbytes:
            _hipcheck(H.hipMemcpyAsync(view.force, force_buf.ptr, force_buf.nbytes, D2D, stream), "hipMemcpyAsync")
        if energy_buf is not None and energy_buf.nbytes:
            _hipcheck(H.hipMemcpyAsync(view.energy, energy_buf.ptr, energy_buf.nbytes, D2D, stream), "hipMemcpyAsync")
        return 0

    return provider


# ---------------------------------------------------------------------------------- C model
def stream_copy(dst_ptr, src_ptr, nbytes):
    """md_rl_stream_copy on raw device addresses (a kernel of known traffic)."""
    _check(lib().md_rl_stream_copy(ctypes.c_void_p(dst_ptr), ctypes.c_void_p(src_ptr),
                                   ctypes.c_size_t(nbytes)))


def cmodel_default():
    cfg = CModelCfg()
    lib().md_rl_cmodel_default(ctypes.byref(cfg))
    return cfg


def cmodel_run(particle, cfg=None, iterations=None, print_lines=False):
    """Runs the reference C model's iterations on the GPU; returns (particle, energies)."""
    if cfg is None:
        cfg = cmodel_default()
    a = _f32(particle, (-1, 11)).copy()
    cfg.n = a.shape[0]
    if iterations is not None:
        cfg.iterations = iterations
    cfg.print = 1 if print_lines else 0
    e = np.zeros(2 * max(cfg.iterations, 1), np.float32)
    _check(lib().md_rl_cmodel_run(ctypes.byref(cfg), _fp(a), _fp(e)))
    return a, e[:2 * cfg.iterations].reshape(-1, 2)


def MD_Evaluation_Model():
    return lib().MD_Evaluation_Model()
