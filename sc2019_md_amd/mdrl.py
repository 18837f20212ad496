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
ACC_FP32, EVAL_ANALYTIC, EVAL_PARITY, KICK_DRIFT, NO_COULOMB = 1, 2, 4, 8, 16
TERM_14, TERM_8, TERM_3, TERM_12, TERM_6, TERM_1 = range(6)

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
        L.md_rl_table_free.restype = None
        L.md_rl_eval_pairs.argtypes = [P, P, P, F, F, F, F, F, F, F, F, F, ctypes.c_int64,
                                       ctypes.c_float, I]
        L.md_rl_eval_pairs_ex.argtypes = [P, F, F, F, F, F, F, F, ctypes.c_int64,
                                          ctypes.c_float, I]
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
        L.md_rl_commit.argtypes = [P]
        L.md_rl_get_system.argtypes = [P, I, F]
        L.md_rl_get_system_forces_fixed.argtypes = [P, I, ctypes.POINTER(ctypes.c_int64)]
        L.md_rl_system_energy.argtypes = [P, I, ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_double)]
        L.md_rl_set_graph.argtypes = [P, I]
        L.md_rl_set_schedule.argtypes = [P, ctypes.POINTER(Schedule)]
        L.md_rl_get_schedule.argtypes = [P, ctypes.POINTER(Schedule)]
        L.md_rl_autotune.argtypes = [P, I, ctypes.POINTER(Schedule), ctypes.POINTER(ctypes.c_double)]
        L.md_rl_cmodel_default.argtypes = [ctypes.POINTER(CModelCfg)]
        L.md_rl_cmodel_default.restype = None
        L.md_rl_cmodel_run.argtypes = [ctypes.POINTER(CModelCfg), F, F]
        L.md_rl_make_system.argtypes = [ctypes.POINTER(SystemDesc), I, I, ctypes.c_uint64]
        L.md_rl_free_system.argtypes = [ctypes.POINTER(SystemDesc)]
        L.md_rl_free_system.restype = None
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

    def energy(self):
        pe, ke = ctypes.c_double(), ctypes.c_double()
        vir = (ctypes.c_double * 9)()
        _check(lib().md_rl_energy(self.handle, ctypes.byref(pe), ctypes.byref(ke), vir))
        return pe.value, ke.value, np.array(vir[:]).reshape(3, 3)

    def stats(self):
        s = Stats()
        _check(lib().md_rl_get_stats(self.handle, ctypes.byref(s)))
        return s

    def set_timing(self, on=True):
        _check(lib().md_rl_set_timing(self.handle, 1 if on else 0))

    def set_graph(self, on=True):
        _check(lib().md_rl_set_graph(self.handle, 1 if on else 0))

    def set_schedule(self, apw=64, tile=256, groups=1, persistent=1):
        sc = Schedule(apw, tile, groups, persistent)
        _check(lib().md_rl_set_schedule(self.handle, ctypes.byref(sc)))

    def schedule(self):
        sc = Schedule()
        _check(lib().md_rl_get_schedule(self.handle, ctypes.byref(sc)))
        return sc

    def autotune(self, reps=3):
        sc = Schedule()
        ms = ctypes.c_double()
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
