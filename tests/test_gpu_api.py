"""GPU: the smaller C-ABI entry points -- batch geometry queries and the streaming copy
used to calibrate the traffic counters."""
import ctypes
import math

import numpy as np
import pytest

from sc2019_md_amd import mdrl

pytestmark = pytest.mark.gpu


def coarse_grid(n, box, rc, given=(0, 0, 0)):
    """The grid of MDRL_COARSE_CELLS, restated: floor(L / rc) per axis (or the count given),
    then one cell fewer on the axis with the most cells among those not given and above one
    cell (ties: the shorter edge, then the lower axis) while the average cell holds fewer
    than 128 atoms."""
    nc = [g if g > 0 else max(1, int(math.floor(float(x) / rc))) for g, x in zip(given, box)]
    while n < 128 * nc[0] * nc[1] * nc[2]:
        free = [d for d in range(3) if given[d] <= 0 and nc[d] > 1]
        if not free:
            break
        dm = free[0]
        for d in free[1:]:
            if nc[d] > nc[dm] or (nc[d] == nc[dm] and float(box[d]) / nc[d] < float(box[dm]) / nc[dm]):
                dm = d
        nc[dm] -= 1
    return nc


def test_coarse_cells_grid():
    """With MDRL_COARSE_CELLS the bonded workload's replicas (3,072 beads, L = 45.3 A,
    rc = 10 A: 4^3 cells of 48) get a grid of 18 cells (about 171 beads each); without it,
    and for a single system given an explicit grid, the grid is as asked."""
    systems = [mdrl.SyntheticSystem(5, k) for k in (0, 1)]
    b = mdrl.Batch.from_synthetic(systems, flags=mdrl.COARSE_CELLS)
    L = mdrl.lib()
    for k, s in enumerate(systems):
        nc = (ctypes.c_int * 3)()
        assert L.md_rl_system_size(b.handle, k, None, None, nc) == 0
        assert [max(1, int(math.floor(float(x) / s.rc))) for x in s.box] == [4, 4, 4]
        assert list(nc) == coarse_grid(s.n, s.box, s.rc)
        assert nc[0] * nc[1] * nc[2] == 18
    b.close()
    b = mdrl.Batch.from_synthetic(systems)
    nc = (ctypes.c_int * 3)()
    assert L.md_rl_system_size(b.handle, 0, None, None, nc) == 0
    assert list(nc) == [4, 4, 4]
    b.close()
    s = systems[0]
    for given in ((4, 4, 4), (4, 0, 0), (0, 0, 0)):
        g = mdrl.System(s.n, s.box, s.rc, s.eps, s.sigma, s.mass, flags=mdrl.COARSE_CELLS, ncell=given)
        nc = (ctypes.c_int * 3)()
        assert L.md_rl_system_size(g.handle, 0, None, None, nc) == 0
        assert list(nc) == coarse_grid(s.n, s.box, s.rc, given), given
        g.close()
    assert coarse_grid(s.n, s.box, s.rc, (4, 4, 4)) == [4, 4, 4]
    assert coarse_grid(s.n, s.box, s.rc, (4, 0, 0)) == [4, 2, 3]


def test_batch_geometry_queries():
    """md_rl_num_systems and md_rl_system_size report each system's atom count, box and
    cell grid (floor(L / rc) cells per axis)."""
    systems = [mdrl.SyntheticSystem(4, k) for k in (0, 3, 8, 21)]
    b = mdrl.Batch.from_synthetic(systems)
    L = mdrl.lib()
    assert L.md_rl_num_systems(b.handle) == len(systems)
    for k, s in enumerate(systems):
        n = ctypes.c_int32()
        box = (ctypes.c_float * 3)()
        nc = (ctypes.c_int * 3)()
        assert L.md_rl_system_size(b.handle, k, ctypes.byref(n), box, nc) == 0
        assert n.value == s.n
        assert list(box) == [float(x) for x in s.box]
        assert list(nc) == [max(1, int(math.floor(float(x) / s.rc))) for x in s.box]
    assert L.md_rl_system_size(b.handle, len(systems), None, None, None) == -1
    b.close()


def test_stream_copy_moves_every_byte():
    """md_rl_stream_copy copies device memory byte for byte; sizes that are not a multiple
    of 16 bytes are refused."""
    a = np.random.default_rng(3).integers(0, 2 ** 32, size=(1 << 20) + 4, dtype=np.uint32)
    src = mdrl.DeviceBuffer(a)
    dst = mdrl.DeviceBuffer(np.zeros_like(a))
    mdrl.stream_copy(dst.ptr.value, src.ptr.value, a.nbytes)
    out = np.zeros_like(a)
    assert mdrl.hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), dst.ptr, a.nbytes, mdrl.D2H) == 0
    assert np.array_equal(out, a)
    with pytest.raises(mdrl.MdrlError):
        mdrl.stream_copy(dst.ptr.value, src.ptr.value, 20)
    src.free()
    dst.free()
