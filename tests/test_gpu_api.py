"""GPU: the smaller C-ABI entry points -- batch geometry queries and the streaming copy
used to calibrate the traffic counters."""
import ctypes
import math

import numpy as np
import pytest

from sc2019_md_amd import mdrl

pytestmark = pytest.mark.gpu


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
