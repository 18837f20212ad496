"""GPU: the summation hook (md_rl_set_long_range) -- forces from a provider outside the
range-limited path summed into the same int64 accumulators before Motion_Update.

Bars: the {x, y, z, q} records handed to the provider equal the oracle's wrapped input
positions and the input charges bit for bit; forces are the oracle's range-limited forces
plus lr_fix of the provider's forces, bit-exact (ties to even, saturation flagged);
velocity-Verlet trajectories with a position-dependent provider bit-exact against the
oracle; a device-only provider captured in graphs gives the same bits as the oracle;
batches hand each system its own slice; the provider's energy joins the potential energy.
"""

import numpy as np
import pytest

from oracle import oracle as O
from sc2019_md_amd import mdrl

from tests.test_oracle_golden import rl_from_synthetic  # noqa: E402

pytestmark = pytest.mark.gpu

FIELD = np.array([0.8, -0.3, 0.45], np.float32)    # kJ/mol/A/e
K_TETHER = np.float32(0.6)                          # kJ/mol/A^2


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def apoa1():
    t = mdrl.Table.load()
    return {k: t.get(k) for k in range(6)}


def tether_forces(xq, anchor, box):
    """Harmonic tether to anchor (minimum image) plus a uniform field on the charges, in
    FP32; energy 0.5 k |d|^2 summed in FP64."""
    b = box.astype(np.float32)
    d = xq[:, :3] - anchor
    d = d - b * np.rint(d / b)
    f = -K_TETHER * d + xq[:, 3:4] * FIELD
    e = 0.5 * float(K_TETHER) * float(np.sum(d.astype(np.float64) ** 2))
    return f.astype(np.float32), e


class Recorder:
    """Host provider of tether forces, keeping the records it was handed."""

    def __init__(self, anchors):
        self.anchors = anchors            # per system, input order
        self.calls = []

    def __call__(self, xq, box, off, want_energy):
        f = np.zeros((xq.shape[0], 3), np.float32)
        e = np.zeros(len(off) - 1)
        for k in range(len(off) - 1):
            sl = slice(off[k], off[k + 1])
            f[sl], e[k] = tether_forces(xq[sl], self.anchors[k], box[k])
        self.calls.append((xq.copy(), box.copy(), off.copy(), want_energy))
        return f, e


def test_records_and_forces_bit_exact(apoa1):
    s = mdrl.SyntheticSystem(4, 3)
    say("long range: replica 3, n=%d, host tether provider" % s.n)
    g = mdrl.System.from_synthetic(s)
    rec = Recorder([s.particle[:, :3].copy()])
    g.set_long_range(mdrl.host_provider(rec))
    got = g.forces_fixed()
    g.close()
    rl = rl_from_synthetic(s, apoa1)
    xq, box, off, we = rec.calls[-1]
    assert not we and off.tolist() == [0, s.n] and np.array_equal(box[0], s.box.astype(np.float32))
    assert np.array_equal(xq[:, :3].view(np.uint32), rl.pos4[:, :3].view(np.uint32))
    assert np.array_equal(xq[:, 3].view(np.uint32), s.charge.astype(np.float32).view(np.uint32))
    fix, _ = rl.forces_fix()
    f, _ = tether_forces(xq, s.particle[:, :3], s.box)
    want = fix + O.lr_fix(f)
    bad = np.nonzero(np.any(got != want, axis=1))[0]
    assert bad.size == 0, "atoms differ: %s" % bad[:10]


def test_conversion_ties_and_saturation(apoa1):
    """A device provider of crafted forces: ties round to even, +-1e15 saturates at 2^62
    and raises the saturation flag."""
    s = mdrl.SyntheticSystem(4, 0)
    f = np.zeros((s.n, 3), np.float32)
    f[1] = [2.0 ** -34, 3 * 2.0 ** -34, -(2.0 ** -34)]
    f[2] = [5 * 2.0 ** -35, 1.0, -0.1]
    f[3:] = np.random.default_rng(3).normal(scale=50.0, size=(s.n - 3, 3)).astype(np.float32)
    buf = mdrl.DeviceBuffer(f)
    g = mdrl.System.from_synthetic(s)
    g.set_long_range(mdrl.device_provider(buf))
    got = g.forces_fixed()
    assert g.stats().flags & 1 == 0
    rl = rl_from_synthetic(s, apoa1)
    fix, _ = rl.forces_fix()
    assert np.array_equal(got, fix + O.lr_fix(f))
    f[0] = [1e15, -1e15, 0.0]
    buf2 = mdrl.DeviceBuffer(f)
    g.set_long_range(mdrl.device_provider(buf2))
    got = g.forces_fixed()
    flags = g.stats().flags
    g.close()
    assert flags & 1
    assert np.array_equal(got, fix + O.lr_fix(f))


def test_fp32_accumulation_mode():
    """With FP32 atomics the provider's forces are added as f * 10 in FP32: the forces
    stay within 1e-5 (relative L2) of the int64 run with the same provider."""
    s = mdrl.SyntheticSystem(4, 3)
    f = (s.charge.astype(np.float32)[:, None] * FIELD).astype(np.float32)
    buf = mdrl.DeviceBuffer(f)
    out = []
    for flags, hook in ((0, True), (mdrl.ACC_FP32, True), (mdrl.ACC_FP32, False)):
        g = mdrl.System.from_synthetic(s, flags=flags)
        if hook:
            g.set_long_range(mdrl.device_provider(buf))
        out.append(g.forces_fixed().astype(np.float64))
        g.close()
    assert np.linalg.norm(out[1] - out[0]) / np.linalg.norm(out[0]) < 1e-5
    # the provider's share in FP32 mode: the field forces, up to the FP32 atomics' order
    lr = O.lr_fix(f).astype(np.float64)
    assert np.linalg.norm((out[1] - out[2]) - lr) / np.linalg.norm(lr) < 1e-2


def test_steps_bit_exact(apoa1):
    s = mdrl.SyntheticSystem(4, 11)
    nsteps = 6
    say("long range: replica 11, n=%d, %d steps with a tether provider" % (s.n, nsteps))
    anchor = s.particle[:, :3].copy()
    g = mdrl.System.from_synthetic(s)
    g.set_long_range(mdrl.host_provider(Recorder([anchor])))
    g.step(nsteps)
    got = g.particles()
    gfix = g.forces_fixed()
    g.close()
    rl = rl_from_synthetic(s, apoa1)
    hk = O.half_kick(s.mass, s.dt)
    pos4, vel = rl.pos4.copy(), s.particle[:, 3:6].copy()
    q = s.charge.astype(np.float32)[:, None]

    def total(p4):
        rl.pos4[:] = p4
        fx, _ = rl.forces_fix()
        f, _ = tether_forces(np.hstack([p4[:, :3], q]), anchor, s.box)
        return fx + O.lr_fix(f)

    fix = total(pos4)
    for k in range(nsteps):
        pos4, vel, _ = O.motion_update(pos4, vel, fix, s.type, hk, s.mass, s.dt, s.box, k == 0)
        fix = total(pos4)
    assert np.array_equal(got[:, :3].view(np.uint32), pos4[:, :3].view(np.uint32))
    assert np.array_equal(gfix, fix)


def test_graph_device_provider_bit_exact(apoa1):
    """18 steps with graphs on (an eager first step, one 16-step graph, one eager step):
    the provider's copies are captured with the step and replayed."""
    s = mdrl.SyntheticSystem(4, 40)
    nsteps = 18
    say("long range: replica 40, n=%d, %d graph steps, device provider" % (s.n, nsteps))
    f = (s.charge.astype(np.float32)[:, None] * FIELD).astype(np.float32)
    buf = mdrl.DeviceBuffer(f)
    got = {}
    for graph in (True, False):
        g = mdrl.System.from_synthetic(s)
        g.set_long_range(mdrl.device_provider(buf))
        g.set_graph(graph)
        g.step(nsteps)
        got[graph] = g.particles()
        g.close()
    assert np.array_equal(got[True].view(np.uint32), got[False].view(np.uint32))
    rl = rl_from_synthetic(s, apoa1)
    hk = O.half_kick(s.mass, s.dt)
    pos4, vel = rl.pos4.copy(), s.particle[:, 3:6].copy()
    lr = O.lr_fix(f)

    def total(p4):
        rl.pos4[:] = p4
        return rl.forces_fix()[0] + lr

    fix = total(pos4)
    for k in range(nsteps):
        pos4, vel, _ = O.motion_update(pos4, vel, fix, s.type, hk, s.mass, s.dt, s.box, k == 0)
        fix = total(pos4)
    assert np.array_equal(got[True][:, :3].view(np.uint32), pos4[:, :3].view(np.uint32))


def test_batch_slices_and_energy():
    systems = [mdrl.SyntheticSystem(4, k) for k in (2, 5, 9)]
    anchors = [s.particle[:, :3].copy() for s in systems]
    say("long range: batch of %s atoms" % [s.n for s in systems])
    b = mdrl.Batch.from_synthetic(systems)
    pe0 = [b.system_energy(k)[0] for k in range(3)]
    rec = Recorder(anchors)
    b.set_long_range(mdrl.host_provider(rec))
    b.step(3)
    xq, box, off, _ = rec.calls[-1]
    assert off.tolist() == [0, systems[0].n, systems[0].n + systems[1].n, sum(s.n for s in systems)]
    for k, s in enumerate(systems):
        assert np.array_equal(box[k], s.box.astype(np.float32))
    for k, s in enumerate(systems):
        g = mdrl.System.from_synthetic(s)
        g.set_long_range(mdrl.host_provider(Recorder([anchors[k]])))
        g.step(3)
        assert np.array_equal(b.system_particles(k).view(np.uint32), g.particles().view(np.uint32)), k
        assert np.array_equal(b.system_forces_fixed(k), g.forces_fixed()), k
        g.close()
    # energy: the provider is asked for its energy, which joins the potential energy
    ncalls = len(rec.calls)
    pe = [b.system_energy(k)[0] for k in range(3)]
    assert len(rec.calls) > ncalls and rec.calls[-1][3]
    xq = rec.calls[-1][0]
    b.set_long_range(None)
    pe_plain = [b.system_energy(k)[0] for k in range(3)]
    b.close()
    for k, s in enumerate(systems):
        sl = slice(off[k], off[k + 1])
        _, e = tether_forces(xq[sl], anchors[k], s.box)
        # pe = range-limited + bonded + provider: the difference carries the rounding of
        # pe itself (unrelaxed replicas reach |pe| ~ 1e13 kJ/mol)
        tol = 4 * float(np.spacing(max(abs(pe[k]), abs(pe_plain[k]))))
        say("system %d: pe %.6e, provider energy %.6f" % (k, pe[k], e))
        assert pe[k] - pe_plain[k] == pytest.approx(e, rel=1e-12, abs=tol)
    assert pe0[0] != pe_plain[0]   # the system moved under the tether and field


def test_provider_errors():
    s = mdrl.SyntheticSystem(4, 0)
    g = mdrl.System.from_synthetic(s)
    with pytest.raises(mdrl.MdrlError) as e:
        g.set_long_range(lambda view, stream: 7)
    assert e.value.code == -1 and "returned 7" in str(e.value)
    # a host round trip cannot run inside a captured graph: the step reports it
    g.set_long_range(mdrl.host_provider(Recorder([s.particle[:, :3].copy()])))
    g.set_graph(True)
    with pytest.raises(mdrl.MdrlError):
        g.step(20)
    g.close()
    # a non-finite provider force is reported like any other
    f = np.zeros((s.n, 3), np.float32)
    f[5, 1] = np.nan
    buf = mdrl.DeviceBuffer(f)
    g = mdrl.System.from_synthetic(s)
    with pytest.raises(mdrl.MdrlError) as e:
        g.set_long_range(mdrl.device_provider(buf))
    assert e.value.code == -6
    g.close()
