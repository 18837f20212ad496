"""GPU: a plain C program (sc2019_md_amd/examples/host_c.c) linking libmdrl in place of the
reference's C model.

Bars: MD_Evaluation_Model(), called by the C host as a reference host would, prints the
oracle's 100 iteration lines character for character; the half-shell system path driven
from C (512-atom argon, 20 velocity-Verlet steps) leaves int64 forces bit-identical to the
oracle's and to the same steps driven from Python.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from sc2019_md_amd import build as B
from sc2019_md_amd import mdrl

from tests.test_oracle_golden import (  # noqa: E402
    cmodel_inputs, parse_pdb_float32, pdb_text, rl_from_synthetic)

pytestmark = pytest.mark.gpu


def host():
    if not os.path.exists(B.HOST_OUT):
        pytest.fail("C host not built: run python sc2019_md_amd/build.py")
    return B.HOST_OUT


def test_c_host_md_evaluation_model(tmp_path):
    text = pdb_text(cmodel_inputs()["uniform"])
    pdb = tmp_path / "in.pdb"
    pdb.write_text(text)
    print("C host: MD_Evaluation_Model on a 500-atom PDB", flush=True)
    r = subprocess.run([host(), "cmodel"], env=dict(os.environ, MDRL_INPUT=str(pdb)),
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    p = np.zeros((500, 11), np.float32)
    p[:, :3] = parse_pdb_float32(text)
    _, _, lines = O.cmodel_run(p, O.cmodel_default_params(), with_lines=True)
    assert len(r.stdout.splitlines()) == 100
    assert r.stdout == lines


def test_c_host_system_steps(tmp_path):
    out = tmp_path / "fix.bin"
    print("C host: 20 steps of the 512-atom system", flush=True)
    r = subprocess.run([host(), "steps", "20", str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(str(out), dtype="<i8").reshape(-1, 3)
    s = mdrl.SyntheticSystem(1)
    g = mdrl.System.from_synthetic(s)
    g.step(20)
    py = g.forces_fixed()
    g.close()
    assert np.array_equal(got, py)
    t = mdrl.Table.load()
    rl = rl_from_synthetic(s, {k: t.get(k) for k in range(6)})
    hk = O.half_kick(s.mass, s.dt)
    pos4, vel = rl.pos4.copy(), s.particle[:, 3:6].copy()
    fix, _ = rl.forces_fix()
    for k in range(20):
        pos4, vel, _ = O.motion_update(pos4, vel, fix, s.type, hk, s.mass, s.dt, s.box, k == 0)
        rl.pos4[:] = pos4
        fix, npairs = rl.forces_fix()
    assert np.array_equal(got, fix)
    assert ("pairs %d " % npairs) in r.stdout
