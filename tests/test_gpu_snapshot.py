"""Snapshots on the GPU: a run that stops, writes a snapshot and resumes in
a fresh engine ends bitwise where an uninterrupted run ends (WC with a
dynamic rigid block, μ(I) whose soil loop reads the boundary v_tmp of the
step before, DP); the headless CLI's --checkpoint / --resume likewise."""
import json

import numpy as np
import pytest

from tests.common import add_rigid_blocks, scene_cfg
from tisphi_amd import _lib as L
from tisphi_amd import snapshot
from tisphi_amd.eng.simulation import Simulation

pytestmark = pytest.mark.gpu

CASES = {
    "wc2d_rigid": ("dam_break_2d", [(10, (1.205, 0.05, 0.0), (0.1, 0.1, 0.1), True)]),
    "mui2d": ("column_collapse_mui", None),
    "dp2d": ("column_collapse_dp", None),
}


def cfg_of(key):
    name, rigid = CASES[key]
    cfg = scene_cfg(name)
    if rigid:
        add_rigid_blocks(cfg, rigid)
    return cfg


def state(case):
    ps = case.ps
    sc = ps.scene
    f = snapshot.state_fields(sc.method, any(r["dynamic"] for r in sc.rigid))
    return {k: ps.engine.download(ps.slot, k) for k in f}


@pytest.mark.parametrize("key", sorted(CASES))
def test_resume_is_bitwise(tmp_path, key):
    print(f"snapshot resume {key}", flush=True)
    k = 40
    straight = Simulation(cfg_of(key))
    straight.solver.steps(2 * k)
    straight.solver.sync()
    want = state(straight)
    first = Simulation(cfg_of(key))
    first.solver.steps(k)
    path = first.ps.save_state(str(tmp_path / "s.npz"), k)
    first.ps.engine.close()
    resumed, step = Simulation.resume(path)
    assert step == k
    resumed.solver.steps(k)
    resumed.solver.sync()
    got = state(resumed)
    for f in want:
        np.testing.assert_array_equal(got[f], want[f], err_msg=f)
    np.testing.assert_array_equal(resumed.ps.globals(), straight.ps.globals())
    assert resumed.ps.engine.status(resumed.ps.slot) == (0, 0)


def test_cli_checkpoint_and_resume(tmp_path):
    from tisphi_amd import run_simulation
    cfg = scene_cfg("dam_break_2d").config
    cfg["Configuration"].update({"exitAtStep": 0, "stepsPerRenderUpdate": 10})
    p = tmp_path / "s.json"
    p.write_text(json.dumps(cfg))
    quiet = dict(log=lambda *a: None)
    _, n0, g0 = run_simulation.run(str(p), max_steps=60, **quiet)
    ck = str(tmp_path / "ck.npz")
    _, n1, _ = run_simulation.run(str(p), max_steps=30, checkpoint=ck, **quiet)
    _, n2, g2 = run_simulation.run(None, max_steps=60, resume=ck, **quiet)
    assert (n0, n1, n2) == (60, 30, 60)
    np.testing.assert_array_equal(g2, g0)
