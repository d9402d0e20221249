"""Exports of a GPU run against the oracle: the 2D dam break stepped 20
times through the reference's facade (Simulation / solver.step), dumped
(ps.dump, particle_system.py:459-545) and written as CSV
(export_csv, ui_sim.py:293-298) and VTU (export_vtk, ui_sim.py:300-304):
read back and matched by id0, every exported column equals the float32
oracle's state after the same 20 steps."""
import numpy as np
import pytest

from oracle.oracle import Oracle, load_scene
from tests.common import scene_cfg
from tisphi_amd import export
from tisphi_amd.eng.simulation import Simulation
from tisphi_amd.scene import build_scene

pytestmark = pytest.mark.gpu

STEPS = 20


def test_csv_and_vtu_hold_the_oracle_state(tmp_path):
    cfg = scene_cfg("dam_break_2d")
    sim = Simulation(cfg)
    for _ in range(STEPS):
        sim.solver.step()
    pos, data = sim.ps.dump()
    csv = export.export_csv(str(tmp_path / "sim.csv.000020.csv"), pos, data)
    vtu = export.export_vtu(str(tmp_path / "sim.vtk.000020"), pos, data)

    sc = build_scene(scene_cfg("dam_break_2d"))
    o = Oracle(sc.desc(), 32)
    load_scene(o, sc)
    o.step(STEPS)
    oid = o.get("id0")
    ref = {"pos": o.get("pos")[np.argsort(oid)], "vel": o.get("vel")[np.argsort(oid)],
           "density": o.get("density")[np.argsort(oid)], "pressure": o.get("pressure")[np.argsort(oid)],
           "material": o.get("type")[np.argsort(oid)]}
    assert np.any(ref["vel"][:, 1] != 0.0)  # the state did move

    arr = np.loadtxt(csv, delimiter=",")
    cols = {c: arr[:, k] for k, c in enumerate(export.CSV_COLUMNS)}
    ic = np.argsort(cols["id0"])
    np.testing.assert_array_equal(cols["id0"][ic], np.arange(sc.n))
    np.testing.assert_array_equal(cols["material"][ic], ref["material"])
    for a, c in enumerate(("pos.x", "pos.y", "pos.z")):
        np.testing.assert_array_equal(cols[c][ic], ref["pos"][:, a], err_msg=c)
    for a, c in enumerate(("vel.x", "vel.y", "vel.z")):
        np.testing.assert_array_equal(cols[c][ic], ref["vel"][:, a], err_msg=c)
    np.testing.assert_array_equal(cols["density"][ic], ref["density"])

    pts, d = export.read_vtu_points(vtu)
    iv = np.argsort(d["id0"])
    np.testing.assert_array_equal(d["id0"][iv], np.arange(sc.n))
    np.testing.assert_array_equal(pts[iv], ref["pos"])
    np.testing.assert_array_equal(np.stack([d["vel.x"], d["vel.y"], d["vel.z"]], axis=1)[iv], ref["vel"])
    np.testing.assert_array_equal(d["density"][iv], ref["density"])
    np.testing.assert_array_equal(d["pressure"][iv], ref["pressure"])
    np.testing.assert_array_equal(d["vel.norm"][iv], np.linalg.norm(ref["vel"], axis=1))
