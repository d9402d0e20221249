"""Batched and served scenes on the GPU: every scene's result is bitwise the
same alone and in a mix; error paths forced from the host."""
import os

import numpy as np
import pytest

from tests.common import SCENES, scene_cfg, tiny_cfg
from tisphi_amd import _lib as L
from tisphi_amd.batch import BatchEngine, build_family, dam_break_family
from tisphi_amd.scene import build_scene
from tisphi_amd.serve import Request, Server, plan_capacity, request_mix

pytestmark = pytest.mark.gpu

FIELDS_WC = ["pos", "vel", "density", "pressure", "d_vel", "d_density", "id0", "cspm_f", "v_tmp"]
FIELDS_SOIL = FIELDS_WC + ["stress", "strain_equ"]


def alone(sc, steps):
    e = L.Engine(sc.method, sc.geo.dim, sc.n, sc.geo.grid_num_total, 1)
    s = e.add_scene(sc.desc())
    L.upload_scene(e, s, sc)
    e.step(steps)
    e.sync()
    return e, s


def test_batch_alone_equals_in_mix():
    scenes = build_family(dam_break_family(12, seed=3))
    # room for one more scene of the family after a retirement
    b = BatchEngine(scenes, spare_particles=10_000, spare_cells=6160, max_scenes=13)
    b.step(40)
    b.sync()
    for k in (0, 5, 11):
        e, s = alone(scenes[k], 40)
        for f in FIELDS_WC:
            np.testing.assert_array_equal(b.engine.download(b.slots[k], f), e.download(s, f), err_msg=f)
        np.testing.assert_array_equal(b.engine.globals(b.slots[k]), e.globals(s))
        e.close()
    # retire a scene and admit another into its range: still identical
    b.engine.remove_scene(b.slots[5])
    new = build_scene(dam_break_family(20, seed=9)[13])
    slot = b.engine.add_scene(new.desc())
    L.upload_scene(b.engine, slot, new)
    b.step(25)
    b.sync()
    e, s = alone(new, 25)
    for f in FIELDS_WC:
        np.testing.assert_array_equal(b.engine.download(slot, f), e.download(s, f), err_msg=f)
    e.close()
    b.close()


def test_soil_batch_alone_equals_in_mix():
    cfgs = []
    for k, w in enumerate((0.06, 0.08, 0.1)):
        c = tiny_cfg(method=3 if k != 1 else 3, dim=2, xsph=True, r=0.002, domain=(0.3, 0.2, 0.06),
                     block=(w, 0.05, 0.04))
        cfgs.append(c)
    scenes = [build_scene(c) for c in cfgs]
    b = BatchEngine(scenes)
    b.step(30)
    b.sync()
    for k in range(3):
        e, s = alone(scenes[k], 30)
        for f in FIELDS_SOIL + ["flag_retmap", "strain_equ_p"]:
            np.testing.assert_array_equal(b.engine.download(b.slots[k], f), e.download(s, f), err_msg=f)
        e.close()
    b.close()


def test_served_results_equal_alone():
    reqs = request_mix(n=9, seed=7, rate=None, steps_range=(20, 60),
                       kinds=("wc2d", "mui2d", "wc3d"))
    # keep the 3D requests small for the test
    small = []
    for r in reqs:
        if r.kind == "wc3d" and r.scene.n > 120_000:
            continue
        small.append(r)
    srv = Server(plan_capacity(small, headroom=0.6), chunk=16, max_scenes=8)
    done = srv.run(small, realtime=False)
    assert len(done) == len(small)
    for r in done:
        assert r.result["steps"] == r.steps
        e, s = alone(r.scene, r.steps)
        np.testing.assert_array_equal(r.result["globals"], e.globals(s))
        e.close()
    srv.close()


def test_escaped_particle_is_reported_not_faulted():
    sc = build_scene(scene_cfg("dam_break_2d"))
    e = L.Engine(sc.method, sc.geo.dim, sc.n, sc.geo.grid_num_total, 1)
    s = e.add_scene(sc.desc())
    L.upload_scene(e, s, sc)
    vel = sc.vel.copy()
    vel[10] = [1.0e5, 0.0, 0.0]  # 10 m in one step: the next hash clamps and counts it
    e.upload(s, "vel", vel)
    e.step(2)
    with pytest.raises(L.ParticleEscaped):
        e.sync()
    assert e.status(s)[0] >= 1
    e.close()


@pytest.mark.parametrize("ti", [1, 2, 4])
def test_nan_is_reported(ti):
    cfg = scene_cfg("dam_break_2d")
    cfg.config["Configuration"]["timeIntegration"] = ti
    sc = build_scene(cfg)
    e = L.Engine(sc.method, sc.geo.dim, sc.n, sc.geo.grid_num_total, 1)
    s = e.add_scene(sc.desc())
    L.upload_scene(e, s, sc)
    rho = sc.density.copy()
    rho[3] = np.nan
    e.upload(s, "density", rho)
    e.step(1)
    with pytest.raises(L.NaNDetected):
        e.sync()
    e.close()


def test_facade_and_headless_cli(tmp_path):
    import json
    from tisphi_amd import export
    from tisphi_amd.eng.simulation import Simulation
    from tisphi_amd.run_simulation import run
    cfg = scene_cfg("dam_break_2d")
    sim = Simulation(cfg)
    assert sim.ps.particle_num[None] == 9144
    assert sim.solver.dt[None] == pytest.approx(1e-4)
    for _ in range(5):
        sim.solver.step()
    pos, data = sim.ps.dump()
    assert set(export.CSV_COLUMNS) <= set(pos) | set(data)
    assert len(pos["pos.x"]) == 9144 and data["id0"].dtype == np.int64
    x = sim.ps.pt.x.to_numpy()
    np.testing.assert_array_equal(x[:, 0], pos["pos.x"])
    # CLI run with exports every 5 steps
    c = json.load(open(os.path.join(SCENES, "dam_break_2d.json")))
    cfg_path = str(tmp_path / "s.json")
    c["Configuration"].update({"exitAtStep": 10, "stepsPerRenderUpdate": 5, "exportEveryRender": 1,
                               "exportCSV": True, "exportVTK": True})
    json.dump(c, open(cfg_path, "w"))
    case, n, g = run(cfg_path, out_dir=str(tmp_path / "out"), log=lambda *a: None)
    assert n == 10
    files = sorted(p.name for p in (tmp_path / "out").iterdir())
    assert "sim.csv.000010.csv" in files and "sim.vtk.000005.vtu" in files and "_info.txt" in files
    arr = np.loadtxt(tmp_path / "out" / "sim.csv.000010.csv", delimiter=",")
    assert arr.shape == (9144, 17)
    pts, d = export.read_vtu_points(str(tmp_path / "out" / "sim.vtk.000010.vtu"))
    np.testing.assert_allclose(pts[:, 0], arr[:, 3], rtol=0, atol=1e-15)


def test_mixed_options_batch_alone_equals_in_mix():
    """Scenes with different per-scene switches (kernel, correction, XSPH,
    boundary, particle size) in one engine: each ends bitwise where it ends alone, although
    the engine's step sequence is the one of the union of the switches."""
    opts = [dict(), dict(xsph=True), dict(kernel=0), dict(kcorr=1), dict(kcorr=2),
            dict(boundary=1), dict(boundary=3), dict(boundary=4, xsph=True, kernel=0),
            dict(r=0.004)]  # another particle size: its own dt, grid and cell range
    scenes = [build_scene(tiny_cfg(**o)) for o in opts]
    b = BatchEngine(scenes)
    b.step(30)
    b.sync()
    for k, sc in enumerate(scenes):
        e, s = alone(sc, 30)
        for f in FIELDS_WC:
            np.testing.assert_array_equal(b.engine.download(b.slots[k], f), e.download(s, f),
                                          err_msg=f"{opts[k]}: {f}")
        e.close()
    b.close()
