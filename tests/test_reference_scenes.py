"""The reference's own scene files (tests/golden/scenes/, copied from its
data/scenes/) load unchanged: method, geometry and time step as the
reference derives them, block particle counts from the np.arange lattice
(particle_func.py:253-259), static rigid intruders as rigid objects without
shape matching, dummy walls present."""
import glob
import math
import os

import numpy as np
import pytest

from tests.common import GOLDEN
from tisphi_amd.eng.configer_builder import SimConfiger
from tisphi_amd.scene import build_scene

SCENES = sorted(glob.glob(os.path.join(GOLDEN, "scenes", "*.json")))


def lattice_count(tr, size, dim, d):
    n = 1
    for a in range(dim):
        off = d if size[a] >= 0 else -d
        n *= len(np.arange(tr[a] + off / 2.0, tr[a] + size[a] + 1e-5, off))
    return n


def test_all_five_present():
    assert [os.path.basename(p) for p in SCENES] == [
        "test1_db_water.json", "test2_cc_sand.json", "test3_ht.json", "test4_in_ver.json",
        "test5_in_hor.json"]


@pytest.mark.parametrize("path", SCENES, ids=[os.path.basename(p)[:-5] for p in SCENES])
def test_reference_scene_loads(path):
    cfg = SimConfiger(path)
    c = cfg.config["Configuration"]
    sc = build_scene(cfg)
    dim = 2 if c["is2D"] else 3
    d = 2 * c["particleRadius"]
    assert sc.method == c["simulationMethod"] and sc.geo.dim == dim
    # d = 2r, h = kh d, r_s = kappa h, grid = ceil(kappa kh) d (particle_system.py:14-58)
    assert sc.geo.particle_diameter == d
    assert sc.geo.smoothing_len == pytest.approx(c["kh"] * d, rel=1e-15)
    assert sc.geo.support_radius == pytest.approx(c["kappa"] * c["kh"] * d, rel=1e-15)
    assert sc.geo.grid_size == pytest.approx(math.ceil(c["kappa"] * c["kh"]) * d, rel=1e-15)
    # dt = max(dt_min, c_dt - c_dt % dt_min), c_dt = 0.2 h / c (solver_sph_base.py:209-212)
    h = c["kh"] * d
    mats = cfg.config["Materials"]
    if sc.method == 1:
        vs = 60.0
    else:
        soil = [m for m in mats if m["matType"] == 2][0]
        vs = 24.0 if sc.method == 2 else math.sqrt(soil["EYoungMod"] / soil["density0"])
    cdt = 0.2 * h / vs
    dt_min = c["timeStepSizeMin"]
    assert sc.dt == max(dt_min, cdt - cdt % dt_min)
    blocks = cfg.get_blocks()
    assert sc.block_counts == [lattice_count(b["translation"], b["size"], dim, d) for b in blocks]
    assert sc.n == sum(sc.block_counts) + sc.n_dummy and sc.n_dummy > 0
    # every block's particles carry its material type and object id, in file order
    at = 0
    for b, n in zip(blocks, sc.block_counts):
        mat = mats[b["materialId"]]
        assert np.all(sc.mat_type[at:at + n] == mat["matType"])
        assert np.all(sc.obj_id[at:at + n] == b["objectId"])
        at += n
    rigid = [b for b in blocks if mats[b["materialId"]]["matType"] == 11]
    assert [r["obj"] for r in sc.rigid] == [b["objectId"] for b in rigid]
    assert all(r["dynamic"] == bool(b.get("isDynamic", False)) for r, b in zip(sc.rigid, rigid))
