"""Autotuned launch shapes leave results bitwise unchanged (GPU)."""
import numpy as np
import pytest

from tests.common import scene_cfg
from tisphi_amd import _lib as L
from tisphi_amd.autotune import tune
from tisphi_amd.scene import build_scene

pytestmark = pytest.mark.gpu

FIELDS = ["cell", "offset", "vel", "density", "pressure", "m_V", "id0", "d_vel", "d_density"]


def _engine(sc):
    e = L.Engine(sc.method, sc.geo.dim, sc.n, sc.geo.grid_num_total, 1)
    s = e.add_scene(sc.desc())
    L.upload_scene(e, s, sc)
    return e, s


def test_tuned_steps_equal_default_steps(tmp_path, monkeypatch):
    monkeypatch.setenv("TSPH_TUNE_CACHE", str(tmp_path / "shapes.json"))
    sc = build_scene(scene_cfg("dam_break_2d"))
    e1, s1 = _engine(sc)
    r = tune(e1, steps=1, quick=False)
    k = r["steps_taken"]
    assert k == 36 * 2 and not r["cached"]
    e1.step(5)
    e1.sync()
    e2, s2 = _engine(sc)
    e2.step(k + 5)
    e2.sync()
    assert e1.steps(s1) == e2.steps(s2) == k + 5
    for f in FIELDS:
        np.testing.assert_array_equal(e1.download(s1, f), e2.download(s2, f), err_msg=f)
    # a second engine of the same class takes the cached shape without stepping
    e3, s3 = _engine(sc)
    r3 = tune(e3)
    assert r3["cached"] and r3["steps_taken"] == 0 and tuple(r3["shape"]) == tuple(r["shape"])
    for e in (e1, e2, e3):
        e.close()
