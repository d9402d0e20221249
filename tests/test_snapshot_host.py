"""State snapshots on the host: the field set per method, the file round
trip (arrays and JSON header, loaded without pickle) and the checks a
restore makes, against an in-memory stand-in for the engine (no GPU)."""
import json

import numpy as np
import pytest

from tests.common import add_rigid_blocks, scene_cfg
from tisphi_amd import _lib as L
from tisphi_amd import snapshot
from tisphi_amd.scene import build_scene


class FakeEngine:
    """download / upload / set_rigid_object over numpy arrays."""

    def __init__(self, sc, seed=0):
        self.method, self.dim = sc.method, sc.geo.dim
        self.n = {0: sc.n}
        rng = np.random.default_rng(seed)
        self.f = {}
        for name, (_, dt, k) in L.FIELDS.items():
            shape = (sc.n, k) if k > 1 else (sc.n,)
            self.f[name] = (rng.integers(0, 1000, shape).astype(dt) if dt == "i4"
                            else rng.normal(size=shape))
        self.rigid = {}

    def sync(self):
        pass

    def download(self, slot, name):
        return self.f[name].copy()

    def upload(self, slot, name, arr):
        assert arr.shape == self.f[name].shape
        self.f[name] = np.array(arr, dtype=self.f[name].dtype)

    def set_rigid_object(self, slot, obj, cm):
        self.rigid[obj] = list(cm)


def test_fields_per_method():
    assert snapshot.state_fields(1, False) == snapshot.STATE
    assert "stress_tmp" in snapshot.state_fields(2, False)
    assert "flag_retmap" not in snapshot.state_fields(2, False)
    f3 = snapshot.state_fields(3, True)
    assert "flag_retmap" in f3 and "d_stress" in f3 and f3[-1] == "x0"
    for f in f3:
        assert f in L.FIELDS


@pytest.mark.parametrize("name", ["dam_break_2d", "column_collapse_mui", "column_collapse_dp"])
def test_round_trip(tmp_path, name):
    cfg = scene_cfg(name)
    if name == "dam_break_2d":
        add_rigid_blocks(cfg, [(10, (1.205, 0.05, 0.0), (0.1, 0.1, 0.1), True)])
    sc = build_scene(cfg)
    a = FakeEngine(sc, 1)
    path = snapshot.save(str(tmp_path / "s.npz"), a, 0, sc, step=123)
    meta, arrays = snapshot.load(path)
    assert meta["step"] == 123 and meta["n"] == sc.n and meta["method"] == sc.method
    assert meta["config"] == json.loads(json.dumps(cfg.config))
    b = FakeEngine(sc, 2)
    snapshot.restore(b, 0, meta, arrays)
    for f in snapshot.state_fields(sc.method, bool(sc.rigid)):
        np.testing.assert_array_equal(b.f[f], a.f[f], err_msg=f)
    if sc.rigid:
        assert b.rigid == {10: [float(v) for v in sc.rigid[0]["rest_cm"]]}


def test_scene_json_drops_derived_keys(tmp_path):
    """Loader-derived keys (particle counts, a dynamic body's rest mesh)
    stay out of the header; the scene JSON itself is kept."""
    from tests.common import add_obj_bodies, box_mesh, write_obj
    cfg = scene_cfg("dam_break_3d")
    cfg.config["Configuration"]["particleRadius"] = 0.01
    path = write_obj(tmp_path / "b.obj", *box_mesh((0, 0, 0), (0.1, 0.1, 0.1)))
    add_obj_bodies(cfg, [(7, path, (1.5, 0.1, 0.2), 1.0, 0.0, (0, 0, 1), True)])
    sc = build_scene(cfg)
    assert "restPosition" in cfg.get_bodies()[0]
    p = snapshot.save(str(tmp_path / "s.npz"), FakeEngine(sc), 0, sc)
    meta, _ = snapshot.load(p)
    body = meta["config"]["Bodies"][0]
    assert "restPosition" not in body and body["geometryFile"] == path
    assert meta["config"]["Configuration"] == cfg.config["Configuration"]


def test_restore_checks(tmp_path):
    sc = build_scene(scene_cfg("dam_break_2d"))
    path = snapshot.save(str(tmp_path / "s.npz"), FakeEngine(sc), 0, sc)
    meta, arrays = snapshot.load(path)
    other = build_scene(scene_cfg("column_collapse_mui"))
    with pytest.raises(ValueError):
        snapshot.restore(FakeEngine(other), 0, meta, arrays)  # particle count
    e = FakeEngine(sc)
    e.method = 2
    with pytest.raises(ValueError):
        snapshot.restore(e, 0, meta, arrays)  # method
    # a file without the header, one of another format, one missing a field
    np.savez(str(tmp_path / "x.npz"), cell=np.zeros(3))
    with pytest.raises(ValueError):
        snapshot.load(str(tmp_path / "x.npz"))
    bad = dict(meta, format=99)
    np.savez(str(tmp_path / "y.npz"), __meta__=np.frombuffer(json.dumps(bad).encode(), np.uint8),
             **arrays)
    with pytest.raises(ValueError):
        snapshot.load(str(tmp_path / "y.npz"))
    part = {k: v for k, v in arrays.items() if k != "vel"}
    np.savez(str(tmp_path / "z.npz"), __meta__=np.frombuffer(json.dumps(meta).encode(), np.uint8),
             **part)
    with pytest.raises(ValueError):
        snapshot.load(str(tmp_path / "z.npz"))
    # object arrays would need pickle: refused by the loader
    np.savez(str(tmp_path / "o.npz"), __meta__=np.frombuffer(json.dumps(meta).encode(), np.uint8),
             obj=np.array([{"a": 1}], dtype=object), **arrays)
    with pytest.raises(ValueError):
        snapshot.load(str(tmp_path / "o.npz"))
