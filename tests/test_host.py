"""Host-side logic that runs without a GPU: exports, request generation,
capacity planning, scene families."""
import numpy as np
import pytest

from tisphi_amd import export
from tisphi_amd.batch import BatchEngine, build_family, dam_break_family
from tisphi_amd.serve import Request, class_key, plan_capacity, request_mix


def _dump(n=50, seed=0):
    rng = np.random.default_rng(seed)
    pos = {k: rng.normal(size=n) for k in ("pos.x", "pos.y", "pos.z")}
    data = {"id0": np.arange(n, dtype=np.int64), "objId": np.zeros(n, np.int64),
            "material": np.ones(n, np.int64)}
    for k in export.CSV_COLUMNS[6:]:
        data[k] = rng.normal(size=n)
    data["pressure"] = rng.normal(size=n)
    return pos, data


def test_csv_columns_and_header(tmp_path):
    pos, data = _dump()
    p = export.export_csv(str(tmp_path / "a.csv"), pos, data)
    first = open(p).readline().strip()
    assert first == "# " + ", ".join(export.CSV_COLUMNS)
    arr = np.loadtxt(p, delimiter=",")
    assert arr.shape == (50, 17)
    np.testing.assert_array_equal(arr[:, 0], data["id0"])
    np.testing.assert_array_equal(arr[:, 3], pos["pos.x"])
    np.testing.assert_array_equal(arr[:, 16], data["strain_equ"])


def test_vtu_round_trip(tmp_path):
    pos, data = _dump(seed=1)
    p = export.export_vtu(str(tmp_path / "b"), pos, data)
    pts, d = export.read_vtu_points(p)
    np.testing.assert_array_equal(pts[:, 1], pos["pos.y"])
    for k, v in data.items():
        np.testing.assert_array_equal(d[k], v)


def test_vtu_appended_layout(tmp_path):
    """The file follows the VTK XML layout pyevtk's pointsToVTK writes:
    every DataArray 'appended' at an offset into one raw AppendedData
    section, each block = UInt64 byte count + little-endian data, blocks in
    header order (points interleaved x, y, z; Int32 connectivity and
    offsets; UInt8 VTK_VERTEX types; the dump arrays)."""
    import xml.etree.ElementTree as ET
    pos, data = _dump(n=37, seed=2)
    p = export.export_vtu(str(tmp_path / "c"), pos, data)
    raw = open(p, "rb").read()
    k = raw.index(b'<AppendedData encoding="raw">\n_')
    base = k + len(b'<AppendedData encoding="raw">\n_')
    assert raw.endswith(b"\n</AppendedData>\n</VTKFile>")
    root = ET.fromstring(raw[:k] + b"<AppendedData/></VTKFile>")
    assert root.attrib == {"type": "UnstructuredGrid", "version": "1.0", "byte_order": "LittleEndian",
                           "header_type": "UInt64"}
    piece = root.find("UnstructuredGrid/Piece")
    assert piece.attrib == {"NumberOfPoints": "37", "NumberOfCells": "37"}
    das = piece.findall(".//DataArray")
    names = [d.get("Name") for d in das]
    assert names == ["points", "connectivity", "offsets", "types"] + list(data)
    assert piece.find("PointData").get("Scalars") == "id0"
    size = {"Float64": 8, "Int64": 8, "Int32": 4, "UInt8": 1}
    dt = {"Float64": "<f8", "Int64": "<i8", "Int32": "<i4", "UInt8": "u1"}
    off = 0
    blocks = {}
    for d in das:
        assert d.get("format") == "appended" and int(d.get("offset")) == off
        nb = int(np.frombuffer(raw, "<u8", 1, base + off)[0])
        ncomp = int(d.get("NumberOfComponents"))
        assert nb == 37 * ncomp * size[d.get("type")]
        blocks[d.get("Name")] = np.frombuffer(raw, dt[d.get("type")], 37 * ncomp, base + off + 8)
        off += 8 + nb
    assert base + off == len(raw) - len(b"\n</AppendedData>\n</VTKFile>")
    pts = blocks["points"].reshape(37, 3)
    for a, key in enumerate(("pos.x", "pos.y", "pos.z")):
        np.testing.assert_array_equal(pts[:, a], pos[key])
    np.testing.assert_array_equal(blocks["connectivity"], np.arange(37))
    np.testing.assert_array_equal(blocks["offsets"], np.arange(1, 38))
    np.testing.assert_array_equal(blocks["types"], np.ones(37))
    for key, v in data.items():
        np.testing.assert_array_equal(blocks[key], v)


def test_stamps():
    assert export.step_stamp(42) == "000042"
    assert export.time_stamp(0.34922) == "time.secx1e6.0349220"


def test_dam_break_family_seeded():
    a = build_family(dam_break_family(16, seed=0))
    b = build_family(dam_break_family(16, seed=0))
    assert [s.n for s in a] == [s.n for s in b]
    for s in a:
        assert 600 <= s.block_counts[0] <= 7200
        assert s.n_dummy == 1944
        assert s.geo.grid_num_total == 6160
    assert len(set(s.n for s in a)) > 8


def test_request_mix_sizes_and_arrivals():
    reqs = request_mix(n=30, seed=1, rate=2.0, steps_range=(500, 5000))
    t = [r.arrival for r in reqs]
    assert t == sorted(t) and t[0] > 0
    for r in reqs:
        nf = r.scene.n_flow
        lo, hi = {"wc2d": (1e3, 5e4), "wc3d": (5e4, 5e5), "mui2d": (1e4, 8e4)}[r.kind]
        assert 0.7 * lo <= nf <= 1.3 * hi, (r.kind, nf)
        assert 500 <= r.steps <= 5000
    cap = plan_capacity(reqs)
    for k, (p, c) in cap.items():
        assert p >= max(r.scene.n for r in reqs if class_key(r.scene) == k)


def test_one_integrator_per_engine_class():
    """An engine steps its residents with one fused sequence, so scenes of
    another time integration go to another serving class and a batch
    refuses them before any engine exists."""
    from tests.common import tiny_cfg
    from tisphi_amd.scene import build_scene
    lf, se = build_scene(tiny_cfg(ti=2)), build_scene(tiny_cfg(ti=1))
    assert class_key(lf) != class_key(se) and class_key(lf)[:2] == class_key(se)[:2]
    cap = plan_capacity([Request(0, lf, 10, 0.0), Request(1, se, 10, 0.0)])
    assert set(cap) == {class_key(lf), class_key(se)}
    with pytest.raises(ValueError):
        BatchEngine([lf, se])
