"""OBJ bodies (eng/particle_func.py:265-294, particle_system.py:82-91,
176-200): the OBJ reader, the rotation, the filled voxelisation and the
body particles of a scene.

The reference voxelises with trimesh, which is not installed here, so the
voxel sets are pinned by what the mathematics fixes instead: an axis-aligned
box gives exactly its lattice points, a convex solid holds every lattice
point deeper than half a voxel diagonal inside it and none farther than that
outside it (brute force over the bounding box), and the rotation agrees with
scipy's.  Which boundary voxels trimesh marks where a split vertex rounds to
a tie is parity unpinned.
"""
import json
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from tests.common import add_obj_bodies, box_mesh, scene_cfg, write_obj
from tisphi_amd.eng.configer_builder import SimConfiger
from tisphi_amd.eng.voxelize import load_body, load_obj, rotation_matrix, voxel_points
from tisphi_amd.scene import build_scene


def tri(verts, faces):
    """Fan-triangulated 0-based faces of 1-based polygons."""
    out = []
    for q in faces:
        for a in range(1, len(q) - 1):
            out.append([q[0] - 1, q[a] - 1, q[a + 1] - 1])
    return np.asarray(verts, float), np.asarray(out)


def as_set(points, pitch):
    return {tuple(k) for k in np.round(points / pitch).astype(np.int64)}


def test_obj_reader(tmp_path):
    p = tmp_path / "m.obj"
    p.write_text("# comment\no mesh\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\n"
                 "vn 0 0 1\nvt 0 0\nf 1/1/1 2/1/1 3/1/1 4/1/1\nf -4 -2 -1\n\n")
    v, f = load_obj(str(p))
    assert v.shape == (4, 3) and v[2].tolist() == [1.0, 1.0, 0.0]
    # the quad fans into two triangles; negative indices count from the end
    assert f.tolist() == [[0, 1, 2], [0, 2, 3], [0, 2, 3]]


def test_degenerate_inputs(tmp_path):
    p = tmp_path / "empty.obj"
    p.write_text("v 0 0 0\nv 1 0 0\n")
    v, f = load_obj(str(p))
    with pytest.raises(ValueError):
        voxel_points(v, f, 0.1)
    v, f = tri(*box_mesh((0, 0, 0), (1, 1, 1)))
    with pytest.raises(ValueError):
        voxel_points(v, f, 0.0)


def test_axis_aligned_box_is_its_lattice():
    for lo, hi, pitch in (((0, 0, 0), (1, 1, 1), 0.1), ((0, 0, 0), (0.5, 0.3, 0.2), 0.1),
                          ((-0.2, 0.1, 0.3), (0.2, 0.5, 0.5), 0.05)):
        v, f = tri(*box_mesh(lo, hi))
        pts = voxel_points(v, f, pitch)
        axes = [np.arange(round(lo[a] / pitch), round(hi[a] / pitch) + 1) for a in range(3)]
        want = {(i, j, k) for i in axes[0] for j in axes[1] for k in axes[2]}
        assert as_set(pts, pitch) == want and len(pts) == len(want)
        np.testing.assert_allclose(pts / pitch, np.round(pts / pitch), atol=1e-12)


def test_rotation_matrix_matches_scipy():
    rng = np.random.default_rng(3)
    for _ in range(20):
        axis = rng.normal(size=3)
        angle = rng.uniform(-math.pi, math.pi)
        point = rng.normal(size=3)
        M = rotation_matrix(angle, axis, point)
        R = Rotation.from_rotvec(axis / np.linalg.norm(axis) * angle).as_matrix()
        np.testing.assert_allclose(M[:3, :3], R, atol=1e-14)
        np.testing.assert_allclose(M[:3, :3] @ point + M[:3, 3], point, atol=1e-14)
        np.testing.assert_array_equal(M[3], [0, 0, 0, 1])


def convex_planes(v, f):
    """(normals, offsets) of the outward face planes of a convex mesh."""
    c = v.mean(axis=0)
    n = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    n /= np.linalg.norm(n, axis=1)[:, None]
    d = np.einsum("ij,ij->i", n, v[f[:, 0]])
    flip = n @ c > d
    n[flip] *= -1
    d[flip] *= -1
    return n, d


@pytest.mark.parametrize("shape", ["rotated_box", "tetrahedron", "octahedron"])
def test_convex_solid_brute_force(shape):
    """Every lattice point deeper than half the voxel diagonal inside the
    solid is a voxel; none farther than that outside it is."""
    pitch = 0.05
    if shape == "rotated_box":
        v, f = tri(*box_mesh((0, 0, 0), (0.4, 0.3, 0.25)))
        R = Rotation.from_rotvec([0.3, -0.5, 0.7]).as_matrix()
        v = v @ R.T + [0.013, -0.021, 0.007]
    elif shape == "tetrahedron":
        v = np.array([[0.0, 0.0, 0.0], [0.5, 0.02, 0.01], [0.1, 0.45, 0.03], [0.05, 0.1, 0.4]])
        f = np.array([[0, 2, 1], [0, 1, 3], [1, 2, 3], [0, 3, 2]])
    else:
        v = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]]) * 0.3
        v = v + [0.011, 0.017, -0.004]
        f = np.array([[0, 2, 4], [2, 1, 4], [1, 3, 4], [3, 0, 4],
                      [2, 0, 5], [1, 2, 5], [3, 1, 5], [0, 3, 5]])
    got = as_set(voxel_points(v, f, pitch), pitch)
    n, d = convex_planes(v, f)
    lo = np.floor(v.min(axis=0) / pitch) - 2
    hi = np.ceil(v.max(axis=0) / pitch) + 2
    grid = np.stack(np.meshgrid(*[np.arange(lo[a], hi[a] + 1) for a in range(3)], indexing="ij"), -1)
    grid = grid.reshape(-1, 3)
    signed = np.max(grid * pitch @ n.T - d, axis=1)  # < 0 inside; a lower bound of the distance outside
    half_diag = pitch * math.sqrt(3) / 2
    deep = {tuple(k) for k in grid[signed < -half_diag].astype(np.int64)}
    far = {tuple(k) for k in grid[signed > half_diag].astype(np.int64)}
    assert deep and deep <= got
    assert not (far & got)


def test_load_body_scale_rotation_translation(tmp_path):
    """load_body: scale, then a rotation about the axis through the point
    made of the first three entries of the per-vertex means (the reference
    passes mesh.vertices.mean(axis=-1)), then the translation."""
    verts, quads = box_mesh((0, 0, 0), (0.2, 0.1, 0.3))
    path = write_obj(tmp_path / "box.obj", verts, quads)
    pitch = 0.02
    body = {"geometryFile": path, "scale": 2.0, "rotationAngle": 90.0, "rotationAxis": [0, 0, 1],
            "translation": [0.5, 0.25, 0.1]}
    pts = load_body(body, pitch)
    sv = verts * 2.0
    c = sv.mean(axis=-1)[:3]  # per-vertex means of vertices 0, 1, 2
    lo, hi = sv.min(axis=0), sv.max(axis=0)
    # x' = cx - (y - cy), y' = cy + (x - cx)
    rlo = np.array([c[0] - (hi[1] - c[1]), c[1] + (lo[0] - c[0]), lo[2]])
    rhi = np.array([c[0] - (lo[1] - c[1]), c[1] + (hi[0] - c[0]), hi[2]])
    want = voxel_points(*tri(*box_mesh(rlo, rhi)), pitch) + body["translation"]
    assert as_set(pts, pitch) == as_set(want, pitch) and len(pts) == len(want)
    # no rotation: the scaled box on its lattice, translated
    body["rotationAngle"] = 0.0
    pts0 = load_body(body, pitch)
    assert len(pts0) == 21 * 11 * 31
    np.testing.assert_allclose(pts0.min(axis=0), body["translation"], atol=1e-12)
    np.testing.assert_allclose(pts0.max(axis=0), np.array([0.4, 0.2, 0.6]) + body["translation"],
                               atol=1e-12)


def body_scene(tmp_path, dyn=True, relative=False, tr=(1.5, 0.1, 0.2)):
    cfg = scene_cfg("dam_break_3d")
    cfg.config["Configuration"]["particleRadius"] = 0.01
    verts, quads = box_mesh((0, 0, 0), (0.1, 0.1, 0.1))
    path = write_obj(tmp_path / "box.obj", verts, quads)
    add_obj_bodies(cfg, [(7, "box.obj" if relative else path, tr, 1.0, 30.0, (0, 0, 1), dyn)])
    if relative:
        sp = tmp_path / "scene.json"
        sp.write_text(json.dumps(cfg.config))
        cfg = SimConfiger(str(sp))
    return cfg


def test_scene_with_body(tmp_path):
    """Particle order blocks -> bodies -> dummy walls; the body's material,
    object id and velocity; a dynamic rigid body registers its mass-weighted
    rest centre."""
    sc = build_scene(body_scene(tmp_path))
    nb = sc.block_counts[0]
    nbody = sc.body_counts[0]
    assert nb == 60 * 30 * 25 and nbody > 100
    assert sc.n == nb + nbody + sc.n_dummy == sc.pos.shape[0]
    body = slice(nb, nb + nbody)
    assert np.all(sc.mat_type[body] == 11) and np.all(sc.obj_id[body] == 7)
    assert np.all(sc.mat_type[:nb] == 1) and np.all(sc.mat_type[nb + nbody:] == -1)
    np.testing.assert_allclose(sc.density[body], 500.0)
    np.testing.assert_array_equal(sc.color0[body],
                                  np.tile(np.float32(np.array([200, 60, 60]) / 255), (nbody, 1)))
    (r,) = sc.rigid
    assert r["obj"] == 7 and r["dynamic"]
    np.testing.assert_allclose(r["rest_cm"], sc.pos[body].mean(axis=0), rtol=1e-12)
    # the voxels are those of load_body
    cfg = body_scene(tmp_path)
    want = load_body(cfg.get_bodies()[0], sc.geo.particle_diameter)
    np.testing.assert_array_equal(sc.pos[body], want)
    # a dynamic body keeps its rest mesh, as the reference's load_body does
    b = cfg.get_bodies()[0]
    assert b["restPosition"].shape == (8, 3)
    np.testing.assert_allclose(b["restCenterOfMass"], b["restPosition"].mean(axis=-1))
    # a static body is a rigid object without shape matching
    sc2 = build_scene(body_scene(tmp_path, dyn=False))
    assert [x["dynamic"] for x in sc2.rigid] == [False]


def test_scene_body_relative_path_and_errors(tmp_path):
    sc = build_scene(body_scene(tmp_path, relative=True))
    assert sc.body_counts[0] > 100
    with pytest.raises(ValueError):
        build_scene(body_scene(tmp_path, tr=(3.2, 0.1, 0.2)))  # leaves the domain
    cfg = scene_cfg("dam_break_2d")
    verts, quads = box_mesh((0, 0, 0), (0.1, 0.1, 0.1))
    add_obj_bodies(cfg, [(7, write_obj(tmp_path / "b.obj", verts, quads), (1.5, 0.1, 0.0), 1.0,
                          0.0, (0, 0, 1), True)])
    with pytest.raises(NotImplementedError):
        build_scene(cfg)


def test_facade_records_body_particle_numbers(tmp_path):
    """particle_system.py:82-91: body["particleNum"] is the voxel count and
    the body joins the object collection (no engine needed to check the
    host bookkeeping)."""
    from tisphi_amd.eng import particle_system as PS
    cfg = body_scene(tmp_path)
    sc = build_scene(cfg)
    bodies = cfg.get_bodies()
    PS.record_objects(cfg.get_blocks(), bodies, sc)
    assert bodies[0]["particleNum"] == sc.body_counts[0]
