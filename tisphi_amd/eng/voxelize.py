"""OBJ bodies: mesh loading, transform and solid voxelisation.

The reference loads a body with trimesh (load_body,
eng/particle_func.py:265-294): load the mesh, scale it, rotate it, then
``mesh.voxelized(pitch=particle_diameter).fill()`` and take the voxel
centres plus the body translation as particle positions.  trimesh is not
available here, so this module restates that pipeline with numpy / scipy:

* voxelized(pitch) (trimesh's "subdivide" voxeliser): every triangle is
  split in four until no edge is longer than pitch / 2; each vertex of the
  split mesh marks the voxel round(v / pitch);
* fill(): the interior of the marked shell, scipy.ndimage.binary_fill_holes;
* points: voxel index * pitch (voxel centres on the pitch lattice through
  the mesh origin).

Which voxels the two implementations mark can differ at the boundary of
the shell where a split vertex lies within rounding of a voxel face; the
counts of closed-form solids are checked in the tests.
"""
import math
import os

import numpy as np
from scipy import ndimage


def load_obj(path):
    """Vertices (n, 3) and triangles (m, 3) of a Wavefront OBJ file
    (``v`` and ``f`` records; polygons fanned into triangles)."""
    verts, faces = [], []
    with open(path) as f:
        for line in f:
            parts = line.split()
            if not parts:
                continue
            if parts[0] == "v":
                verts.append([float(x) for x in parts[1:4]])
            elif parts[0] == "f":
                idx = []
                for tok in parts[1:]:
                    k = int(tok.split("/")[0])
                    idx.append(k - 1 if k > 0 else len(verts) + k)
                for a in range(1, len(idx) - 1):
                    faces.append([idx[0], idx[a], idx[a + 1]])
    return np.array(verts, dtype=np.float64).reshape(-1, 3), np.array(faces, dtype=np.int64).reshape(-1, 3)


def rotation_matrix(angle, direction, point):
    """4x4 rotation by ``angle`` about the axis ``direction`` through
    ``point`` (the usual Rodrigues form, as trimesh.transformations)."""
    d = np.asarray(direction, dtype=np.float64)[:3]
    d = d / np.linalg.norm(d)
    s, c = math.sin(angle), math.cos(angle)
    R = np.diag([c, c, c]) + np.outer(d, d) * (1.0 - c)
    R += np.array([[0.0, -d[2], d[1]], [d[2], 0.0, -d[0]], [-d[1], d[0], 0.0]]) * s
    M = np.identity(4)
    M[:3, :3] = R
    p = np.asarray(point, dtype=np.float64)[:3]
    M[:3, 3] = p - R @ p
    return M


def _surface_samples(v, faces, max_edge):
    """Vertices of every triangle split in four until no edge is longer
    than ``max_edge`` (the barycentric lattice of 2^k divisions: the
    vertices that repeated midpoint splitting reaches)."""
    out = [v]
    for a, b, c in v[faces]:
        longest = max(np.linalg.norm(b - a), np.linalg.norm(c - b), np.linalg.norm(a - c))
        k = 0
        while longest / (2 ** k) > max_edge:
            k += 1
        n = 2 ** k
        if n == 1:
            continue
        i, j = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
        m = i + j <= n
        i, j = i[m], j[m]
        w = np.stack([(n - i - j), i, j], axis=1) / n
        out.append(w[:, :1] * a + w[:, 1:2] * b + w[:, 2:3] * c)
    return np.concatenate(out)


def voxel_points(vertices, faces, pitch):
    """Centres of the filled voxelisation of a closed mesh at ``pitch``."""
    if len(faces) == 0:
        raise ValueError("the body mesh has no faces")
    if not pitch > 0.0:
        raise ValueError(f"voxel pitch must be positive, got {pitch}")
    pts = _surface_samples(vertices, faces, pitch / 2.0)
    hit = np.round(pts / pitch).astype(np.int64)
    lo = hit.min(axis=0)
    shape = hit.max(axis=0) - lo + 1
    grid = np.zeros(shape, dtype=bool)
    grid[tuple((hit - lo).T)] = True
    grid = ndimage.binary_fill_holes(grid)
    idx = np.argwhere(grid) + lo
    return idx.astype(np.float64) * pitch


def load_body(body, vox_len, base_dir=None):
    """Particle positions of an OBJ body (load_body,
    eng/particle_func.py:265-294): scale, rotate, voxelise and fill, then
    translate."""
    path = body["geometryFile"]
    if not os.path.isabs(path) and not os.path.exists(path) and base_dir is not None:
        path = os.path.join(base_dir, path)
    v, f = load_obj(path)
    v = v * np.asarray(body.get("scale", 1.0), dtype=np.float64)
    angle = float(body.get("rotationAngle", 0.0)) / 180.0 * math.pi
    if angle != 0.0:
        # the reference passes mesh.vertices.mean(axis=-1) as the point, of
        # which the rotation uses the first three entries
        point = v.mean(axis=-1)[:3]
        M = rotation_matrix(angle, body.get("rotationAxis", [0, 0, 1]), point)
        v = v @ M[:3, :3].T + M[:3, 3]
    offset = np.asarray(body["translation"], dtype=np.float64)
    if body.get("isDynamic", False):
        # the rest mesh the reference keeps on a dynamic body
        # (particle_func.py:278-284; "restCenterOfMass" is its per-vertex
        # mean, mesh.vertices.mean(axis=-1), as there)
        body["restPosition"] = v + offset
        body["restCenterOfMass"] = body["restPosition"].mean(axis=-1)
    return voxel_points(v, f, vox_len) + offset
