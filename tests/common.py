"""Shared helpers of the test suite (scene builders, tolerances)."""
import copy
import json
import os

import numpy as np

from tisphi_amd.eng.configer_builder import SimConfiger
from tisphi_amd.scene import build_scene

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCENES = os.path.join(ROOT, "tisphi_amd", "scenes")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def scene_cfg(name):
    return SimConfiger(os.path.join(SCENES, name + ".json"))


def add_rigid_blocks(cfg, blocks, density=500.0):
    """Append a rigid material (matType 11) and rigid blocks to a scene
    config: ``blocks`` = [(objectId, translation, size, isDynamic), ...]."""
    c = cfg.config
    mat_id = len(c["Materials"])
    c["Materials"].append({"matId": mat_id, "matType": 11, "color": [120, 80, 40],
                           "density0": density})
    for oid, tr, size, dyn in blocks:
        c["Blocks"].append({"objectId": oid, "translation": list(tr), "size": list(size),
                            "velocity": [0.0, 0.0, 0.0], "materialId": mat_id,
                            "isDynamic": bool(dyn)})
    return cfg


def box_mesh(lo, hi):
    """Vertices and quad faces (1-based, outward) of an axis-aligned box."""
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    v = np.array([[(hi if (k >> a) & 1 else lo)[a] for a in range(3)] for k in range(8)])
    quads = [[5, 7, 3, 1], [4, 8, 6, 2], [2, 6, 5, 1], [7, 8, 4, 3], [3, 4, 2, 1], [6, 8, 7, 5]]
    return v, quads


def write_obj(path, verts, faces):
    """A Wavefront OBJ file of ``verts`` and 1-based polygon ``faces``."""
    with open(path, "w") as f:
        f.write("# test mesh\n")
        for x in verts:
            f.write("v %.17g %.17g %.17g\n" % tuple(x))
        for q in faces:
            f.write("f " + " ".join(str(int(k)) for k in q) + "\n")
    return str(path)


def add_obj_bodies(cfg, bodies, density=500.0):
    """Append a rigid material and OBJ bodies to a scene config:
    ``bodies`` = [(objectId, path, translation, scale, angle_deg, axis,
    isDynamic), ...]."""
    c = cfg.config
    mat_id = len(c["Materials"])
    c["Materials"].append({"matId": mat_id, "matType": 11, "color": [200, 60, 60],
                           "density0": density})
    c.setdefault("Bodies", [])
    for oid, path, tr, scale, angle, axis, dyn in bodies:
        c["Bodies"].append({"objectId": oid, "geometryFile": path, "translation": list(tr),
                            "scale": scale, "rotationAngle": angle, "rotationAxis": list(axis),
                            "velocity": [0.0, 0.0, 0.0], "materialId": mat_id,
                            "isDynamic": bool(dyn)})
    return cfg


def tiny_cfg(method=1, dim=2, xsph=False, r=None, domain=(0.2, 0.2, 0.06), block=(0.08, 0.08, 0.04),
             boundary=2, ti=2, kernel=1, kcorr=0):
    """A small dam-break / column scene for brute-force checks."""
    base = json.load(open(os.path.join(SCENES, "dam_break_2d.json" if method == 1 else
                                       "column_collapse_mui.json")))
    c = base["Configuration"]
    c.update({"domainEnd": list(domain), "particleRadius": r or (0.005 if dim == 2 else 0.01),
              "is2D": dim == 2, "simulationMethod": method, "xsph": xsph, "boundary": boundary,
              "timeIntegration": ti, "kernel": kernel, "kernelCorrection": kcorr})
    base["Blocks"][0]["size"] = list(block)
    return SimConfiger(config=base)


# Per-element parity tolerance |a - b| <= tol * max(|b|, floor) with the
# floors of the measurement plan (rho0 for densities, rho0*|g|*H for
# pressure and stress, |g| for accelerations, max|ref| for rates).
TOL = 1e-5


def floors(scene, H):
    rho0 = scene.material["density0"]
    g = float(np.linalg.norm(scene.gravity))
    return dict(density=rho0, density_tmp=rho0, pressure=rho0 * g * H, stress=rho0 * g * H,
                stress_tmp=rho0 * g * H, d_vel=g)


def parity_err(a, b, floor):
    """max |a-b| / max(|b|, floor)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if b.size == 0:
        return 0.0
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))
