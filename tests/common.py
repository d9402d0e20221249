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


FLUID_T, SOIL_T, RIGID_T = 1, 2, 11


def _list_keep(ti, tj):
    """Which for_all_neighbors pairs a recorded list keeps: every neighbour
    of a flow particle, flow neighbours of a boundary particle, flow and
    rigid neighbours of a rigid one."""
    flow_i = (ti == FLUID_T) | (ti == SOIL_T)
    flow_j = (tj == FLUID_T) | (tj == SOIL_T)
    bdy_i = ti < 0
    return flow_i | (bdy_i & flow_j) | ((ti == RIGID_T) & (flow_j | (tj == RIGID_T)))


def compare_neighbor_lists(eng, slot, o32, o64=None, fp64_stride=1, log=print):
    """The engine's recorded neighbour lists against the oracle's neighbour
    walk on the same (sorted) state:

    * float32 oracle: for every particle the same neighbours in the same
      order (ids and stencil codes, bit-exact), after the type filter of the
      lists; a particle whose list overflowed (count -1) must have more
      neighbours than any list holds;
    * FP64 oracle (optional, every ``fp64_stride``-th particle): the
      neighbour sets may differ only in knife-edge pairs, |x_ij| within 4
      float32 ulps of r_s; their count is returned and logged.

    Returns (entries compared, knife-edge pairs, differing pairs vs FP64)."""
    cg, ids_g, codes_g = eng.neighbor_lists(slot)
    cnt, ids_o, codes_o = o32.neighbor_csr()
    n = cnt.size
    assert cg.size == n
    typ = o32.get("type")
    id0 = o32.get("id0")
    type_by_id0 = np.empty(n, dtype=typ.dtype)
    type_by_id0[id0] = typ
    owner = np.repeat(np.arange(n), cnt)
    keep = _list_keep(typ[owner], type_by_id0[ids_o])
    kcnt = np.bincount(owner[keep], minlength=n)
    over = cg < 0
    cap = 160 if o32.desc["dim"] == 3 else 48  # entries a list of the engine holds
    assert np.all(kcnt[over] > cap), "a list overflowed below its capacity"
    np.testing.assert_array_equal(cg[~over], kcnt[~over], err_msg="neighbour counts")
    sel = keep & ~over[owner]
    np.testing.assert_array_equal(ids_g, ids_o[sel], err_msg="neighbour ids / order")
    np.testing.assert_array_equal(codes_g, codes_o[sel], err_msg="stencil codes")
    knife = diff = 0
    if o64 is not None:
        c64, ids64, _ = o64.neighbor_csr()
        rs = o32.desc["support_radius"]
        ulp4 = 4.0 * float(np.spacing(np.float32(rs)))
        pos = o64.get("pos")
        pos_by_id0 = np.empty_like(pos)
        pos_by_id0[o64.get("id0")] = pos
        assert np.array_equal(o64.get("id0"), id0)
        pick = np.zeros(n, dtype=bool)
        pick[::fp64_stride] = True
        own32 = owner
        own64 = np.repeat(np.arange(n), c64)
        a = own32[pick[own32]].astype(np.int64) * n + ids_o[pick[own32]]
        b = own64[pick[own64]].astype(np.int64) * n + ids64[pick[own64]]
        x = np.setxor1d(a, b, assume_unique=True)
        diff = x.size
        both = np.union1d(a, b)
        i_own, j_id = both // n, both % n
        r = np.linalg.norm(pos[i_own] - pos_by_id0[j_id], axis=1)
        edge = np.abs(r - rs) <= ulp4
        knife = int(edge.sum())
        if diff:
            xi, xj = x // n, x % n
            rx = np.linalg.norm(pos[xi] - pos_by_id0[xj], axis=1)
            assert np.all(np.abs(rx - rs) <= ulp4), \
                f"{int(np.sum(np.abs(rx - rs) > ulp4))} pairs differ from FP64 away from r_s"
    log(f"neighbour lists: {ids_g.size} entries bit-exact vs float32 oracle, {int(over.sum())} "
        f"overflowed; vs FP64: {diff} differing pairs, all among {knife} knife-edge pairs "
        f"(|r - r_s| <= 4 ulp)")
    return ids_g.size, knife, diff
