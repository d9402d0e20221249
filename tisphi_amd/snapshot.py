"""Scene state snapshots: save a resident scene to an ``.npz`` file and
restore it into an engine slot, so a run stops and resumes with the same
bits as one that never stopped.

The reference has no checkpoint (its reset is broken, eng/ui_sim.py:202-209,
and its export is write-only, :293-304); this is an engine addition.  A
snapshot holds every field a step reads from the previous one, in the
engine's current particle order (cell id + float32 offset, no rounding
through absolute positions), including the temporaries a solver carries
over (the μ(I) soil loop reads the boundary particles' v_tmp of the step
before, solver_sph_muI.py:62-132), the initial positions of shape matching,
the dynamic rigid objects' rest centres, the scene JSON and the step count.
Loading uses ``numpy.load(allow_pickle=False)``: a snapshot is arrays and a
JSON header, nothing that executes.
"""
import json

import numpy as np

FORMAT = 1

STATE = ["cell", "offset", "vel", "density", "mass", "m_V", "type", "id0", "obj_id", "pressure",
         "density_tmp", "v_tmp", "cspm_f", "d_density", "d_vel"]
SOIL_STATE = ["stress", "stress_tmp", "strain_equ", "d_strain_equ"]
DP_STATE = ["strain_equ_p", "flag_retmap", "d_stress", "d_strain_equ_p"]


def state_fields(method, dynamic_rigid):
    """Engine fields of a snapshot for a scene of ``method`` (1 WC, 2 μ(I),
    3 DP); ``x0`` (initial positions by id0) when shape matching runs."""
    f = list(STATE)
    if method != 1:
        f += SOIL_STATE
    if method == 3:
        f += DP_STATE
    if dynamic_rigid:
        f.append("x0")
    return f


# keys the loaders add to Blocks / Bodies entries (particle counts, the
# rest mesh of a dynamic body); rebuilt when the scene is built again
DERIVED = ("particleNum", "restPosition", "restCenterOfMass", "voxelizedPoints", "mesh")


def scene_json(config):
    """The scene JSON without the keys the loaders derive from it."""
    out = dict(config)
    for sec in ("Blocks", "Bodies"):
        if sec in out:
            out[sec] = [{k: v for k, v in b.items() if k not in DERIVED} for b in out[sec]]
    return out


def save(path, eng, slot, scene, step=0):
    """Write the state of ``slot`` (built from ``scene``) after ``step``
    steps to ``path``; returns the path written."""
    rigid = [dict(obj=int(r["obj"]), dynamic=bool(r["dynamic"]),
                  rest_cm=[float(v) for v in r["rest_cm"]]) for r in scene.rigid]
    dyn = any(r["dynamic"] for r in rigid)
    eng.sync()
    arrays = {f: eng.download(slot, f) for f in state_fields(scene.method, dyn)}
    meta = dict(format=FORMAT, method=int(scene.method), dim=int(scene.geo.dim), n=int(scene.n),
                step=int(step), dt=float(scene.dt), rigid=rigid, config=scene_json(scene.cfg.config),
                scene_file=getattr(scene.cfg, "scene_file_path", None))
    header = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    with open(path, "wb") as f:
        np.savez(f, __meta__=header, **arrays)
    return path


def load(path):
    """(meta, arrays) of a snapshot file."""
    with np.load(path, allow_pickle=False) as z:
        if "__meta__" not in z.files:
            raise ValueError(f"{path}: not a scene snapshot")
        meta = json.loads(bytes(z["__meta__"]).decode())
        arrays = {k: z[k] for k in z.files if k != "__meta__"}
    if meta.get("format") != FORMAT:
        raise ValueError(f"{path}: snapshot format {meta.get('format')}, expected {FORMAT}")
    want = state_fields(meta["method"], any(r["dynamic"] for r in meta["rigid"]))
    missing = [f for f in want if f not in arrays]
    if missing:
        raise ValueError(f"{path}: missing fields {missing}")
    return meta, arrays


def restore(eng, slot, meta, arrays):
    """Upload a loaded snapshot into ``slot`` (a scene of the same method,
    dimension and particle count) and register its dynamic rigid objects."""
    if int(eng.n[slot]) != int(meta["n"]):
        raise ValueError(f"snapshot holds {meta['n']} particles, the slot {eng.n[slot]}")
    if int(eng.method) != int(meta["method"]) or int(eng.dim) != int(meta["dim"]):
        raise ValueError("snapshot of another method or dimension")
    for f in state_fields(meta["method"], any(r["dynamic"] for r in meta["rigid"])):
        eng.upload(slot, f, arrays[f])
    for r in meta["rigid"]:
        if r["dynamic"]:
            eng.set_rigid_object(slot, r["obj"], r["rest_cm"])
