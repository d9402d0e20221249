"""ParticleSystem with the reference's interface, backed by the HIP engine.

Reference: eng/particle_system.py.  Geometry and particle generation run on
the host in float64 exactly as in the reference (particle_system.py:12-133);
particle storage, the cell hash, prefix sum and counting sort
(``initialize_particle_system``, :254-257) and every neighbour loop run in
libtisphi_hip.so on the GPU.  The Taichi struct field ``pt`` becomes an
accessor that downloads one field at a time as numpy arrays.
"""
import numpy as np

from .. import _lib
from ..scene import build_scene
from .particle_func import MAT_DUMMY, MAT_FLUID, MAT_REP, MAT_RIGID, MAT_SOIL


class ScalarField:
    """Stand-in for a 0-d Taichi field: ``f[None]`` reads / writes a value."""

    def __init__(self, getter, setter=None):
        self._get, self._set = getter, setter

    def __getitem__(self, idx):
        return self._get()

    def __setitem__(self, idx, value):
        if self._set is None:
            raise AttributeError("read-only field")
        self._set(value)


class FieldView:
    """Snapshot of one particle field (``to_numpy()`` like a Taichi field)."""

    def __init__(self, arr):
        self._a = arr

    def to_numpy(self):
        return self._a.copy()

    def __array__(self, dtype=None, copy=None):
        return self._a if dtype is None else self._a.astype(dtype)

    def __getitem__(self, i):
        return self._a[i]

    def __len__(self):
        return len(self._a)

    @property
    def shape(self):
        return self._a.shape


def _stress33(s6):
    """(n, 6) xx,yy,zz,xy,yz,zx -> (n, 3, 3)."""
    n = s6.shape[0]
    m = np.zeros((n, 3, 3))
    m[:, 0, 0], m[:, 1, 1], m[:, 2, 2] = s6[:, 0], s6[:, 1], s6[:, 2]
    m[:, 0, 1] = m[:, 1, 0] = s6[:, 3]
    m[:, 1, 2] = m[:, 2, 1] = s6[:, 4]
    m[:, 2, 0] = m[:, 0, 2] = s6[:, 5]
    return m


def record_objects(blocks, bodies, sc):
    """Object collection by object id, with each block's and body's
    particle count in its "particleNum" (particle_system.py:72-91)."""
    coll = {}
    for b, n in zip(blocks, sc.block_counts):
        b["particleNum"] = n
        coll[b["objectId"]] = b
    for b, n in zip(bodies, sc.body_counts):
        b["particleNum"] = n
        coll[b["objectId"]] = b
    return coll


class _ParticleAccessor:
    """``ps.pt.<member>`` -> FieldView of the current device state."""

    _map = {"x": "pos", "v": "vel", "density": "density", "mass": "mass", "m_V": "m_V",
            "pressure": "pressure", "mat_type": "type", "id0": "id0", "obj_id": "obj_id",
            "strain_equ": "strain_equ", "strain_equ_p": "strain_equ_p",
            "d_strain_equ": "d_strain_equ", "d_strain_equ_p": "d_strain_equ_p",
            "flag_retmap": "flag_retmap", "CSPM_f": "cspm_f", "d_density": "d_density",
            "d_vel": "d_vel", "density_tmp": "density_tmp", "v_tmp": "v_tmp",
            "grid_ids": "cell", "val": "val", "color": "color"}
    _mat = {"stress": "stress", "stress_tmp": "stress_tmp", "d_stress": "d_stress"}

    def __init__(self, ps):
        self._ps = ps

    def __getattr__(self, name):
        ps = self._ps
        if name in self._map:
            return FieldView(ps.engine.download(ps.slot, self._map[name]))
        if name in self._mat:
            return FieldView(_stress33(ps.engine.download(ps.slot, self._mat[name])))
        if name == "mat_id":
            return FieldView(ps._mat_id_now())
        raise AttributeError(name)


class ParticleSystem:
    def __init__(self, config, engine=None, slot=None):
        """``engine``/``slot``: place the scene in an existing engine (batched
        use); by default a one-scene engine is created."""
        self.cfg = config
        sc = build_scene(config)
        self.scene = sc
        g = sc.geo
        self.dim = g.dim
        self.dim3 = 3
        self.domain_start = g.domain_start
        self.domain_end = g.domain_end
        self.domain_size = g.domain_size
        self.color_title = config.get_cfg_default("colorTitle", 0)
        self.color_group = config.get_cfg_default("colorGroup", 0)
        self.flag_boundary = config.get_cfg("boundary")
        self.show_bdy = config.get_cfg_default("showBdyPts", False)
        self.mat_dummy_type, self.mat_rep_type = MAT_DUMMY, MAT_REP
        self.mat_fluid_type, self.mat_soil_type, self.mat_rigid_type = MAT_FLUID, MAT_SOIL, MAT_RIGID
        self.bdy_none, self.bdy_collision, self.bdy_dummy, self.bdy_rep, self.bdy_dummy_rep = 0, 1, 2, 3, 4
        self.particle_radius = g.particle_radius
        self.kappa = g.kappa
        self.kh = g.kh
        self.particle_diameter = g.particle_diameter
        self.smoothing_len = g.smoothing_len
        self.support_radius = g.support_radius
        self.m_V0 = g.m_V0
        self.grid_size = g.grid_size
        self.vdomain_start = g.vdomain_start
        self.vdomain_end = g.vdomain_end
        self.vdomain_size = g.vdomain_size
        self.grid_num = g.grid_num
        self.grid_num_total = g.grid_num_total
        self.mat_index, self.mat_fluid, self.mat_soil, self.mat_rigid = (
            sc.mat_index, sc.mat_fluid, sc.mat_soil, sc.mat_rigid)
        self.blocks = config.get_blocks()
        self.bodies = config.get_bodies()
        self.object_collection = record_objects(self.blocks, self.bodies, sc)
        self.object_id_rigid = {r["obj"] for r in sc.rigid}
        self.rigid_rest_cm = {r["obj"]: r["rest_cm"] for r in sc.rigid if r["dynamic"]}
        self.particle_max_num = sc.n
        self._vmax, self._vmin = 0.0, 0.0
        self._colour_inputs = False
        self._colour_calls = 0
        self.vmax = ScalarField(lambda: self._vmax)
        self.vmin = ScalarField(lambda: self._vmin)
        self.particle_num = ScalarField(lambda: self.particle_max_num)
        if engine is None:
            engine = _lib.Engine(sc.method, g.dim, sc.n, g.grid_num_total, 1)
        self.engine = engine
        self.slot = engine.add_scene(sc.desc()) if slot is None else slot
        self.pt = _ParticleAccessor(self)
        self.initialize_particles()

    # ------------------------------------------------------------- particles
    def initialize_particles(self):
        """Upload the generated lattice (blocks, then dummy walls)."""
        _lib.upload_scene(self.engine, self.slot, self.scene)

    def save_state(self, path, step=0):
        """Write this scene's state after ``step`` steps to an ``.npz``
        snapshot (an engine addition: the reference has no checkpoint)."""
        from .. import snapshot
        return snapshot.save(path, self.engine, self.slot, self.scene, step)

    def load_state(self, path):
        """Replace this scene's state by a snapshot of the same scene;
        returns the snapshot's header (step count, dt, config)."""
        from .. import snapshot
        meta, arrays = snapshot.load(path)
        snapshot.restore(self.engine, self.slot, meta, arrays)
        return meta

    def set_id0(self):
        self.engine.upload(self.slot, "id0", np.arange(self.particle_max_num, dtype=np.int32))

    def initialize_particle_system(self):
        """update_grid_id + prefix sum + counting sort (one device pass)."""
        self.engine.run("sort")

    def _mat_id_now(self):
        order = self.engine.download(self.slot, "id0")
        return self.scene.mat_id[order]

    # ------------------------------------------------------------ read-back
    def _upload_colour_inputs(self):
        """Initial positions (colour title 44) and initial colours (kept by
        particles outside the colour group), indexed by id0, once."""
        if not self._colour_inputs:
            self.engine.upload(self.slot, "x0", self.scene.pos)
            self.engine.upload(self.slot, "color0", self.scene.color0.astype(np.float64))
            self._colour_inputs = True

    def v_maxmin(self, givenmax=-1, givenmin=-1, fixmax=0, fixmin=0):
        """Max/min of the colour value over flow particles, then the
        given/fix rule (particle_system.py:387-397), on the device."""
        self._upload_colour_inputs()
        self._vmax, self._vmin = self.engine.v_maxmin(self.slot, givenmax, givenmin, fixmax, fixmin)

    def set_color(self):
        """Jet colour of the colour group (particle_system.py:399-407), on
        the device (``ps.pt.color`` reads it)."""
        self._upload_colour_inputs()
        self.engine.set_color(self.slot, self.color_group, self.show_bdy)

    def dump(self):
        """(positions, data) dictionaries of particle_system.py:459-545, in
        the current particle order."""
        e, s = self.engine, self.slot
        pos = e.download(s, "pos")
        vel = e.download(s, "vel")
        st = e.download(s, "stress")
        return {
            "pos.x": pos[:, 0].copy(), "pos.y": pos[:, 1].copy(), "pos.z": pos[:, 2].copy(),
        }, {
            "id0": e.download(s, "id0").astype(np.int64),
            "objId": e.download(s, "obj_id").astype(np.int64),
            "material": e.download(s, "type").astype(np.int64),
            "density": e.download(s, "density"),
            "vel.x": vel[:, 0].copy(), "vel.y": vel[:, 1].copy(), "vel.z": vel[:, 2].copy(),
            "vel.norm": np.linalg.norm(vel, axis=1),
            "stress.xx": st[:, 0].copy(), "stress.yy": st[:, 1].copy(), "stress.zz": st[:, 2].copy(),
            "stress.xy": st[:, 3].copy(), "stress.yz": st[:, 4].copy(), "stress.zx": st[:, 5].copy(),
            "stress.hydro": (st[:, 0] + st[:, 1] + st[:, 2]) / 3.0,
            "strain_equ": e.download(s, "strain_equ"),
            "strain_equ_p": e.download(s, "strain_equ_p"),
            "pressure": e.download(s, "pressure"),
            "plas_behav": e.download(s, "flag_retmap").astype(np.float64),
        }

    def globals(self):
        """FP64 (surge front, potential energy, kinetic energy, mass)."""
        return self.engine.globals(self.slot)
