"""ctypes binding of libtisphi_hip.so (the C ABI of include/tisphi.h).

The library is the only compute path: if it is missing, or no HIP device is
visible, every entry point raises instead of falling back to anything else.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtisphi_hip.so")

OK = 0
ERR_INVALID_ARG, ERR_OOM, ERR_HIP, ERR_ESCAPED, ERR_NAN, ERR_UNSUPPORTED, ERR_CAPACITY = (
    -1, -2, -3, -4, -5, -6, -7)

# name -> (id, host dtype, components)
FIELDS = dict(pos=(0, "f8", 3), vel=(1, "f8", 3), density=(2, "f8", 1), mass=(3, "f8", 1),
              m_V=(4, "f8", 1), pressure=(5, "f8", 1), type=(6, "i4", 1), id0=(7, "i4", 1),
              obj_id=(8, "i4", 1), stress=(9, "f8", 6), stress_tmp=(10, "f8", 6),
              d_stress=(11, "f8", 6), strain_equ=(12, "f8", 1), strain_equ_p=(13, "f8", 1),
              d_strain_equ=(14, "f8", 1), d_strain_equ_p=(15, "f8", 1),
              flag_retmap=(16, "i4", 1), cspm_f=(17, "f8", 1), d_density=(18, "f8", 1),
              d_vel=(19, "f8", 3), density_tmp=(20, "f8", 1), v_tmp=(21, "f8", 3),
              cell=(22, "i4", 1), offset=(23, "f8", 3), cell_start=(24, "i4", 1),
              cspm_L=(25, "f8", 9), val=(26, "f8", 1), color=(27, "f8", 3), x0=(28, "f8", 3),
              color0=(29, "f8", 3))
PASSES = dict(sort=0, shepard=1, init_tmp=2, one_step=3, lf_half=4, lf_full=5,
              advect_pos=6, advect_something=7, step=8, se=9, enforce_boundary=10, solve_rigid=11)

EXPORTS = ["tsph_abi_version", "tsph_last_error", "tsph_device_count", "tsph_engine_create",
           "tsph_engine_destroy", "tsph_engine_bytes", "tsph_add_scene", "tsph_remove_scene",
           "tsph_upload", "tsph_download", "tsph_step", "tsph_sync", "tsph_run_pass",
           "tsph_scene_steps", "tsph_set_graph", "tsph_set_block", "tsph_set_neighbor_list",
           "tsph_set_cell_tile", "tsph_globals",
           "tsph_status", "tsph_kernel_time", "tsph_timing", "tsph_neighbor_stats",
           "tsph_neighbor_lists", "tsph_neighbor_stats_noalloc", "tsph_check_sqrt", "tsph_create",
           "tsph_destroy", "tsph_assign_value_color", "tsph_v_maxmin", "tsph_set_color",
           "tsph_set_rigid_object"]


def check_sqrt():
    """(differences from sqrtf at arguments >= 2^-96, out-of-range results
    below) of the pair-distance root over every finite non-negative float."""
    out = (ctypes.c_uint64 * 2)()
    check(load().tsph_check_sqrt(out))
    return int(out[0]), int(out[1])


class SceneDesc(ctypes.Structure):
    """tsph_scene_desc."""
    _fields_ = [("dim", ctypes.c_int32), ("method", ctypes.c_int32), ("boundary", ctypes.c_int32),
                ("kernel", ctypes.c_int32), ("kernel_corr", ctypes.c_int32),
                ("time_integration", ctypes.c_int32), ("xsph", ctypes.c_int32),
                ("grid_num", ctypes.c_int32 * 3), ("n_particles", ctypes.c_int64),
                ("vdomain_start", ctypes.c_double * 3), ("domain_start", ctypes.c_double * 3),
                ("domain_end", ctypes.c_double * 3), ("grid_size", ctypes.c_double),
                ("particle_radius", ctypes.c_double), ("particle_diameter", ctypes.c_double),
                ("smoothing_len", ctypes.c_double), ("support_radius", ctypes.c_double),
                ("gravity", ctypes.c_double * 3), ("dt", ctypes.c_double),
                ("density0", ctypes.c_double), ("viscosity", ctypes.c_double),
                ("stiffness", ctypes.c_double), ("exponent", ctypes.c_double),
                ("cohesion", ctypes.c_double), ("friction", ctypes.c_double),
                ("young", ctypes.c_double), ("poisson", ctypes.c_double),
                ("dilatancy", ctypes.c_double), ("vsound", ctypes.c_double)]

    @classmethod
    def from_dict(cls, d):
        st = cls()
        for name, _ in cls._fields_:
            f = getattr(st, name)
            if isinstance(f, ctypes.Array):
                for k in range(len(f)):
                    f[k] = d[name][k]
            else:
                setattr(st, name, d[name])
        return st


class TsphError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"tisphi_hip error {code}: {msg}")
        self.code = code


class ParticleEscaped(TsphError):
    pass


class NaNDetected(TsphError):
    pass


_lib = None


def load():
    """Load the engine library (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run tisphi_amd.build.build()")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64, dp = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.POINTER(ctypes.c_double)
    sig = {
        "tsph_abi_version": ([], i32),
        "tsph_last_error": ([], ctypes.c_char_p),
        "tsph_device_count": ([ctypes.POINTER(i32)], i32),
        "tsph_engine_create": ([i32, i32, i64, i64, i32, ctypes.POINTER(vp)], i32),
        "tsph_engine_destroy": ([vp], i32),
        "tsph_engine_bytes": ([vp, ctypes.POINTER(i64)], i32),
        "tsph_add_scene": ([vp, ctypes.POINTER(SceneDesc), ctypes.POINTER(i32)], i32),
        "tsph_remove_scene": ([vp, i32], i32),
        "tsph_upload": ([vp, i32, i32, vp, ctypes.c_size_t], i32),
        "tsph_download": ([vp, i32, i32, vp, ctypes.c_size_t], i32),
        "tsph_step": ([vp, i32], i32),
        "tsph_sync": ([vp], i32),
        "tsph_run_pass": ([vp, i32], i32),
        "tsph_scene_steps": ([vp, i32, ctypes.POINTER(i64)], i32),
        "tsph_set_graph": ([vp, i32], i32),
        "tsph_set_block": ([vp, i32, i32], i32),
        "tsph_set_neighbor_list": ([vp, i32], i32),
        "tsph_set_cell_tile": ([vp, i32], i32),
        "tsph_globals": ([vp, i32, dp], i32),
        "tsph_status": ([vp, i32, ctypes.POINTER(i64)], i32),
        "tsph_kernel_time": ([vp, ctypes.c_char_p, dp, ctypes.POINTER(i64)], i32),
        "tsph_timing": ([vp, i32], i32),
        "tsph_neighbor_stats": ([vp, i32, vp, vp, vp], i32),
        "tsph_neighbor_lists": ([vp, i32, vp, vp, vp, i64, ctypes.POINTER(i64)], i32),
        "tsph_neighbor_stats_noalloc": ([vp, i32, vp, vp, vp], i32),
        "tsph_check_sqrt": ([ctypes.POINTER(ctypes.c_uint64)], i32),
        "tsph_create": ([ctypes.POINTER(SceneDesc), ctypes.POINTER(vp)], i32),
        "tsph_destroy": ([vp], i32),
        "tsph_assign_value_color": ([vp, i32, i32, i32, ctypes.c_uint32], i32),
        "tsph_v_maxmin": ([vp, i32, ctypes.c_double, ctypes.c_double, i32, i32, dp, dp], i32),
        "tsph_set_color": ([vp, i32, i32, i32], i32),
        "tsph_set_rigid_object": ([vp, i32, i32, dp], i32),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def check(code):
    if code == OK:
        return
    msg = load().tsph_last_error().decode(errors="replace")
    if code == ERR_ESCAPED:
        raise ParticleEscaped(code, msg)
    if code == ERR_NAN:
        raise NaNDetected(code, msg)
    if code == ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    if code == ERR_INVALID_ARG:
        raise ValueError(msg)
    raise TsphError(code, msg)


def device_count():
    n = ctypes.c_int(0)
    check(load().tsph_device_count(ctypes.byref(n)))
    return n.value


class Engine:
    """A set of resident scenes of one method and dimension on one GPU."""

    def __init__(self, method, dim, particle_capacity, cell_capacity, max_scenes=1):
        L = load()
        if device_count() < 1:
            raise RuntimeError("no HIP device visible: the SPH engine runs on the GPU only")
        h = ctypes.c_void_p()
        check(L.tsph_engine_create(int(method), int(dim), int(particle_capacity),
                                   int(cell_capacity), int(max_scenes), ctypes.byref(h)))
        self.h = h
        self.method, self.dim = int(method), int(dim)
        self.n = {}
        self.ncell = {}

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            load().tsph_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def bytes(self):
        v = ctypes.c_int64()
        check(load().tsph_engine_bytes(self.h, ctypes.byref(v)))
        return v.value

    def add_scene(self, desc):
        st = SceneDesc.from_dict(desc)
        slot = ctypes.c_int(-1)
        check(load().tsph_add_scene(self.h, ctypes.byref(st), ctypes.byref(slot)))
        self.n[slot.value] = int(desc["n_particles"])
        self.ncell[slot.value] = int(np.prod(desc["grid_num"]))
        return slot.value

    def remove_scene(self, slot):
        check(load().tsph_remove_scene(self.h, int(slot)))
        self.n.pop(slot, None)

    def upload(self, slot, name, arr):
        fid, dt, k = FIELDS[name]
        a = np.ascontiguousarray(arr, dtype=dt).reshape(-1)
        if a.size != self.n[slot] * k:
            raise ValueError(f"{name}: expected {self.n[slot] * k} values, got {a.size}")
        check(load().tsph_upload(self.h, int(slot), fid, a.ctypes.data, a.nbytes))

    def download(self, slot, name):
        fid, dt, k = FIELDS[name]
        cnt = self.ncell[slot] + 1 if name == "cell_start" else self.n[slot] * k
        a = np.zeros(cnt, dtype=dt)
        check(load().tsph_download(self.h, int(slot), fid, a.ctypes.data, a.nbytes))
        return a.reshape(-1, k) if k > 1 else a

    def step(self, n=1):
        check(load().tsph_step(self.h, int(n)))

    def sync(self):
        check(load().tsph_sync(self.h))

    def run(self, pass_name):
        check(load().tsph_run_pass(self.h, PASSES[pass_name]))

    def steps(self, slot):
        v = ctypes.c_int64()
        check(load().tsph_scene_steps(self.h, int(slot), ctypes.byref(v)))
        return v.value

    def set_graph(self, on):
        check(load().tsph_set_graph(self.h, 1 if on else 0))

    def set_neighbor_list(self, mode):
        """0/False: stencil walk in every pass; 1/True: per-step lists;
        2: per-step lists built from LDS-staged cell windows."""
        check(load().tsph_set_neighbor_list(self.h, int(mode)))

    def set_block(self, pair, elem):
        check(load().tsph_set_block(self.h, int(pair), int(elem)))

    def set_cell_tile(self, stage):
        """Candidates one wave stages in LDS per row offset in the mode-2
        list build: 128, 256 (default) or 512."""
        check(load().tsph_set_cell_tile(self.h, int(stage)))

    def globals(self, slot=0):
        out = (ctypes.c_double * 4)()
        check(load().tsph_globals(self.h, int(slot), out))
        return np.array(list(out))

    def status(self, slot=0):
        out = (ctypes.c_int64 * 2)()
        check(load().tsph_status(self.h, int(slot), out))
        return int(out[0]), int(out[1])

    def timing(self, on):
        check(load().tsph_timing(self.h, 1 if on else 0))

    def neighbor_stats(self, slot=0):
        n = self.n[slot]
        c = np.zeros(n, dtype=np.int32)
        s1 = np.zeros(n, dtype=np.int64)
        s2 = np.zeros(n, dtype=np.int64)
        check(load().tsph_neighbor_stats(self.h, int(slot), c.ctypes.data, s1.ctypes.data,
                                         s2.ctypes.data))
        return c, s1, s2

    def neighbor_stats_noalloc(self, slot=0):
        """neighbor_stats through the entry point that allocates nothing
        (a slot of at most 262,144 particles)."""
        n = self.n[slot]
        c = np.zeros(n, dtype=np.int32)
        s1 = np.zeros(n, dtype=np.int64)
        s2 = np.zeros(n, dtype=np.int64)
        check(load().tsph_neighbor_stats_noalloc(self.h, int(slot), c.ctypes.data, s1.ctypes.data,
                                                 s2.ctypes.data))
        return c, s1, s2

    def neighbor_lists(self, slot=0):
        """(count, ids, codes) of the recorded neighbour lists of ``slot``:
        count[i] entries of particle i (-1: it walks the stencil), the
        neighbours' id0 and stencil codes concatenated in list order."""
        n = self.n[slot]
        c = np.zeros(n, dtype=np.int32)
        total = ctypes.c_int64(0)
        st = load().tsph_neighbor_lists(self.h, int(slot), c.ctypes.data, None, None, 0,
                                        ctypes.byref(total))
        if st not in (0, -1):
            check(st)
        ids = np.zeros(max(total.value, 1), dtype=np.int32)
        codes = np.zeros(max(total.value, 1), dtype=np.int8)
        check(load().tsph_neighbor_lists(self.h, int(slot), c.ctypes.data, ids.ctypes.data,
                                         codes.ctypes.data, ids.size, ctypes.byref(total)))
        return c, ids[:total.value], codes[:total.value]

    def set_rigid_object(self, slot, obj, rest_cm):
        """Register a dynamic rigid object (isDynamic block) of ``slot``."""
        cm = (ctypes.c_double * 3)(*[float(v) for v in rest_cm])
        check(load().tsph_set_rigid_object(self.h, int(slot), int(obj), cm))

    # colour feed (solver.assign_value_color, ps.v_maxmin, ps.set_color)
    def assign_value_color(self, slot, color_title, show_bdy=False, seed=0):
        check(load().tsph_assign_value_color(self.h, int(slot), int(color_title),
                                             1 if show_bdy else 0, int(seed) & 0xFFFFFFFF))

    def v_maxmin(self, slot, givenmax=-1.0, givenmin=-1.0, fixmax=0, fixmin=0):
        hi, lo = ctypes.c_double(), ctypes.c_double()
        check(load().tsph_v_maxmin(self.h, int(slot), float(givenmax), float(givenmin),
                                   1 if fixmax else 0, 1 if fixmin else 0, ctypes.byref(hi),
                                   ctypes.byref(lo)))
        return hi.value, lo.value

    def set_color(self, slot, color_group, show_bdy=False):
        check(load().tsph_set_color(self.h, int(slot), int(color_group), 1 if show_bdy else 0))

    def kernel_time(self, name):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        check(load().tsph_kernel_time(self.h, name.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value


def upload_scene(eng, slot, scene, id0=None):
    """Upload a generated Scene (generation order) into ``slot``."""
    eng.upload(slot, "pos", scene.pos)
    eng.upload(slot, "vel", scene.vel)
    eng.upload(slot, "density", scene.density)
    eng.upload(slot, "mass", scene.mass)
    eng.upload(slot, "m_V", scene.m_V)
    eng.upload(slot, "type", scene.mat_type)
    eng.upload(slot, "id0", np.arange(scene.n, dtype=np.int32) if id0 is None else id0)
    eng.upload(slot, "obj_id", scene.obj_id)
    if scene.method != 1:
        eng.upload(slot, "stress", scene.stress)
    dyn = [r for r in getattr(scene, "rigid", []) if r["dynamic"]]
    if dyn:
        # shape matching reads the initial positions (by id0)
        eng.upload(slot, "x0", scene.pos)
        for r in dyn:
            eng.set_rigid_object(slot, r["obj"], r["rest_cm"])
