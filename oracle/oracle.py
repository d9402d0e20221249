"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker and by bench.py's cpu_baseline leg; never by the product package.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

# Field / pass ids: identical numbering to include/tisphi.h.
FIELDS = dict(pos=(0, "f8", 3), vel=(1, "f8", 3), density=(2, "f8", 1), mass=(3, "f8", 1),
              m_V=(4, "f8", 1), pressure=(5, "f8", 1), type=(6, "i4", 1), id0=(7, "i4", 1),
              obj_id=(8, "i4", 1), stress=(9, "f8", 6), stress_tmp=(10, "f8", 6),
              d_stress=(11, "f8", 6), strain_equ=(12, "f8", 1), strain_equ_p=(13, "f8", 1),
              d_strain_equ=(14, "f8", 1), d_strain_equ_p=(15, "f8", 1),
              flag_retmap=(16, "i4", 1), cspm_f=(17, "f8", 1), d_density=(18, "f8", 1),
              d_vel=(19, "f8", 3), density_tmp=(20, "f8", 1), v_tmp=(21, "f8", 3),
              cell=(22, "i4", 1), offset=(23, "f8", 3), cell_start=(24, "i4", 1),
              cspm_L=(25, "f8", 9), x0=(28, "f8", 3))
PASSES = dict(sort=0, shepard=1, init_tmp=2, one_step=3, lf_half=4, lf_full=5,
              advect_pos=6, advect_something=7, step=8, se=9, enforce_boundary=10,
              solve_rigid=11)


class OrcParams(ctypes.Structure):
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


def fill_struct(st, d):
    for name, _ in st._fields_:
        v = d[name]
        f = getattr(st, name)
        if isinstance(f, ctypes.Array):
            for k in range(len(f)):
                f[k] = v[k]
        else:
            setattr(st, name, v)
    return st


_lib = None


def build(force=False):
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "liboracle.so"], check=True)


def use_library(path):
    """Load the oracle from ``path`` (e.g. the host-native CPU-baseline
    build) instead of liboracle.so; takes effect for later calls."""
    global _lib, LIB_PATH
    LIB_PATH = path
    _lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.POINTER(OrcParams), ctypes.c_int]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        for fn in (L.orc_set, L.orc_get):
            fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
            fn.restype = ctypes.c_int
        L.orc_pass.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_globals.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
        L.orc_neighbors.argtypes = [ctypes.c_void_p, ctypes.c_int64,
                                    ctypes.POINTER(ctypes.c_int64), ctypes.c_int64]
        L.orc_neighbors.restype = ctypes.c_int64
        L.orc_neighbor_csr.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_int64]
        L.orc_neighbor_csr.restype = ctypes.c_int64
        L.orc_escaped.argtypes = [ctypes.c_void_p]
        L.orc_escaped.restype = ctypes.c_int64
        L.orc_set_threads.argtypes = [ctypes.c_int]
        P = ctypes.POINTER(OrcParams)
        L.orc_kernel_eval.argtypes = [P, ctypes.c_int, ctypes.c_void_p, ctypes.c_int64,
                                      ctypes.c_void_p, ctypes.c_void_p]
        L.orc_dp_adapt.argtypes = [P, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
        L.orc_dp_rate.argtypes = [P, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_get_threads.restype = ctypes.c_int
        L.orc_set_rigid_object.argtypes = [ctypes.c_void_p, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_double)]
        L.orc_polar.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        _lib = L
    return _lib


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


class Oracle:
    """One scene on the CPU oracle.  ``precision`` 32 mirrors the engine's
    arithmetic; 64 is the FP64 oracle."""

    def __init__(self, desc, precision=32):
        self.desc = dict(desc)
        self.n = int(desc["n_particles"])
        self.C = int(np.prod(desc["grid_num"]))
        self._p = fill_struct(OrcParams(), self.desc)
        self.h = lib().orc_create(ctypes.byref(self._p), int(precision))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def set(self, name, arr):
        fid, dt, k = FIELDS[name]
        a = np.ascontiguousarray(arr, dtype=dt).reshape(-1)
        assert a.size == self.n * k, (name, a.size, self.n * k)
        if lib().orc_set(self.h, fid, a.ctypes.data) != 0:
            raise ValueError(name)

    def get(self, name):
        fid, dt, k = FIELDS[name]
        cnt = (self.C + 1) if name == "cell_start" else self.n * k
        a = np.zeros(cnt, dtype=dt)
        if lib().orc_get(self.h, fid, a.ctypes.data) != 0:
            raise ValueError(name)
        return a.reshape(self.n, k) if k > 1 else a

    def run(self, pass_name):
        if lib().orc_pass(self.h, PASSES[pass_name]) != 0:
            raise ValueError(pass_name)

    def step(self, n=1):
        lib().orc_step(self.h, int(n))

    def globals(self):
        out = (ctypes.c_double * 4)()
        lib().orc_globals(self.h, out)
        return np.array(list(out))

    def neighbors(self, i, cap=4096):
        buf = (ctypes.c_int64 * cap)()
        k = lib().orc_neighbors(self.h, int(i), buf, cap)
        return np.array(buf[:min(k, cap)], dtype=np.int64)

    def neighbor_csr(self):
        """(counts, id0 of each neighbour, stencil codes) of every particle,
        concatenated in particle and walk order."""
        counts = np.zeros(self.n, dtype=np.int64)
        total = lib().orc_neighbor_csr(self.h, counts.ctypes.data, None, None, 0)
        ids = np.zeros(max(total, 1), dtype=np.int32)
        codes = np.zeros(max(total, 1), dtype=np.int8)
        lib().orc_neighbor_csr(self.h, counts.ctypes.data, ids.ctypes.data, codes.ctypes.data, total)
        return counts, ids[:total], codes[:total]

    def escaped(self):
        return lib().orc_escaped(self.h)

    def set_rigid_object(self, obj, rest_cm):
        cm = (ctypes.c_double * 3)(*[float(v) for v in rest_cm])
        if lib().orc_set_rigid_object(self.h, int(obj), cm) != 0:
            raise ValueError("at most 4 dynamic rigid objects")


def load_scene(oracle, scene):
    """Upload a freshly generated tisphi_amd Scene (generation order)."""
    n = scene.n
    oracle.set("pos", scene.pos)
    oracle.set("vel", scene.vel)
    oracle.set("density", scene.density)
    oracle.set("mass", scene.mass)
    oracle.set("m_V", scene.m_V)
    oracle.set("type", scene.mat_type)
    oracle.set("id0", np.arange(n, dtype=np.int32))
    oracle.set("obj_id", scene.obj_id)
    oracle.set("stress", scene.stress)
    oracle.set("x0", scene.pos)
    for r in getattr(scene, "rigid", []):
        if r["dynamic"]:
            oracle.set_rigid_object(r["obj"], r["rest_cm"])


def polar(A, dim):
    """The oracle's orthogonal polar factor of a 3x3 matrix (row-major)."""
    A = np.ascontiguousarray(A, dtype=np.float64).reshape(9)
    R = np.zeros(9)
    lib().orc_polar(A.ctypes.data, int(dim), R.ctypes.data)
    return R.reshape(3, 3)


def kernel_eval(desc, r, precision=64):
    """(W(r), grad factor(r)) of the oracle's smoothing kernel."""
    r = np.ascontiguousarray(r, dtype=np.float64)
    W = np.zeros_like(r)
    G = np.zeros_like(r)
    p = fill_struct(OrcParams(), desc)
    lib().orc_kernel_eval(ctypes.byref(p), int(precision), r.ctypes.data, r.size, W.ctypes.data,
                          G.ctypes.data)
    return W, G


def dp_adapt(desc, st6):
    """(flags of the input state, adapted stress) of the DP return mapping."""
    st = np.ascontiguousarray(np.array(st6, dtype=np.float64).reshape(-1, 6))
    fl = np.zeros(st.shape[0], dtype=np.int32)
    p = fill_struct(OrcParams(), desc)
    lib().orc_dp_adapt(ctypes.byref(p), st.ctypes.data, st.shape[0], fl.ctypes.data)
    return fl, st


def dp_rate(desc, L, st6):
    """(d_stress, d_strain_equ, d_strain_equ_p) of the Bui (2008) update."""
    L = np.ascontiguousarray(np.array(L, dtype=np.float64).reshape(-1, 9))
    st = np.ascontiguousarray(np.array(st6, dtype=np.float64).reshape(-1, 6))
    n = L.shape[0]
    ds = np.zeros((n, 6))
    dq = np.zeros(n)
    dqp = np.zeros(n)
    p = fill_struct(OrcParams(), desc)
    lib().orc_dp_rate(ctypes.byref(p), L.ctypes.data, st.ctypes.data, n, ds.ctypes.data,
                      dq.ctypes.data, dqp.ctypes.data)
    return ds, dq, dqp
