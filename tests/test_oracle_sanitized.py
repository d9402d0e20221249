"""The CPU oracle under AddressSanitizer + UndefinedBehaviorSanitizer.

oracle/sanitize_main.cpp drives the oracle from a file; built with
``make -C oracle sanitize`` (every UB report fatal), it steps small scenes
of every solver and option family and must exit cleanly and compute the
same state as the ordinary -O3 library on the same input (per-particle
fields bitwise: both builds make the same IEEE roundings, contraction
off; FP64 global sums to 1e-12, their OpenMP summation order may differ).
"""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from tests.common import add_rigid_blocks, scene_cfg, tiny_cfg
from tisphi_amd.scene import build_scene

ORACLE_DIR = os.path.dirname(os.path.abspath(O.__file__))

INPUTS = ["pos", "vel", "density", "mass", "m_V", "type", "id0", "obj_id", "stress", "x0"]
OUTPUTS = ["cell", "offset", "id0", "vel", "density", "pressure", "d_vel", "d_density", "stress",
           "strain_equ", "flag_retmap"]


def scenes():
    out = {}
    out["wc2d_xsph_rk4"] = build_scene(tiny_cfg(xsph=True, ti=4))
    out["wc2d_cspm_se_cubic"] = build_scene(tiny_cfg(ti=1, kcorr=1, kernel=0))
    out["wc2d_dummy_rep"] = build_scene(tiny_cfg(boundary=4))
    out["wc2d_collision"] = build_scene(tiny_cfg(boundary=1))
    out["mui2d"] = build_scene(tiny_cfg(method=2, r=0.002, domain=(0.1, 0.06, 0.01),
                                        block=(0.03, 0.02, 0.01)))
    cfg = tiny_cfg(method=3, r=0.002, domain=(0.1, 0.06, 0.01), block=(0.03, 0.02, 0.01))
    add_rigid_blocks(cfg, [(10, (0.05, 0.005, 0.0), (0.01, 0.01, 0.01), True)])
    out["dp2d_rigid"] = build_scene(cfg)
    cfg = tiny_cfg(dim=3, domain=(0.3, 0.2, 0.12), block=(0.1, 0.08, 0.06))
    add_rigid_blocks(cfg, [(10, (0.16, 0.02, 0.03), (0.04, 0.04, 0.04), True)])
    out["wc3d_rigid"] = build_scene(cfg)
    return out


def desc_bytes(sc):
    st = O.fill_struct(O.OrcParams(), dict(sc.desc()))
    return bytes(st)


def write_input(path, sc, steps):
    fields = {"pos": sc.pos, "vel": sc.vel, "density": sc.density, "mass": sc.mass, "m_V": sc.m_V,
              "type": sc.mat_type, "id0": np.arange(sc.n, dtype=np.int32), "obj_id": sc.obj_id,
              "stress": sc.stress, "x0": sc.pos}
    dyn = [r for r in sc.rigid if r["dynamic"]]
    with open(path, "wb") as f:
        f.write(desc_bytes(sc))
        f.write(struct.pack("<iii", 32, steps, len(dyn)))
        for r in dyn:
            f.write(struct.pack("<i3d", r["obj"], *[float(v) for v in r["rest_cm"]]))
        f.write(struct.pack("<i", len(INPUTS)))
        for name in INPUTS:
            fid, dt, k = O.FIELDS[name]
            a = np.ascontiguousarray(fields[name], dtype=dt).reshape(-1)
            f.write(struct.pack("<iq", fid, a.nbytes))
            f.write(a.tobytes())
        f.write(struct.pack("<i", len(OUTPUTS)))
        for name in OUTPUTS:
            fid, dt, k = O.FIELDS[name]
            f.write(struct.pack("<iq", fid, sc.n * k * np.dtype(dt).itemsize))


def read_output(path, sc):
    out = {}
    with open(path, "rb") as f:
        for name in OUTPUTS:
            fid, nbytes = struct.unpack("<iq", f.read(12))
            _, dt, k = O.FIELDS[name]
            assert fid == O.FIELDS[name][0]
            a = np.frombuffer(f.read(nbytes), dtype=dt)
            out[name] = a.reshape(sc.n, k) if k > 1 else a
        out["globals"] = np.frombuffer(f.read(32), dtype=np.float64)
    return out


@pytest.fixture(scope="module")
def sanitized(tmp_path_factory):
    if shutil.which(os.environ.get("CXX", "g++")) is None:
        pytest.skip("no C++ compiler")
    d = tmp_path_factory.mktemp("san")
    exe = str(d / "sanitize_main")
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, "sanitize", f"OUT={exe}"], check=True)
    return exe


def test_sanitized_oracle_matches_library(sanitized, tmp_path):
    O.build()
    steps = 25
    env = dict(os.environ, ASAN_OPTIONS="abort_on_error=0:halt_on_error=1:detect_leaks=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="1")
    for name, sc in scenes().items():
        inp, outp = str(tmp_path / f"{name}.in"), str(tmp_path / f"{name}.out")
        write_input(inp, sc, steps)
        r = subprocess.run([sanitized, inp, outp], env=env, capture_output=True, text=True,
                           timeout=600)
        assert r.returncode == 0, f"{name}: exit {r.returncode}\n{r.stderr[-3000:]}"
        assert "runtime error" not in r.stderr and "Sanitizer" not in r.stderr, r.stderr[-3000:]
        got = read_output(outp, sc)
        o = O.Oracle(sc.desc(), 32)
        O.load_scene(o, sc)
        o.step(steps)
        for f in OUTPUTS:
            np.testing.assert_array_equal(got[f], o.get(f), err_msg=f"{name}: {f}")
        np.testing.assert_allclose(got["globals"], o.globals(), rtol=1e-12, err_msg=name)
        assert np.all(np.isfinite(got["vel"])), name
