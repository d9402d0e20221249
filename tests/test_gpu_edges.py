"""Edge cases on the GPU: an empty scene beside a full one, a scene of
boundary particles only, one isolated particle in free fall, and the
capacity limits (refused with TSPH_ERR_CAPACITY, the engine stays usable;
out-of-range arguments with TSPH_ERR_INVALID_ARG)."""
import numpy as np
import pytest

from tests.common import scene_cfg, tiny_cfg
from tisphi_amd import _lib as L
from tisphi_amd.scene import build_scene

pytestmark = pytest.mark.gpu

STATE = ["pos", "vel", "density", "pressure", "d_vel", "id0"]


def run_alone(sc, steps):
    eng = L.Engine(sc.method, sc.geo.dim, sc.n, sc.geo.grid_num_total, 1)
    slot = eng.add_scene(sc.desc())
    L.upload_scene(eng, slot, sc)
    eng.step(steps)
    eng.sync()
    out = {f: eng.download(slot, f) for f in STATE}
    out["globals"] = eng.globals(slot)
    eng.close()
    return out


def test_empty_scene_beside_a_full_one():
    """A scene with no particles takes a slot, steps with the others and
    changes nothing: the full scene's state is bitwise the one it reaches
    alone; the empty scene's globals are the empty reductions."""
    sc = build_scene(tiny_cfg())
    alone = run_alone(sc, 20)
    d0 = sc.desc()
    d0["n_particles"] = 0
    eng = L.Engine(sc.method, 2, sc.n + 64, 2 * sc.geo.grid_num_total, 3)
    s0 = eng.add_scene(d0)
    s1 = eng.add_scene(sc.desc())
    L.upload_scene(eng, s1, sc)
    eng.step(20)
    eng.sync()
    for f in STATE:
        np.testing.assert_array_equal(eng.download(s1, f), alone[f], err_msg=f)
    g0 = eng.globals(s0)
    assert g0[0] == -np.inf and g0[1] == 0.0 and g0[2] == 0.0 and g0[3] == 0.0
    assert eng.download(s0, "pos").shape == (0, 3)
    assert eng.status(s0) == (0, 0) and eng.status(s1) == (0, 0)
    eng.close()


def test_boundary_only_scene():
    """No flow particle: the dummy walls stay where they are, nothing is
    flagged, the flow reductions are empty."""
    cfg = scene_cfg("dam_break_2d")
    sc = build_scene(cfg)
    keep = sc.mat_type == -1
    for k in ("pos", "vel", "density", "mass", "m_V", "mat_type", "obj_id", "mat_id", "stress",
              "color0"):
        setattr(sc, k, getattr(sc, k)[keep])
    sc.block_counts = [0]
    eng = L.Engine(sc.method, 2, sc.n, sc.geo.grid_num_total, 1)
    slot = eng.add_scene(sc.desc())
    L.upload_scene(eng, slot, sc)
    eng.step(10)
    eng.sync()
    pos = eng.download(slot, "pos")
    order = eng.download(slot, "id0")
    np.testing.assert_allclose(pos, sc.pos[order], rtol=0, atol=1e-7)
    assert eng.status(slot) == (0, 0)
    g = eng.globals(slot)
    assert g[0] == -np.inf and g[3] == 0.0
    eng.close()


def test_single_particle_free_fall():
    """One fluid particle with no neighbour and no walls (boundary 0): the
    velocity is g t (any step of a constant acceleration adds dt g), the
    density never changes, and the position equals the float32 oracle's."""
    from oracle.oracle import Oracle, load_scene
    sc = build_scene(tiny_cfg(boundary=0, block=(0.01, 0.01, 0.01)))
    assert sc.n == 1
    steps, dt = 200, sc.desc()["dt"]
    out = run_alone(sc, steps)
    np.testing.assert_allclose(out["vel"][0, 1], -9.81 * steps * dt, rtol=1e-5)
    assert out["vel"][0, 0] == 0.0
    assert out["density"][0] == np.float32(sc.density[0])
    o = Oracle(sc.desc(), 32)
    load_scene(o, sc)
    o.step(steps)
    np.testing.assert_array_equal(out["pos"], o.get("pos"))
    np.testing.assert_array_equal(out["vel"], o.get("vel"))


def test_capacity_limits_are_errors():
    sc = build_scene(tiny_cfg())
    eng = L.Engine(sc.method, 2, sc.n, sc.geo.grid_num_total, 1)
    slot = eng.add_scene(sc.desc())
    with pytest.raises(L.TsphError):
        eng.add_scene(sc.desc())  # no free slot
    d = sc.desc()
    d["method"] = 2
    with pytest.raises(ValueError):
        eng.add_scene(d)  # another method: an argument error
    d = sc.desc()
    d["time_integration"] = 1
    with pytest.raises(ValueError):
        eng.add_scene(d)  # another integrator than the resident scene's
    L.upload_scene(eng, slot, sc)
    eng.step(3)
    eng.sync()
    assert eng.status(slot) == (0, 0)
    eng.close()
    eng = L.Engine(sc.method, 2, sc.n, sc.geo.grid_num_total, 2)
    eng.add_scene(sc.desc())
    with pytest.raises(L.TsphError):
        eng.add_scene(sc.desc())  # particle capacity exhausted
    eng.close()
    with pytest.raises(ValueError):
        L.Engine(sc.method, 2, (1 << 26) + 1, 100, 1)


def test_out_of_domain_inputs_are_argument_errors():
    """A position outside the cell grid over the virtual domain (the
    reference's pos_to_index would address a cell that does not exist) or
    a cell id or an id0 out of range is refused at upload; the engine stays
    usable."""
    sc = build_scene(tiny_cfg())
    eng = L.Engine(sc.method, 2, sc.n, sc.geo.grid_num_total, 1)
    slot = eng.add_scene(sc.desc())
    L.upload_scene(eng, slot, sc)
    bad = sc.pos.copy()
    bad[5, 0] = sc.geo.vdomain_start[0] - 1e-3
    with pytest.raises(ValueError):
        eng.upload(slot, "pos", bad)
    bad = sc.pos.copy()
    g = sc.geo
    bad[7, 1] = g.vdomain_start[1] + g.grid_num[1] * g.grid_size + 1e-3  # past the last cell
    with pytest.raises(ValueError):
        eng.upload(slot, "pos", bad)
    cell = eng.download(slot, "cell")
    cell[3] = sc.geo.grid_num_total
    with pytest.raises(ValueError):
        eng.upload(slot, "cell", cell)
    # id0 indexes the scene's initial positions and colours on the device:
    # a value outside [0, n) is refused before any kernel reads it
    good = eng.download(slot, "id0")
    for v in (-1, sc.n):
        ids = good.copy()
        ids[11] = v
        with pytest.raises(ValueError):
            eng.upload(slot, "id0", ids)
        np.testing.assert_array_equal(eng.download(slot, "id0"), good)
    L.upload_scene(eng, slot, sc)
    eng.step(3)
    eng.sync()
    assert eng.status(slot) == (0, 0)
    eng.close()


def test_pair_distance_root_equals_sqrtf():
    """The pair kernels take |x_ij| from v_sqrt_f32 with a one-ulp residual
    correction instead of the correctly rounded sqrtf: over every finite
    non-negative float it must give sqrtf's bits at arguments >= 2^-96 and
    a value below 1e-8 (a zero kernel value either way) under it."""
    bad, small = L.check_sqrt()
    assert (bad, small) == (0, 0)
