"""The colour restatement (oracle/colour.py) against values worked by hand,
and the facade plumbing of the colour feed (no GPU)."""
import numpy as np
import pytest

from oracle import colour as C


def test_jet_closed_form():
    """colormap.py:35-48 by hand: channel = clamp((0.37 - |clamp(x) - c|) /
    0.37 * 1.5), c = 0.75 / 0.5 / 0.25; k = 1.5 * 0.12 / 0.37."""
    k = 1.5 * (0.37 - 0.25) / 0.37
    want = {
        0.0: (0.0, 0.0, k),
        0.25: (0.0, k, 1.0),
        0.5: (k, 1.0, k),
        0.75: (1.0, k, 0.0),
        1.0: (k, 0.0, 0.0),
        -3.0: (0.0, 0.0, k),   # clamped to 0
        7.0: (k, 0.0, 0.0),    # clamped to 1
    }
    x = np.array(list(want))
    got = C.jet(x)
    assert got.dtype == np.float32
    np.testing.assert_array_equal(got, np.array(list(want.values()), dtype=np.float32))
    # the band edges: a channel rises from 0 at c - 0.37 to 1 at c - 0.37/3
    c = 0.5
    np.testing.assert_allclose(C.jet(np.array([c - 0.37, c - 0.37 / 3.0]))[:, 1], [0.0, 1.0],
                               atol=1e-7)


def test_v_maxmin_rule():
    """particle_system.py:395-396 over flow particles only."""
    val = np.array([5.0, -2.0, 100.0, 1.0])
    typ = np.array([1, 2, -1, 1])  # the dummy's 100 is not flow
    assert C.v_maxmin(val, typ) == (5.0, -2.0)
    assert C.v_maxmin(val, typ, givenmax=3.0) == (3.0, -2.0)   # vmax >= given: given
    assert C.v_maxmin(val, typ, givenmax=8.0) == (5.0, -2.0)   # below given, not fixed
    assert C.v_maxmin(val, typ, givenmax=8.0, fixmax=1) == (8.0, -2.0)
    assert C.v_maxmin(val, typ, givenmin=-1.5) == (5.0, -1.5)  # vmin <= given: given
    assert C.v_maxmin(val, typ, givenmin=-9.0) == (5.0, -2.0)
    assert C.v_maxmin(val, typ, givenmin=-9.0, fixmin=1) == (5.0, -9.0)
    assert C.v_maxmin(val, np.array([-1, -1, -1, -1])) == (-np.inf, np.inf)


def test_value_titles_and_groups():
    """Title table of solver_sph_base.py:724-789 on a hand-made state; dummy
    values only with show_bdy; colour groups of particle_system.py:403-405."""
    n = 4
    typ = np.array([1, 2, -1, -2])
    f = dict(type=typ, id0=np.array([3, 2, 1, 0]), vel=np.array([[3.0, 4.0, 0.0]] * n),
             d_vel=np.array([[1.0, -2.0, 2.0]] * n), pos=np.array([[1.0, 2.0, 2.0]] * n),
             x0_by_id=np.zeros((n, 3)), density=np.full(n, 1000.0), d_density=np.ones(n),
             pressure=np.full(n, 7.0), cell=np.arange(n), v_tmp=np.zeros((n, 3)),
             stress=np.array([[-1.0, -2.0, -3.0, 0.5, 0.6, 0.7]] * n))
    np.testing.assert_array_equal(C.assign_value_color(f, 3, False), [5.0, 5.0, 0.0, 0.0])
    np.testing.assert_array_equal(C.assign_value_color(f, 3, True), [5.0, 5.0, 5.0, 0.0])
    np.testing.assert_array_equal(C.assign_value_color(f, 32, False), [-4.0, -4.0, 0, 0])
    np.testing.assert_array_equal(C.assign_value_color(f, 34, False), [3.0, 3.0, 0, 0])
    np.testing.assert_array_equal(C.assign_value_color(f, 44, False), [3.0, 3.0, 0, 0])
    np.testing.assert_array_equal(C.assign_value_color(f, 52, False), [2.0, 2.0, 0, 0])
    np.testing.assert_array_equal(C.assign_value_color(f, 57, False), [2.0, 2.0, 0, 0])
    np.testing.assert_array_equal(C.assign_value_color(f, 1, True), [3.0, 2.0, 1.0, 0.0])
    np.testing.assert_array_equal(C.assign_value_color(f, 999, True), np.zeros(n))
    c0 = np.array([[0.1, 0.2, 0.3], [0.4, 0.5, 0.6], [0.7, 0.8, 0.9], [1.0, 0.0, 0.5]])
    val = np.array([0.0, 1.0, 0.5, 0.5])
    col = C.set_color(val, typ, f["id0"], c0, 1.0, 0.0, 0, False)
    np.testing.assert_array_equal(col[0], C.jet(np.array([0.0]))[0])
    np.testing.assert_array_equal(col[1], C.jet(np.array([1.0]))[0])
    np.testing.assert_array_equal(col[2], c0[1].astype(np.float32))  # dummy: initial colour
    np.testing.assert_array_equal(col[3], c0[0].astype(np.float32))
    col = C.set_color(val, typ, f["id0"], c0, 1.0, 0.0, 2, True)
    np.testing.assert_array_equal(col[2], C.jet(np.array([0.5]))[0])  # dummy shown
    np.testing.assert_array_equal(col[3], c0[0].astype(np.float32))   # repulsive never


def test_scene_initial_colours():
    """Material colour / 255 as float32 for blocks (particle_system.py:159),
    [153, 153, 255] / 255 for dummy and [170, 17, 255] / 255 for repulsive
    walls (:146-148)."""
    from tests.common import tiny_cfg
    from tisphi_amd.scene import build_scene
    sc = build_scene(tiny_cfg(boundary=4))
    assert sc.color0.shape == (sc.n, 3) and sc.color0.dtype == np.float32
    fl = sc.mat_type == 1
    mat = sc.mat_fluid[0]
    np.testing.assert_array_equal(sc.color0[fl][0], np.array([c / 255 for c in mat["color"]], np.float32))
    np.testing.assert_array_equal(sc.color0[sc.mat_type == -1][0],
                                  (np.array([153, 153, 255]) / 255).astype(np.float32))
    np.testing.assert_array_equal(sc.color0[sc.mat_type == -2][0],
                                  (np.array([170, 17, 255]) / 255).astype(np.float32))


def test_colour_entry_points_exported():
    pytest.importorskip("ctypes")
    from tisphi_amd import _lib as L
    lib = L.load()
    for name in ("tsph_assign_value_color", "tsph_v_maxmin", "tsph_set_color"):
        assert hasattr(lib, name)
    assert L.FIELDS["val"][0] == 26 and L.FIELDS["color0"][0] == 29
