"""The device colour feed (tsph_assign_value_color / tsph_v_maxmin /
tsph_set_color) against the CPU restatement in oracle/colour.py, computed
from the same engine state: values, range and colours equal."""
import numpy as np
import pytest

from oracle import colour as C
from tests.common import scene_cfg
from tisphi_amd import _lib as L
from tisphi_amd.scene import build_scene

pytestmark = pytest.mark.gpu

FIELDS = ["type", "id0", "vel", "d_vel", "pos", "density", "d_density", "pressure", "cell",
          "v_tmp", "cspm_L"]


def stepped(name, mods=None, steps=60):
    cfg = scene_cfg(name)
    if mods:
        cfg.config["Configuration"].update(mods)
    sc = build_scene(cfg)
    eng = L.Engine(sc.method, sc.geo.dim, sc.n, sc.geo.grid_num_total, 1)
    slot = eng.add_scene(sc.desc())
    L.upload_scene(eng, slot, sc)
    eng.upload(slot, "x0", sc.pos)
    eng.upload(slot, "color0", sc.color0.astype(np.float64))
    eng.step(steps)
    eng.sync()
    f = {k: eng.download(slot, k) for k in FIELDS}
    if sc.method != 1:
        for k in ("stress", "strain_equ", "strain_equ_p", "flag_retmap"):
            f[k] = eng.download(slot, k)
    f["x0_by_id"] = sc.pos
    f["kernel_corr"] = sc.kernel_corr
    return sc, eng, slot, f


def check_frame(sc, eng, slot, f, title, group, show_bdy, given=(-1.0, -1.0, 0, 0)):
    eng.assign_value_color(slot, title, show_bdy)
    val = eng.download(slot, "val")
    want = C.assign_value_color(f, title, show_bdy)
    if title in (3, 34, 4, 44, 103):  # FP64 square roots
        np.testing.assert_allclose(val, want, rtol=1e-15, atol=0, err_msg=f"title {title}")
        val_ref = val
    else:
        np.testing.assert_array_equal(val, want, err_msg=f"title {title}")
        val_ref = want
    vmax, vmin = eng.v_maxmin(slot, *given)
    assert (vmax, vmin) == C.v_maxmin(val_ref, f["type"], *given), f"title {title}"
    eng.set_color(slot, group, show_bdy)
    col = eng.download(slot, "color").astype(np.float32)
    want_c = C.set_color(val_ref, f["type"], f["id0"], sc.color0, vmax, vmin, group, show_bdy)
    np.testing.assert_array_equal(col, want_c, err_msg=f"title {title} group {group}")


def test_colour_feed_wc():
    sc, eng, slot, f = stepped("dam_break_2d", {"kernelCorrection": 1})
    for title in (1, 2, 21, 3, 31, 32, 33, 34, 35, 36, 37, 4, 41, 42, 43, 44, 7, 100, 102, 103,
                  51, 61, 8, 999):
        check_frame(sc, eng, slot, f, title, group=0, show_bdy=False)
    for group in (1, 2):
        for show in (False, True):
            check_frame(sc, eng, slot, f, 41, group, show)
    # the given / fixed range rule
    check_frame(sc, eng, slot, f, 41, 0, False, given=(1.0, 0.2, 0, 0))
    check_frame(sc, eng, slot, f, 41, 0, False, given=(5.0, -3.0, 1, 1))
    eng.close()


def test_colour_feed_dp():
    sc, eng, slot, f = stepped("column_collapse_dp", steps=40)
    for title in (51, 52, 53, 54, 55, 56, 57, 61, 62, 8, 7, 3, 21):
        check_frame(sc, eng, slot, f, title, group=0, show_bdy=False)
    check_frame(sc, eng, slot, f, 57, 2, True)
    eng.close()


def test_colour_random_title():
    """Title 101 (ti.random, solver_sph_base.py:782-783): a uniform in
    [0, 1) per particle, fixed by (seed, id0), other seeds other values."""
    sc, eng, slot, f = stepped("dam_break_2d", steps=5)
    eng.assign_value_color(slot, 101, False, seed=1)
    a = eng.download(slot, "val")
    eng.assign_value_color(slot, 101, False, seed=1)
    b = eng.download(slot, "val")
    eng.assign_value_color(slot, 101, False, seed=2)
    c = eng.download(slot, "val")
    real = f["type"] > 0
    np.testing.assert_array_equal(a, b)
    assert np.all((a[real] >= 0.0) & (a[real] < 1.0)) and not np.any(a[~real])
    assert np.mean(a[real] != c[real]) > 0.99
    assert abs(np.mean(a[real]) - 0.5) < 0.02 and abs(np.var(a[real]) - 1 / 12) < 0.01
    eng.close()


def test_headless_run_colours_frames(tmp_path):
    """The headless loop colours the frames as the window loop does
    (eng/ui_sim.py:94, 107, 282-286)."""
    import json
    from tisphi_amd import run_simulation
    cfg = scene_cfg("dam_break_2d").config
    cfg["Configuration"].update({"exitAtStep": 20, "stepsPerRenderUpdate": 10, "colorTitle": 3})
    p = tmp_path / "s.json"
    p.write_text(json.dumps(cfg))
    case, steps, _ = run_simulation.run(str(p), log=lambda *a: None)
    assert steps == 20
    col = case.ps.pt.color.to_numpy()
    val = case.ps.pt.val.to_numpy()
    t = case.ps.pt.mat_type.to_numpy()
    fl = t == 1
    assert np.all(val[fl] >= 0) and case.ps.vmax[None] == np.max(val[fl])
    assert np.all(col[fl] >= 0) and np.all(col[fl] <= 1) and np.any(col[fl] != col[fl][0])
