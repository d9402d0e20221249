"""The reference's five scene files on the GPU: ten fused steps of each
(WC dam break, DP column collapse with XSPH, hydrostatic tank, DP with a
vertical and with horizontal static rigid intruders) against the float32
oracle from the same input: particle order and cell ids bit-exact, every
state field within the parity tolerance."""
import glob
import os

import numpy as np
import pytest

from oracle.oracle import Oracle, load_scene
from tests.common import GOLDEN, TOL, floors, parity_err
from tisphi_amd import _lib as L
from tisphi_amd.eng.configer_builder import SimConfiger
from tisphi_amd.scene import build_scene

pytestmark = pytest.mark.gpu

SCENES = sorted(glob.glob(os.path.join(GOLDEN, "scenes", "*.json")))
STEPS = 10


@pytest.mark.parametrize("path", SCENES, ids=[os.path.basename(p)[:-5] for p in SCENES])
def test_reference_scene_vs_oracle(path):
    sc = build_scene(SimConfiger(path))
    print(f"{os.path.basename(path)}: {sc.n} particles, method {sc.method}", flush=True)
    eng = L.Engine(sc.method, sc.geo.dim, sc.n, sc.geo.grid_num_total, 1)
    slot = eng.add_scene(sc.desc())
    L.upload_scene(eng, slot, sc)
    eng.step(STEPS)
    eng.sync()
    assert eng.status(slot) == (0, 0)
    o = Oracle(sc.desc(), 32)
    load_scene(o, sc)
    o.step(STEPS)
    np.testing.assert_array_equal(eng.download(slot, "id0"), o.get("id0"))
    np.testing.assert_array_equal(eng.download(slot, "cell"), o.get("cell"))
    H = float(sc.cfg.get_blocks()[0]["size"][1])
    fl = floors(sc, H)
    fl.update({"offset": sc.geo.grid_size, "m_V": sc.geo.m_V0})
    fields = ["offset", "vel", "density", "pressure", "d_vel", "d_density", "m_V"]
    if sc.method != 1:
        fields += ["stress", "strain_equ"]
    if sc.method == 3:
        fields += ["strain_equ_p"]
    bad = []
    for f in fields:
        b = o.get(f)
        floor = fl.get(f, max(float(np.max(np.abs(b))) if b.size else 0.0, 1e-30))
        e = parity_err(eng.download(slot, f), b, floor)
        print(f"  {f}: {e:.2e}", flush=True)
        if not e <= TOL:
            bad.append(f"{f}: {e:.2e}")
    if sc.method == 3:
        np.testing.assert_array_equal(eng.download(slot, "flag_retmap"), o.get("flag_retmap"))
    eng.close()
    assert not bad, "parity above tolerance: " + ", ".join(bad)
