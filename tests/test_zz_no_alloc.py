"""tsph_neighbor_stats_noalloc (scratch in a static device array, no
allocation) against tsph_neighbor_stats on the same state: two calls in a
row, the second on a smaller scene, and the size limit refused."""
import numpy as np
import pytest

from tests.common import tiny_cfg
from tisphi_amd import _lib as L
from tisphi_amd.scene import build_scene

pytestmark = pytest.mark.gpu


def _engine(cfg, steps):
    sc = build_scene(cfg)
    eng = L.Engine(sc.method, sc.geo.dim, sc.n, sc.geo.grid_num_total, 2)
    slot = eng.add_scene(sc.desc())
    L.upload_scene(eng, slot, sc)
    eng.step(steps)
    eng.sync()
    return sc, eng, slot


def test_noalloc_neighbor_stats_equal_allocating_ones():
    big = tiny_cfg(method=1, dim=3, r=0.005, domain=(0.3, 0.3, 0.2), block=(0.15, 0.15, 0.12))
    small = tiny_cfg(method=1, dim=2)
    sc1, e1, s1 = _engine(big, 5)
    sc2, e2, s2 = _engine(small, 3)
    assert sc2.n < sc1.n
    print(f"no-alloc neighbour stats: {sc1.n} then {sc2.n} particles", flush=True)
    for eng, slot in ((e1, s1), (e2, s2)):
        want = eng.neighbor_stats(slot)
        got = eng.neighbor_stats_noalloc(slot)
        for a, b in zip(got, want):
            np.testing.assert_array_equal(a, b)
        assert want[0].sum() > 0
    # the first engine again, after the smaller call: still equal
    for a, b in zip(e1.neighbor_stats_noalloc(s1), e1.neighbor_stats(s1)):
        np.testing.assert_array_equal(a, b)
    e1.close()
    e2.close()


def test_noalloc_neighbor_stats_refuses_large_slots():
    sc = build_scene(tiny_cfg(method=1, dim=3, r=0.002, domain=(0.4, 0.4, 0.2), block=(0.3, 0.3, 0.15)))
    assert sc.n > 262144
    eng = L.Engine(sc.method, 3, sc.n, sc.geo.grid_num_total, 1)
    slot = eng.add_scene(sc.desc())
    L.upload_scene(eng, slot, sc)
    with pytest.raises(ValueError):
        eng.neighbor_stats_noalloc(slot)
    eng.step(1)  # the engine stays usable
    eng.sync()
    eng.close()
