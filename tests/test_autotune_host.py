"""Autotuner host logic (CPU): candidate sets, cache keys, cache use."""
import json

from tisphi_amd import autotune


class FakeEngine:
    def __init__(self, n=(7200, 900), method=1, dim=2):
        self.n = {i: v for i, v in enumerate(n)}
        self.method, self.dim = method, dim
        self.shape = (256, 256)
        self.nl = 1
        self.tile = 256
        self.steps_done = 0
        self.log = []

    def set_block(self, p, e):
        self.shape = (p, e)
        self.log.append((p, e))

    def set_neighbor_list(self, m):
        self.nl = m

    def set_cell_tile(self, c):
        self.tile = c

    def step(self, k=1):
        self.steps_done += k
        # the shape (64, 256) with LDS-built lists and 512-candidate cell
        # tiles is the fast one
        import time
        time.sleep(0.0005 if (self.shape, self.nl, self.tile) == ((64, 256), 2, 512) else 0.004)

    def sync(self):
        pass


def test_candidates():
    full = autotune.candidates(quick=False)
    assert len(full) == 36 and len(set(full)) == 36
    assert {c for p, e, m, c in full if m == 2} == {128, 256, 512}
    assert {c for p, e, m, c in full if m == 1} == {256}  # the tile is unused there
    q = autotune.candidates(quick=True)
    assert q[0] == (256, 256, 2, 256) and len(q) == 8 and len(set(q)) == 8
    assert (256, 256, 1, 256) in q and (256, 256, 2, 128) in q and (256, 256, 2, 512) in q


def test_key_buckets():
    assert autotune.tune_key(FakeEngine((7200,))) == "m1/d2/p12/s0"
    assert autotune.tune_key(FakeEngine(tuple([1000] * 256), method=2)) == "m2/d2/p17/s8"


def test_tune_picks_fastest_and_caches(tmp_path, monkeypatch):
    path = tmp_path / "c.json"
    monkeypatch.setenv("TSPH_TUNE_CACHE", str(path))
    e = FakeEngine()
    r = autotune.tune(e, steps=2, quick=False)
    assert r["shape"] == (64, 256, 2, 512) and e.shape == (64, 256) and e.nl == 2 and e.tile == 512
    assert r["steps_taken"] == e.steps_done == 36 * 3
    assert "64x256/nl2/tile512" in r["trials"]
    d = json.load(open(path))
    assert d[autotune.tune_key(e)]["shape"] == [64, 256, 2, 512]
    e2 = FakeEngine()
    r2 = autotune.tune(e2)
    assert r2["cached"] and r2["steps_taken"] == 0 and e2.steps_done == 0
    assert e2.shape == (64, 256) and e2.nl == 2 and e2.tile == 512


def test_cached_shape_without_cell_tile_is_retuned(tmp_path, monkeypatch):
    path = tmp_path / "c.json"
    monkeypatch.setenv("TSPH_TUNE_CACHE", str(path))
    e = FakeEngine()
    path.write_text(json.dumps({autotune.tune_key(e): {"shape": [64, 256, 2], "ms_per_step": 1.0}}))
    r = autotune.tune(e, steps=1, quick=True)
    assert not r["cached"] and r["steps_taken"] == 8 * 2
    assert len(json.load(open(path))[autotune.tune_key(e)]["shape"]) == 4


def test_unwritable_cache_is_not_fatal(tmp_path, monkeypatch):
    blocker = tmp_path / "file"
    blocker.write_text("not a directory")
    monkeypatch.setenv("TSPH_TUNE_CACHE", str(blocker / "sub" / "c.json"))
    e = FakeEngine()
    r = autotune.tune(e, steps=1, quick=False)
    assert r["shape"] == (64, 256, 2, 512)
