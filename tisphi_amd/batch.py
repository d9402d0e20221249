"""Batched scenes: many independent scenes advanced by one launch sequence.

The reference runs one scene per process (run_simulation.py:23).  Here an
engine holds up to thousands of resident scenes of one method and
dimension; every kernel of a step covers all of them (one 64-particle tile
per wave, tiles of every scene in one grid), so a step of 256 scenes costs
the same ~30 launches as a step of one.  Per-scene arithmetic does not
depend on the other residents: each scene keeps its own particle and cell
ranges, its own sort order and its own tile-aligned reductions, so results
are bitwise identical alone and in a batch.
"""
import copy
import json
import os
import time

import numpy as np

from . import _lib
from .eng.configer_builder import SimConfiger
from .scene import build_scene

SCENES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "scenes")


def dam_break_family(n=256, seed=0, base="dam_break_2d"):
    """Scene configurations of the batched 2D dam breaks: column width
    U[0.3, 1.2] m and height U[0.2, 0.6] m drawn from ``seed``; everything
    else as the 2D dam break (0.01 m lattice, same WCSPH material)."""
    rng = np.random.default_rng(seed)
    cfg0 = json.load(open(os.path.join(SCENES, base + ".json")))
    out = []
    for k in range(n):
        w, h = rng.uniform(0.3, 1.2), rng.uniform(0.2, 0.6)
        c = copy.deepcopy(cfg0)
        c["Blocks"][0]["size"] = [float(w), float(h), 0.5]
        out.append(SimConfiger(config=c))
    return out


class BatchEngine:
    """Scenes of one method, dimension and time integration resident in one
    engine."""

    def __init__(self, scenes, spare_particles=0, spare_cells=0, max_scenes=None):
        if not scenes:
            raise ValueError("no scenes")
        m, d, ti = scenes[0].method, scenes[0].geo.dim, scenes[0].time_integration
        if any(s.method != m or s.geo.dim != d or s.time_integration != ti for s in scenes):
            raise ValueError("a batch holds scenes of one method, dimension and time integration")
        npart = sum(s.n for s in scenes) + spare_particles
        ncell = sum(s.geo.grid_num_total for s in scenes) + spare_cells
        self.engine = _lib.Engine(m, d, npart, ncell, max_scenes or len(scenes))
        self.scenes = list(scenes)
        self.slots = []
        for s in scenes:
            slot = self.engine.add_scene(s.desc())
            _lib.upload_scene(self.engine, slot, s)
            self.slots.append(slot)

    def step(self, n=1):
        self.engine.step(n)

    def sync(self):
        self.engine.sync()

    def globals(self):
        return np.array([self.engine.globals(s) for s in self.slots])

    def close(self):
        self.engine.close()


def run_batched(scenes, steps, graph=True):
    """Advance all scenes together; returns (seconds, engine)."""
    b = BatchEngine(scenes)
    b.engine.set_graph(graph)
    b.step(1)
    b.sync()
    t0 = time.perf_counter()
    b.step(steps)
    b.sync()
    return time.perf_counter() - t0, b


def run_plain(scenes, steps, graph=False):
    """One engine per scene, each scene's step launched on its own (the
    one-scene-per-launch path).  Returns (seconds, engines)."""
    engs = []
    for s in scenes:
        e = _lib.Engine(s.method, s.geo.dim, s.n, s.geo.grid_num_total, 1)
        slot = e.add_scene(s.desc())
        _lib.upload_scene(e, slot, s)
        e.set_graph(graph)
        e.step(1)
        e.sync()
        engs.append(e)
    t0 = time.perf_counter()
    for _ in range(steps):
        for e in engs:
            e.step(1)
            e.sync()
    dt = time.perf_counter() - t0
    return dt, engs


def build_family(cfgs):
    return [build_scene(c) for c in cfgs]
