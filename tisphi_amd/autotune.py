"""Launch-shape autotuner.

Picks the launch shape of an engine's resident scenes by timing real
steps: the threads per block of the pair kernels and of the per-particle
kernels, how the per-step neighbour lists are built (stencil walk over
global memory, or over cell windows staged in LDS), and the cell tile of
the LDS build (how many candidate particles a wave stages per row offset:
128, 256 or 512).  Every shape computes the same per-particle arithmetic in
the same order (a wave always owns one 64-particle tile; the shape only
changes how waves are grouped into workgroups and where candidates are
read from), so the results are bitwise identical whatever shape is chosen:
the tuning steps are ordinary simulation steps, and the caller counts them
as such.

Choices are cached per (device, method, dimension, particle-count bucket,
scene-count bucket) in a JSON file (``$TSPH_TUNE_CACHE`` or
``~/.cache/tisphi_amd/launch_shapes.json``); a cached choice is applied
without stepping.
"""
import json
import math
import os
import time

BLOCKS = (64, 128, 256)
CELL_TILES = (128, 256, 512)


def _cache_path():
    p = os.environ.get("TSPH_TUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "tisphi_amd", "launch_shapes.json")


def _load_cache(path):
    try:
        with open(path) as f:
            d = json.load(f)
        return d if isinstance(d, dict) else {}
    except (OSError, ValueError):
        return {}


def _store_cache(path, d):
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except (OSError, ValueError):
        pass  # an unwritable cache location only loses the cache


def tune_key(eng):
    """Cache key: method, dimension, log2 bucket of resident particles and
    of resident scenes."""
    n = max(1, sum(eng.n.values()))
    s = max(1, len(eng.n))
    return f"m{eng.method}/d{eng.dim}/p{int(math.log2(n))}/s{int(math.log2(s))}"


NL_MODES = (1, 2)  # lists built from global memory / from LDS windows


def candidates(quick=False):
    """Shapes to try: (pair_block, elem_block, list mode, cell tile).
    ``quick`` varies one at a time around (256, 256, 2, 256); the cell tile
    only matters to list mode 2."""
    if quick:
        return [(256, 256, 2, 256), (128, 256, 2, 256), (64, 256, 2, 256), (256, 128, 2, 256),
                (256, 64, 2, 256), (256, 256, 1, 256), (256, 256, 2, 128), (256, 256, 2, 512)]
    return ([(p, e, 1, 256) for p in BLOCKS for e in BLOCKS] +
            [(p, e, 2, c) for c in CELL_TILES for p in BLOCKS for e in BLOCKS])


def apply(eng, shape):
    pair, elem, nl, tile = shape
    eng.set_block(pair, elem)
    eng.set_neighbor_list(nl)
    eng.set_cell_tile(tile)


def _time_steps(eng, steps):
    eng.sync()
    t0 = time.perf_counter()
    eng.step(steps)
    eng.sync()
    return (time.perf_counter() - t0) / steps


def tune(eng, steps=2, quick=True, use_cache=True, log=None):
    """Choose and apply the fastest launch shape for ``eng``'s residents.

    Returns ``dict(shape=(pair, elem, nl, cell_tile), ms_per_step=..., steps_taken=k,
    cached=bool, trials={shape: ms})``.  ``steps_taken`` steps were advanced
    (0 when the choice came from the cache)."""
    say = log or (lambda *a: None)
    path = _cache_path()
    key = tune_key(eng)
    cache = _load_cache(path) if use_cache else {}
    if key in cache and len(cache[key].get("shape", ())) == 4:
        shape = tuple(int(v) for v in cache[key]["shape"])
        apply(eng, shape)
        say(f"autotune {key}: cached shape {shape}")
        return dict(shape=shape, ms_per_step=cache[key].get("ms_per_step"), steps_taken=0,
                    cached=True, trials={})
    trials = {}
    taken = 0
    for shape in candidates(quick):
        apply(eng, shape)
        eng.step(1)  # first step with a new shape also captures its graphs
        taken += 1
        ms = _time_steps(eng, steps) * 1e3
        taken += steps
        trials[shape] = ms
        say(f"autotune {key}: shape {shape} {ms:.3f} ms/step")
    best = min(trials, key=trials.get)
    apply(eng, best)
    if use_cache:
        cache = _load_cache(path)
        cache[key] = {"shape": list(best), "ms_per_step": trials[best]}
        _store_cache(path, cache)
    return dict(shape=best, ms_per_step=trials[best], steps_taken=taken, cached=False,
                trials={f"{p}x{e}/nl{m}/tile{c}": v for (p, e, m, c), v in trials.items()})
