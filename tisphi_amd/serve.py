"""Continuous-batching scene server.

Requests (a scene plus a step budget) arrive over time.  The server keeps
one engine per (method, dimension, time integration) class, each sized from
an HBM budget and
driven on its own HIP stream.  Between iterations it admits waiting requests
into free scene slots / particle / cell ranges (in arrival order per class;
a request that does not fit yet is passed over by later ones that do, a
bounded number of times), advances every engine that has residents, and
retires scenes whose budget is spent.  Particle state stays resident on the device from admission
to retirement; the engines' launches overlap on the GPU.

Iteration length: each engine advances by the smallest remaining budget of
its residents (capped by ``chunk``), so every scene stops exactly at its
budget and results do not depend on what else is resident.
"""
import copy
import json
import math
import os
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .eng.configer_builder import SimConfiger
from .scene import build_scene

SCENES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "scenes")


def class_key(scene):
    """Engine class of a scene: an engine steps scenes of one method,
    dimension and time integration."""
    return (scene.method, scene.geo.dim, scene.time_integration)


@dataclass
class Request:
    rid: int
    scene: object          # tisphi_amd.scene.Scene (generated)
    steps: int
    arrival: float         # seconds after the stream starts
    kind: str = ""
    admitted: float = None
    finished: float = None
    result: dict = field(default_factory=dict)

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival


def _resized(base, n_target, dims):
    """Scale the particle radius of ``base`` so its block holds ~n_target
    particles (``dims`` = number of lattice axes)."""
    c = copy.deepcopy(json.load(open(os.path.join(SCENES, base + ".json"))))
    size = c["Blocks"][0]["size"]
    vol = float(np.prod(size[:dims]))
    d = (vol / n_target) ** (1.0 / dims)
    c["Configuration"]["particleRadius"] = d / 2
    return c


def request_mix(n=64, seed=1, rate=None, steps_range=(500, 5000), scale_steps=1.0,
                kinds=("wc2d", "wc3d", "mui2d")):
    """The served stream: kinds in equal shares, fluid/soil particle counts
    log-uniform (2D WC 1k-50k, 3D WC 50k-500k, 2D muI 10k-80k), step budgets
    uniform in ``steps_range`` (times ``scale_steps``), Poisson arrivals at
    ``rate`` requests/s (all at t=0 when rate is None)."""
    rng = np.random.default_rng(seed)
    # arrival times from a generator of their own: the same seed gives the
    # same scenes and budgets at any rate (or none)
    rng_t = np.random.default_rng([seed, 1])
    spec = {"wc2d": ("dam_break_2d", 2, 1e3, 5e4), "wc3d": ("dam_break_3d", 3, 5e4, 5e5),
            "mui2d": ("column_collapse_mui", 2, 1e4, 8e4)}
    t = 0.0
    out = []
    for k in range(n):
        kind = kinds[int(rng.integers(len(kinds)))]
        base, dims, lo, hi = spec[kind]
        nt = math.exp(rng.uniform(math.log(lo), math.log(hi)))
        cfg = SimConfiger(config=_resized(base, nt, dims))
        steps = max(1, int(round(rng.integers(steps_range[0], steps_range[1] + 1) * scale_steps)))
        if rate:
            t += rng_t.exponential(1.0 / rate)
        out.append(Request(k, build_scene(cfg), steps, t, kind))
    return out


class _Class:
    def __init__(self, method, dim, particles, cells, max_scenes):
        self.engine = _lib.Engine(method, dim, particles, cells, max_scenes)
        self.resident = {}  # slot -> (request, steps done)


class Server:
    def __init__(self, capacity, chunk=100, max_scenes=256, max_bypass=8):
        """``capacity``: {(method, dim, time integration): (particles,
        cells)} — the HBM budget split per class (see ``plan_capacity``);
        ``max_bypass``: how often a waiting request may be passed over by
        later ones that fit."""
        self.chunk = chunk
        self.max_bypass = max_bypass
        self._bypassed = {}
        self.classes = {k: _Class(k[0], k[1], p, c, max_scenes) for k, (p, c) in capacity.items()}

    @property
    def hbm_bytes(self):
        return sum(c.engine.bytes for c in self.classes.values())

    def _admit(self, cls, req, now):
        sc = req.scene
        try:
            slot = cls.engine.add_scene(sc.desc())
        except _lib.TsphError as e:
            if e.code == _lib.ERR_CAPACITY:
                return False
            raise
        try:
            _lib.upload_scene(cls.engine, slot, sc)
        except Exception:
            cls.engine.remove_scene(slot)  # the slot and its ranges go back
            raise
        cls.resident[slot] = [req, 0]
        req.admitted = now
        return True

    def _admit_waiting(self, cls, q, now):
        """Admit waiting requests of a class in arrival order; one that does
        not fit is passed over for later ones that do (backfill), at most
        ``max_bypass`` times, after which the class admits nothing until it
        fits (so a large request is not starved by a stream of small ones)."""
        k = 0
        while k < len(q):
            r = q[k]
            if self._admit(cls, r, now):
                q.pop(k)
                continue
            if self._bypassed.get(r.rid, 0) >= self.max_bypass:
                break  # the oldest blocked request holds the class
            self._bypassed[r.rid] = self._bypassed.get(r.rid, 0) + 1
            k += 1

    def run(self, requests, realtime=True, log=None):
        """Serve ``requests`` (sorted by arrival).  Returns them with
        admission / completion times and final globals filled in; ``log``
        (optional) is called with a line for every retired request."""
        pending = sorted(requests, key=lambda r: r.arrival)
        waiting = {k: [] for k in self.classes}
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0  # noqa: E731
        done = []
        total_steps = 0
        while pending or any(waiting.values()) or any(c.resident for c in self.classes.values()):
            tn = now() if realtime else float("inf")
            while pending and pending[0].arrival <= tn:
                r = pending.pop(0)
                waiting[class_key(r.scene)].append(r)
            for k, cls in self.classes.items():
                self._admit_waiting(cls, waiting[k], now())
            active = [c for c in self.classes.values() if c.resident]
            if not active:
                if pending:
                    time.sleep(max(0.0, min(pending[0].arrival - now(), 0.05)))
                    continue
                if any(waiting.values()):
                    raise RuntimeError("a waiting request does not fit any engine")
                break
            adv = {}
            for c in active:
                k = min(min(r.steps - s for r, s in c.resident.values()), self.chunk)
                c.engine.step(k)  # asynchronous: engines overlap on the device
                adv[id(c)] = k
            for c in active:
                c.engine.sync()
            for c in active:
                k = adv[id(c)]
                for slot in list(c.resident):
                    r, s = c.resident[slot]
                    s += k
                    total_steps += k
                    c.resident[slot][1] = s
                    if s >= r.steps:
                        r.result = {"globals": c.engine.globals(slot), "steps": c.engine.steps(slot)}
                        r.finished = now()
                        c.engine.remove_scene(slot)
                        del c.resident[slot]
                        done.append(r)
                        if log:
                            log(f"request {r.rid} ({r.kind}, {r.scene.n} particles, {r.steps} steps) "
                                f"done at {r.finished:.2f} s")
        self.scene_steps = total_steps
        self.wall = now()
        return done

    def close(self):
        for c in self.classes.values():
            c.engine.close()


def plan_capacity(requests, hbm_budget_bytes=None, headroom=1.0):
    """Per-class capacity: room for the largest ``headroom`` x concurrent
    load of the class (by default every request of the class at once),
    bounded by an HBM budget split in proportion to particle counts."""
    need, big = {}, {}
    for r in requests:
        k = class_key(r.scene)
        p, c = need.get(k, (0, 0))
        need[k] = (p + r.scene.n, c + r.scene.geo.grid_num_total)
        bp, bc = big.get(k, (0, 0))
        big[k] = (max(bp, r.scene.n), max(bc, r.scene.geo.grid_num_total))
    # every class can hold at least its largest scene
    cap = {k: (max(int(p * headroom), big[k][0]) + 64, max(int(c * headroom), big[k][1]) + 1)
           for k, (p, c) in need.items()}
    if hbm_budget_bytes:
        per_particle = 1400  # upper bound of engine bytes per particle (3D DP)
        tot = sum(p for p, _ in cap.values())
        scale = min(1.0, hbm_budget_bytes / (per_particle * max(tot, 1)))
        cap = {k: (max(int(p * scale), big[k][0] + 64), c) for k, (p, c) in cap.items()}
    return cap


def percentile(v, q):
    return float(np.percentile(np.asarray(v, dtype=np.float64), q)) if len(v) else float("nan")


def run_alone(scene, steps, graph=False):
    """``scene`` alone in its own engine for ``steps`` steps, one step
    launched and synchronised at a time (the one-scene-per-launch path).
    Returns (seconds of stepping, FP64 globals after the steps)."""
    e = _lib.Engine(scene.method, scene.geo.dim, scene.n, scene.geo.grid_num_total, 1)
    try:
        s = e.add_scene(scene.desc())
        _lib.upload_scene(e, s, scene)
        e.set_graph(graph)
        e.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            e.step(1)
            e.sync()
        return time.perf_counter() - t0, e.globals(s)
    finally:
        e.close()


def plain_run(requests, log=None):
    """The plain path on the same requests, run: each request alone, its
    whole step budget one step at a time, one request after the other.
    Returns {kind: (scene-steps, seconds)} and the total."""
    per = {}
    for r in requests:
        t, _ = run_alone(r.scene, r.steps)
        st, sec = per.get(r.kind, (0, 0.0))
        per[r.kind] = (st + r.steps, sec + t)
        if log:
            log(f"plain: request {r.rid} ({r.kind}, {r.scene.n} particles, {r.steps} steps) {t:.2f} s")
    tot = (sum(v[0] for v in per.values()), sum(v[1] for v in per.values()))
    return per, tot


def serve_bench(n=24, seed=1, load=0.7, chunk=100, log=None, fp64_steps=None):
    """The served-stream measurement on one request mix (``n`` requests,
    ``seed``):

    * capacity: the whole mix admitted at once (saturated), scene-steps/s;
    * a Poisson stream of the same requests at ``load`` x capacity:
      throughput, p50 / p99 completion latency;
    * per kind (2D WC, 3D WC, 2D mu(I)): that kind's requests served at
      once against the same requests run plain, one after the other, each
      for its whole budget (measured, not extrapolated);
    * for the smallest request of each kind, the globals of its GPU run
      alone after its budget (``fp64_steps`` caps the steps per kind), for
      the caller to check against an FP64 run; where the whole budget ran,
      whether the served result equals it.

    Returns (results, samples) with samples = [(kind, scene, steps,
    globals)]."""
    say = log or (lambda *a: None)
    reqs = request_mix(n=n, seed=seed, rate=None)
    cap = plan_capacity(reqs)
    srv = Server(cap, chunk=chunk, max_scenes=n)
    done = srv.run(reqs, realtime=False, log=say)
    t_all = srv.wall
    steps_all = sum(r.steps for r in done)
    capacity = n / t_all
    say(f"serve: all {n} requests at once in {t_all:.2f} s ({steps_all / t_all:.0f} scene-steps/s)")
    hbm = srv.hbm_bytes
    srv.close()
    rate = load * capacity
    reqs2 = request_mix(n=n, seed=seed, rate=rate)
    srv2 = Server(cap, chunk=chunk, max_scenes=n)
    done2 = srv2.run(reqs2, realtime=True, log=say)
    lat = [r.latency for r in done2]
    served_rate = sum(r.steps for r in done2) / srv2.wall
    say(f"serve: Poisson stream at {rate:.3f} req/s done in {srv2.wall:.2f} s")
    srv2.close()
    kinds = sorted({r.kind for r in reqs})
    per_kind = {}
    for k in kinds:
        rk = [r for r in request_mix(n=n, seed=seed, rate=None) if r.kind == k]
        sk = Server(plan_capacity(rk), chunk=chunk, max_scenes=n)
        dk = sk.run(rk, realtime=False)
        per_kind[k] = dict(requests=len(rk), scene_steps=sum(r.steps for r in dk),
                           served_scene_steps_per_s=sum(r.steps for r in dk) / sk.wall)
        sk.close()
        say(f"serve: {k} alone: {len(rk)} requests in {sk.wall:.2f} s")
    plain_k, (plain_steps, plain_s) = plain_run(reqs, log=say)
    for k, (st, sec) in plain_k.items():
        per_kind[k]["plain_scene_steps_per_s"] = st / sec
        per_kind[k]["speedup_vs_plain"] = per_kind[k]["served_scene_steps_per_s"] / (st / sec)
    plain = plain_steps / plain_s
    say(f"serve: plain path {plain:.1f} scene-steps/s ({plain_s:.1f} s)")
    caps = fp64_steps or {"wc2d": None, "mui2d": 500, "wc3d": 20}
    sample, samples = {}, []
    for k in kinds:
        r = min((q for q in reqs if q.kind == k), key=lambda q: q.scene.n)
        st = r.steps if caps.get(k) is None else min(r.steps, caps[k])
        _, g = run_alone(r.scene, st)
        served = next(q for q in done2 if q.rid == r.rid)
        entry = dict(request=r.rid, particles=r.scene.n, steps=st, of_budget=r.steps)
        if st == r.steps:  # the served result itself equals the alone run
            entry["served_equals_alone"] = bool(np.array_equal(served.result["globals"], g))
        sample[k] = entry
        samples.append((k, r.scene, st, g))
    out = dict(requests=n, arrival_rate_per_s=rate, load=load,
               capacity_requests_per_s=capacity,
               saturated_scene_steps_per_s=steps_all / t_all,
               served_scene_steps_per_s=served_rate,
               p50_latency_s=percentile(lat, 50), p99_latency_s=percentile(lat, 99),
               plain_scene_steps_per_s=plain, plain_measured_s=plain_s,
               speedup_vs_plain=(steps_all / t_all) / plain,
               per_kind=per_kind, fp64_sample=sample, hbm_bytes=hbm)
    return out, samples
