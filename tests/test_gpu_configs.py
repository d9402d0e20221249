"""The batched and served workloads at their stated shape on the GPU.

* 256 batched 2D dam breaks (column width U[0.3, 1.2] m, height
  U[0.2, 0.6] m, seed 0) in one engine: every scene advances; for 8 evenly
  spaced scenes the state equals the scene run alone, bitwise, and the FP64
  globals are within 1 % of the FP64 oracle (printed);
* the served stream (24 requests of 2D WC 1k-50k, 3D WC 50k-500k and 2D
  mu(I) 10k-80k flow particles, seed 1; 3D requests up to ~1 M particles
  with walls) arriving as a Poisson stream in real time, admitted by an HBM
  budget while others run: every request's result equals its run alone,
  bitwise.  Step budgets are cut to 2 % (10-100 steps) for test time; the
  scene count and sizes are the stated ones.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle, load_scene
from tisphi_amd import _lib as L
from tisphi_amd.batch import BatchEngine, build_family, dam_break_family
from tisphi_amd.serve import Server, plan_capacity, request_mix

pytestmark = pytest.mark.gpu

FIELDS = ["pos", "vel", "density", "pressure", "d_vel", "d_density", "id0"]


def alone(sc, steps):
    e = L.Engine(sc.method, sc.geo.dim, sc.n, sc.geo.grid_num_total, 1)
    s = e.add_scene(sc.desc())
    L.upload_scene(e, s, sc)
    e.step(steps)
    e.sync()
    return e, s


def test_256_batched_dam_breaks():
    scenes = build_family(dam_break_family(256, seed=0))
    assert len(scenes) == 256 and min(s.n_flow for s in scenes) >= 600
    steps = 100
    b = BatchEngine(scenes)
    g0 = b.globals()
    b.step(steps)
    b.sync()
    g = b.globals()
    for k, sl in enumerate(b.slots):
        assert b.engine.status(sl) == (0, 0)
        assert b.engine.steps(sl) == steps
    assert np.all(g[:, 3] == g0[:, 3])  # mass
    assert np.all(g[:, 1] < g0[:, 1])  # every column fell
    print(f"256 scenes, {sum(s.n for s in scenes)} particles, {steps} steps", flush=True)
    for k in range(0, 256, 32):
        e, s = alone(scenes[k], steps)
        for f in FIELDS:
            np.testing.assert_array_equal(b.engine.download(b.slots[k], f), e.download(s, f),
                                          err_msg=f"scene {k}: {f}")
        e.close()
        o = Oracle(scenes[k].desc(), 64)
        load_scene(o, scenes[k])
        o.step(steps)
        ref = o.globals()
        rel = np.abs(g[k] - ref) / np.abs(ref)
        print(f"  scene {k} ({scenes[k].n} particles): FP64 error front {rel[0]:.1e} "
              f"PE {rel[1]:.1e} KE {rel[2]:.1e} mass {rel[3]:.1e}", flush=True)
        assert np.all(rel[:2] <= 0.01) and rel[2] <= 0.01 and rel[3] <= 1e-6
    b.close()


def test_served_poisson_stream_results_equal_alone():
    reqs = request_mix(n=24, seed=1, rate=None, scale_steps=0.02)
    big3d = [r for r in reqs if r.kind == "wc3d" and r.scene.n >= 200_000]
    assert len(big3d) >= 3
    cap = plan_capacity(reqs)
    # capacity first (all at once), then the stream at 0.7 x that rate
    srv = Server(cap, chunk=25, max_scenes=24)
    srv.run(reqs, realtime=False)
    rate = 0.7 * len(reqs) / srv.wall
    srv.close()
    stream = request_mix(n=24, seed=1, rate=rate, scale_steps=0.02)
    assert [r.scene.n for r in stream] == [r.scene.n for r in reqs]
    srv = Server(cap, chunk=25, max_scenes=24)
    done = srv.run(stream, realtime=True)
    srv.close()
    assert len(done) == 24
    overlap = sum(1 for r in done for q in done
                  if q is not r and q.admitted < r.finished and r.admitted < q.finished)
    assert overlap > 0  # requests were on the GPU together
    lat = sorted(r.latency for r in done)
    print(f"served 24 requests at {rate:.2f} req/s: p50 {lat[11]:.3f} s, p99 {lat[-1]:.3f} s", flush=True)
    for r in done:
        assert r.result["steps"] == r.steps
        e, s = alone(r.scene, r.steps)
        np.testing.assert_array_equal(r.result["globals"], e.globals(s),
                                      err_msg=f"request {r.rid} ({r.kind}, {r.scene.n} particles)")
        e.close()
