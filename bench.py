"""Benchmark: fluid particle-steps/s of the 3D WCSPH dam break on one MI355X.

Workload (the configuration the headline metric is quoted on): 3D dam
break, water block 1.2 x 0.6 x 0.5 m in a 3.22 x 1.6 x 0.5 m box, particle
radius 0.0025 m (240 x 120 x 100 = 2,880,000 fluid particles) plus
1,646,700 dummy-wall particles, rho0 1000, Tait B 5e4 / gamma 7,
viscosity 0.005, Wendland C2 (kappa 2, kh 1.5), leap-frog, dt 1e-4, f32.
Inputs are generated on the host and resident in HBM before timing starts.

Prints one JSON line; see the keys below.  ``--gpus`` > 1 is refused: this
engine runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s at 2.9M 3D WCSPH; batched/served scene p99 latency vs plain"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
PAIR_KERNEL = "wc_fluid"


def algorithmic_bytes_fluid_pass(dim, n_all, n_fluid):
    """Bytes per launch of the fluid pair pass: every particle's record is
    read once (cell 4, offset 4*dim, v_tmp 4*dim, rho_tmp 4, p 4, m_V 4,
    type 1) and d_density + d_vel written once per fluid particle."""
    rec = 41 if dim == 3 else 33
    out = 16 if dim == 3 else 12
    return rec * n_all + out * n_fluid


def algorithmic_flops_fluid_pass(dim, n_fluid):
    """FP32 flops per launch: candidates x 8 + neighbours x 51 per fluid
    particle (27 cells x 27 / 9 cells x 9 candidates; 92 / 24 lattice
    neighbours)."""
    cand, nb = (729, 92) if dim == 3 else (81, 24)
    return n_fluid * (cand * 8 + nb * 51)


def read_traffic(path):
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def cpu_baseline(scene, max_steps, budget_s=20.0):
    """The FP64 oracle (C++/OpenMP restatement of the reference) timed on
    the host cores on a bounded sample of steps of the same scene: one
    warm-up step, then as many steps as fit in ``budget_s`` (at least one,
    at most ``max_steps``).  Progress goes to stderr after every step."""
    from oracle import oracle as orc
    cores = min(len(os.sched_getaffinity(0)), 16)
    orc.set_threads(cores)
    o = orc.Oracle(scene.desc(), 64)
    orc.load_scene(o, scene)
    t0 = time.perf_counter()
    o.step(1)  # warm-up (first sort of the generated order)
    t1 = time.perf_counter() - t0
    print(f"bench: cpu baseline warm-up step {t1:.1f} s on {cores} threads", file=sys.stderr, flush=True)
    steps = int(max(1, min(max_steps, budget_s // max(t1, 1e-3))))
    t0 = time.perf_counter()
    for k in range(steps):
        o.step(1)
        print(f"bench: cpu baseline step {k + 1}/{steps} at {time.perf_counter() - t0:.1f} s",
              file=sys.stderr, flush=True)
    dt = time.perf_counter() - t0
    return dict(value=scene.n_flow * steps / dt, unit="particle-steps/s", cores=cores, kind="port",
                sample=f"{steps} step(s) of the full workload after 1 warm-up step, FP64 oracle, "
                       f"{dt:.1f} s, {cores} OpenMP threads of {os.cpu_count()} logical CPUs")


def granular(steps=500):
    """The 2D granular column collapse (80,000 soil + 4,818 dummy particles,
    dt 1e-5) under mu(I) and Drucker-Prager: particle-steps/s of the fused
    step (hipGraph replay), after 2 warm-up steps."""
    from tisphi_amd import _lib as L
    from tisphi_amd.eng.configer_builder import SimConfiger
    from tisphi_amd.scene import build_scene
    out = {}
    for key, name in (("mui", "column_collapse_mui"), ("dp", "column_collapse_dp")):
        sc = build_scene(SimConfiger(os.path.join(ROOT, "tisphi_amd", "scenes", name + ".json")))
        e = L.Engine(sc.method, sc.geo.dim, sc.n, sc.geo.grid_num_total, 1)
        s = e.add_scene(sc.desc())
        L.upload_scene(e, s, sc)
        e.step(2)
        e.sync()
        t0 = time.perf_counter()
        e.step(steps)
        e.sync()
        t = time.perf_counter() - t0
        out[key] = dict(particles=sc.n, soil_particles=sc.n_flow, steps=steps, ms_per_step=t * 1e3 / steps,
                        particle_steps_per_s=sc.n_flow * steps / t)
        e.close()
        print(f"bench: granular {key} {t * 1e3 / steps:.3f} ms/step", file=sys.stderr, flush=True)
    return out


def batched_vs_plain(n_scenes=256, steps=100, plain_steps=10):
    """256 uneven 2D dam breaks (seed 0): scene-steps/s advanced together in
    one engine (hipGraph replay) and one scene per engine, one after the
    other (the one-scene-per-launch path, best of eager and graph)."""
    from tisphi_amd.batch import build_family, dam_break_family, run_batched, run_plain
    scenes = build_family(dam_break_family(n_scenes, seed=0))
    t_b, b = run_batched(scenes, steps, graph=True)
    b.close()
    print(f"bench: batched {n_scenes} scenes x {steps} steps in {t_b:.3f} s", file=sys.stderr, flush=True)
    best = None
    for graph in (False, True):
        t_p, engs = run_plain(scenes, plain_steps, graph=graph)
        for e in engs:
            e.close()
        rate = n_scenes * plain_steps / t_p
        best = rate if best is None or rate > best else best
        print(f"bench: plain graph={graph} {n_scenes} scenes x {plain_steps} steps in {t_p:.3f} s",
              file=sys.stderr, flush=True)
    batched = n_scenes * steps / t_b
    return dict(scenes=n_scenes, particles=sum(s.n for s in scenes),
                batched_scene_steps_per_s=batched, plain_scene_steps_per_s=best,
                speedup=batched / best)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--scene", default="dam_break_3d")
    ap.add_argument("--graph", action="store_true", help="time with hipGraph replay (no per-kernel events)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-vs-plain measurement")
    ap.add_argument("--no-serve", action="store_true", help="skip the served-stream measurement")
    ap.add_argument("--no-granular", action="store_true", help="skip the granular (mu(I), DP) measurement")
    ap.add_argument("--nl", type=int, default=2, choices=(0, 1, 2),
                    help="neighbour search: 0 stencil walk per pass, 1 lists, 2 LDS-built lists (default)")
    ap.add_argument("--block", type=str, default="256,256", help="pair,elem threads per block")
    ap.add_argument("--tune", action="store_true",
                    help="pick block sizes and list mode with the autotuner (untimed steps) instead")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy a 256 MiB buffer 3 times (reference kernel for the HBM byte counters)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        a = torch.ones(64 * 1024 * 1024, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print("copied 256 MiB x3", flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    if int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: multi-process launch is not supported (one GPU)", file=sys.stderr)
        sys.exit(2)

    from tisphi_amd import _lib as L
    from tisphi_amd.eng.configer_builder import SimConfiger
    from tisphi_amd.scene import build_scene

    sc = build_scene(SimConfiger(os.path.join(ROOT, "tisphi_amd", "scenes", args.scene + ".json")))
    print(f"bench: {sc.n} particles generated", file=sys.stderr, flush=True)
    dim = sc.geo.dim
    say = lambda m: print("bench: " + m, file=sys.stderr, flush=True)  # noqa: E731
    eng = L.Engine(sc.method, dim, sc.n, sc.geo.grid_num_total, 1)
    say(f"engine holds {eng.bytes / 2**30:.2f} GiB of device memory")
    slot = eng.add_scene(sc.desc())
    L.upload_scene(eng, slot, sc)
    say("scene uploaded")
    eng.set_graph(args.graph)
    # first steps one at a time, each reported (the first also captures graphs)
    for k in range(2):
        t1 = time.perf_counter()
        eng.step(1)
        eng.sync()
        say(f"step {k + 1} in {1e3 * (time.perf_counter() - t1):.1f} ms")
    if args.tune:
        # launch shape: bitwise-neutral, picked by timing real steps (untimed)
        from tisphi_amd.autotune import tune
        tuned = tune(eng, steps=2, quick=True, use_cache=False, log=say)
        shape = tuned["shape"]
    else:
        pair, elem = (int(v) for v in args.block.split(","))
        shape = (pair, elem, args.nl)
        eng.set_block(pair, elem)
        eng.set_neighbor_list(shape[2])
    eng.step(args.warmup)
    eng.sync()
    say(f"{args.warmup} warm-up steps done")
    if not args.graph:
        eng.timing(True)  # HIP events around every launch on the engine stream
    eng.sync()
    t0 = time.perf_counter()
    eng.step(args.steps)
    eng.sync()
    t = time.perf_counter() - t0
    value = sc.n_flow * args.steps / t
    print(f"bench: gpu {args.steps} steps in {t:.3f} s", file=sys.stderr, flush=True)
    ms_pair, n_pair = eng.kernel_time(PAIR_KERNEL) if not args.graph else (0.0, 0)
    kernels = {}
    if not args.graph:
        for k in ("hash", "scan", "scatter", "rank_move", "shepard", "nl_build", "init_tmp", "wc_eos", "wc_bdy",
                  "wc_fluid", "lf_half", "lf_full", "advect_pos", "chk_density"):
            ms, n = eng.kernel_time(k)
            if n:
                kernels[k] = round(ms * n / args.steps, 5)
    status = eng.status(slot)
    g = eng.globals(slot)
    eng.close()

    alg = algorithmic_bytes_fluid_pass(dim, sc.n, sc.n_flow)
    flops = algorithmic_flops_fluid_pass(dim, sc.n_flow)
    traffic = read_traffic(os.path.join(ROOT, "profiles", f"wc_fluid_{args.scene}_pmc.txt"))
    roof = None
    if ms_pair > 0:
        ach = alg / (ms_pair * 1e-3) / 1e9
        roof = dict(bound="hbm", achieved=round(ach, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(ach / HBM_PEAK_GBS, 5), traffic=traffic,
                    kernel=f"k_wc_fluid<{dim}>", launches=n_pair, avg_launch_ms=round(ms_pair, 5),
                    algorithmic_bytes_per_launch=alg,
                    valu=dict(achieved_tflops=round(flops / (ms_pair * 1e-3) / 1e12, 3),
                              peak_tflops=FP32_PEAK_TFLOPS,
                              frac=round(flops / (ms_pair * 1e-3) / 1e12 / FP32_PEAK_TFLOPS, 5),
                              flops_per_launch=flops))
    out = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": t * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: lattice dam break generated from the scene JSON",
        "config": {"workload": f"3D WCSPH dam break, {sc.n_flow:,} fluid + {sc.n - sc.n_flow:,} dummy "
                               f"particles, r={sc.geo.particle_radius:g} m, dt={sc.dt:g} s (CFL 0.2 h/c), "
                               f"Wendland C2, leap-frog"
                   if args.scene == "dam_break_3d" else args.scene,
                   "particles": sc.n, "fluid_particles": sc.n_flow,
                   "cells": sc.geo.grid_num_total, "parallelism": "single GPU",
                   "launch_shape": {"pair_block": shape[0], "elem_block": shape[1],
                                    "neighbour_lists": ["walk", "global", "lds"][shape[2]]}},
        "roofline": roof,
        "kernel_ms_per_step": kernels,
        "status": {"escaped": status[0], "nan": status[1]},
        "globals_after": {"front": g[0], "potential": g[1], "kinetic": g[2], "mass": g[3]},
    }
    if not args.no_granular:
        out["granular"] = granular()
    if not args.no_batch:
        out["batched"] = batched_vs_plain()
    if not args.no_serve:
        from tisphi_amd.serve import serve_bench
        out["served"] = serve_bench(log=lambda m: print("bench: " + m, file=sys.stderr, flush=True))
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sc, args.cpu_steps)
        out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
