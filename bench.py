"""Benchmark: fluid particle-steps/s of the 3D WCSPH dam break on one MI355X.

Workload (the configuration the headline metric is quoted on): 3D dam
break, water block 1.2 x 0.6 x 0.5 m in a 3.22 x 1.6 x 0.5 m box, particle
radius 0.0025 m (240 x 120 x 100 = 2,880,000 fluid particles) plus
1,646,700 dummy-wall particles, rho0 1000, Tait B 5e4 / gamma 7,
viscosity 0.005, Wendland C2 (kappa 2, kh 1.5), leap-frog, f32.  The time
step is the solver's CFL step 0.2 h / c = 2.5e-5 s (1e-4 s is unstable at
this resolution in the reference's algorithm itself); a step's work does not
depend on it.  Inputs are generated on the host and resident in HBM before
timing starts.

Beside the headline: the 2D granular column collapse (mu(I), DP), 256
batched 2D dam breaks against one scene per engine, and a served request
stream against the plain path.  The CPU-baseline leg times the C++ oracle
(the reference step restated, built for the host with -march=native) and,
as the checker, states the FP64-oracle error of every GPU result above.

Prints one JSON line; see the keys below.  ``--gpus`` > 1 is refused: this
engine runs on one GPU.
"""
import argparse
import json
import os
import statistics
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s at 2.9M 3D WCSPH; batched/served scene p99 latency vs plain"
HBM_PEAK_GBS = 8000.0
# FP32 VALU peak without packed FP32 (the library is built without it):
# 256 CUs x 4 SIMDs x 16 lanes x 2 flops (FMA) x 2.4 GHz
FP32_PEAK_TFLOPS_NONPACKED = 78.6
PAIR_KERNEL = "wc_fluid"
FP64_STEPS_3D = 22  # FP64-oracle check of the 3D run: 2 warm-up + 20 timed CPU steps


def say(m):
    print("bench: " + m, file=sys.stderr, flush=True)


def algorithmic_bytes_fluid_pass(dim, n_all, n_fluid):
    """Bytes per launch of the fluid pair pass: every particle's record is
    read once (cell 4, offset 4*dim, v_tmp 4*dim, rho_tmp 4, p 4, m_V 4,
    type 1) and d_density + d_vel written once per fluid particle."""
    rec = 41 if dim == 3 else 33
    out = 16 if dim == 3 else 12
    return rec * n_all + out * n_fluid


def algorithmic_bytes_step(dim, n_fluid, n_dummy, cells):
    """Bytes per WC leap-frog step when every reference pass reads its
    inputs and writes its outputs once per particle (f32 SoA fields; the
    pass-by-pass worksheet: 561 B per fluid and 381 B per dummy particle in
    3D, 453 / 309 B in 2D, plus 8 B per cell)."""
    f, d = (561, 381) if dim == 3 else (453, 309)
    return f * n_fluid + d * n_dummy + 8 * cells


def read_kv(path):
    """``key: value`` lines of a committed profile summary (numbers only)."""
    out = {}
    try:
        with open(path) as f:
            for line in f:
                if ":" not in line:
                    continue
                k, v = line.split(":", 1)
                try:
                    out[k.strip()] = float(v.split()[0])
                except (ValueError, IndexError):
                    pass
    except OSError:
        pass
    return out


# ------------------------------------------------------------- GPU legs
def granular(steps=500, check_steps=300):
    """The 2D granular column collapse (80,000 soil + 4,818 dummy particles,
    dt 1e-5) under mu(I) and Drucker-Prager: particle-steps/s of the fused
    step (hipGraph replay) after 2 warm-up steps; then the same scene again
    from its initial state for ``check_steps`` steps, whose globals the CPU
    leg checks against FP64."""
    from tisphi_amd import _lib as L
    from tisphi_amd.eng.configer_builder import SimConfiger
    from tisphi_amd.scene import build_scene
    out, samples = {}, []
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
        L.upload_scene(e, s, sc)
        e.step(check_steps)
        e.sync()
        samples.append((("granular", key), sc, check_steps, e.globals(s)))
        e.close()
        say(f"granular {key} {t * 1e3 / steps:.3f} ms/step")
    return out, samples


STEP_KERNELS = ("hash", "scan", "scatter", "rank_move", "shepard", "nl_build", "init_tmp", "wc_eos", "wc_bdy",
                "wc_fluid", "lf_half", "lf_full", "advect_pos", "chk_density", "finish")


def kernel_breakdown(eng, steps):
    """Microseconds per step of each kernel of the step, eager launches with
    HIP events around each (the engine is left in graph mode)."""
    eng.set_graph(False)
    eng.timing(True)
    eng.step(steps)
    eng.sync()
    out = {}
    for k in STEP_KERNELS:
        ms, n = eng.kernel_time(k)
        if n:
            out[k] = round(1e3 * ms * n / steps, 2)
    eng.timing(False)
    eng.set_graph(True)
    return out


def batched_vs_plain(n_scenes=256, steps=100, plain_steps=10, n_check=8):
    """256 uneven 2D dam breaks (seed 0): scene-steps/s advanced together in
    one engine (hipGraph replay) and one scene per engine, one after the
    other (the one-scene-per-launch path, best of eager and graph); the
    globals of ``n_check`` evenly spaced scenes after the batched run go to
    the CPU leg's FP64 check."""
    from tisphi_amd.batch import build_family, dam_break_family, run_batched, run_plain
    scenes = build_family(dam_break_family(n_scenes, seed=0))
    t_b, b = run_batched(scenes, steps, graph=True)
    picks = list(range(0, n_scenes, n_scenes // n_check))[:n_check]
    samples = [(("batched", k), scenes[k], steps + 1, b.engine.globals(b.slots[k])) for k in picks]
    # where the batched step's time goes: its kernels one by one (the sum
    # against the graph step says what launch boundaries cost)
    kern = kernel_breakdown(b.engine, 20)
    b.close()
    say(f"batched {n_scenes} scenes x {steps} steps in {t_b:.3f} s")
    best = None
    for graph in (False, True):
        t_p, engs = run_plain(scenes, plain_steps, graph=graph)
        for e in engs:
            e.close()
        rate = n_scenes * plain_steps / t_p
        best = rate if best is None or rate > best else best
        say(f"plain graph={graph} {n_scenes} scenes x {plain_steps} steps in {t_p:.3f} s")
    batched = n_scenes * steps / t_b
    return dict(scenes=n_scenes, particles=sum(s.n for s in scenes),
                batched_scene_steps_per_s=batched, plain_scene_steps_per_s=best,
                speedup=batched / best, graph_step_us=round(1e6 * t_b / steps, 2),
                kernel_us_per_step=kern, kernel_sum_us=round(sum(kern.values()), 2),
                fp64_check_scenes=picks), samples


# ----------------------------------------------------- CPU-baseline leg
def build_native_oracle():
    """Compile the oracle for this host (-O3 -march=native -fopenmp, default
    FP contraction); fall back to the portable test build."""
    lib = os.path.join(ROOT, "oracle", "_native", "liboracle.so")
    try:
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "native"], check=True,
                       capture_output=True, timeout=120)
        return lib, "g++ -O3 -march=native -fopenmp (host build)"
    except (OSError, subprocess.SubprocessError) as e:
        say(f"native oracle build failed ({e}); timing the portable build")
        return os.path.join(ROOT, "oracle", "liboracle.so"), "g++ -O3 -march=x86-64-v3 -ffp-contract=off"


def rel_err(g, ref):
    g, ref = np.asarray(g, dtype=np.float64), np.asarray(ref, dtype=np.float64)
    e = np.abs(g - ref) / np.maximum(np.abs(ref), 1e-300)
    return dict(zip(("front", "potential", "kinetic", "mass"), (float(x) for x in e)))


def cpu_baseline(scene, gpu_globals_3d, samples, steps=20, runs=3):
    """The CPU leg.  Baseline: the oracle (the reference step restated in
    C++/OpenMP; Taichi, the reference's own CPU backend, is not installed)
    built for this host, on its cores: float32 (the engine's arithmetic) for
    ``runs`` consecutive runs of ``steps`` steps after 2 warm-up steps,
    median reported; float64 for one run of ``steps`` steps after the same
    warm-up.  Checker: the FP64 state after those 2 + ``steps`` steps gives
    the 3D result's FP64 error, and an FP64 run of every sample (kind,
    scene, steps, GPU globals) gives theirs."""
    from oracle import oracle as orc  # test infrastructure: baseline and checker only
    lib, flags = build_native_oracle()
    orc.use_library(lib)
    cores = min(len(os.sched_getaffinity(0)), 16)
    orc.set_threads(cores)
    res = {}
    for prec in (32, 64):
        o = orc.Oracle(scene.desc(), prec)
        orc.load_scene(o, scene)
        o.step(2)  # warm-up: the first sorts of the generated order
        times = []
        for k in range(runs if prec == 32 else 1):
            t0 = time.perf_counter()
            for _ in range(steps):
                o.step(1)
            times.append(time.perf_counter() - t0)
            say(f"cpu baseline f{prec} run {k + 1}: {steps} steps in {times[-1]:.1f} s")
        res[prec] = times
        if prec == 64:
            fp64_3d = rel_err(gpu_globals_3d, o.globals())
        del o
    med = statistics.median(res[32])
    out = dict(value=scene.n_flow * steps / med, unit="particle-steps/s", cores=cores, kind="port",
               sample=f"float32 oracle (the reference step restated in C++/OpenMP), {runs} runs of "
                      f"{steps} steps of the full workload after 2 warm-up steps, median "
                      f"{med:.1f} s; {cores} threads of {os.cpu_count()} logical CPUs; {flags}",
               f32_run_s=res[32], f64_value=scene.n_flow * steps / res[64][0], f64_run_s=res[64][0],
               build=flags)
    errs = {}
    for key, sc, st, g in samples:
        o = orc.Oracle(sc.desc(), 64)
        orc.load_scene(o, sc)
        o.step(st)
        errs[key] = dict(steps=st, error=rel_err(g, o.globals()))
        del o
        say(f"fp64 check {key}: {errs[key]['error']}")
    return out, fp64_3d, errs


# ------------------------------------------------------------------ main
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--cpu-runs", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the FP64 checks")
    ap.add_argument("--scene", default="dam_break_3d")
    ap.add_argument("--graph", action="store_true", help="time with hipGraph replay (no per-kernel events)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-vs-plain measurement")
    ap.add_argument("--no-serve", action="store_true", help="skip the served-stream measurement")
    ap.add_argument("--no-granular", action="store_true", help="skip the granular (mu(I), DP) measurement")
    ap.add_argument("--nl", type=int, default=2, choices=(0, 1, 2),
                    help="neighbour search: 0 stencil walk per pass, 1 lists, 2 LDS-built lists (default)")
    ap.add_argument("--block", type=str, default="256,256", help="pair,elem threads per block")
    ap.add_argument("--cell-tile", type=int, default=256, choices=(128, 256, 512),
                    help="candidates per LDS-staged window of the list build")
    ap.add_argument("--tune", action="store_true",
                    help="pick block sizes, list mode and cell tile with the autotuner (untimed steps) instead")
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
    say(f"{sc.n} particles generated")
    dim = sc.geo.dim
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
        shape = (pair, elem, args.nl, args.cell_tile)
        eng.set_block(pair, elem)
        eng.set_neighbor_list(shape[2])
        eng.set_cell_tile(shape[3])
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
    say(f"gpu {args.steps} steps in {t:.3f} s")
    ms_pair, n_pair = eng.kernel_time(PAIR_KERNEL) if not args.graph else (0.0, 0)
    kernels = {}
    if not args.graph:
        for k in STEP_KERNELS:
            ms, n = eng.kernel_time(k)
            if n:
                kernels[k] = round(ms * n / args.steps, 5)
        eng.timing(False)
    status = eng.status(slot)
    g = eng.globals(slot)
    # the FP64 check of this workload: the same scene from its initial state
    # for as many steps as the CPU leg's FP64 run takes
    L.upload_scene(eng, slot, sc)
    eng.step(FP64_STEPS_3D)
    eng.sync()
    g_check = eng.globals(slot)
    eng.close()

    alg = algorithmic_bytes_fluid_pass(dim, sc.n, sc.n_flow)
    prof = os.path.join(ROOT, "profiles")
    pmc = read_kv(os.path.join(prof, f"wc_fluid_{args.scene}_pmc.txt"))
    cnt = read_kv(os.path.join(prof, f"wc_fluid_{args.scene}_counters.txt"))
    roof = None
    step_bytes = algorithmic_bytes_step(dim, sc.n_flow, sc.n - sc.n_flow, sc.geo.grid_num_total)
    if ms_pair > 0:
        ach = alg / (ms_pair * 1e-3) / 1e9
        roof = dict(bound="hbm", achieved=round(ach, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(ach / HBM_PEAK_GBS, 5), traffic=pmc.get("traffic_bytes_per_launch"),
                    traffic_source=f"profiles/wc_fluid_{args.scene}_pmc.txt: rocprofv3 FETCH_SIZE and "
                                   f"WRITE_SIZE runs of bench.py on this workload, scaled by a copy kernel",
                    kernel=f"k_wc_fluid<{dim}>", launches=n_pair, avg_launch_ms=round(ms_pair, 5),
                    algorithmic_bytes_per_launch=alg)
        if cnt:
            # what does bound the pass (committed counter runs of this kernel)
            roof["binding"] = dict(
                source=f"profiles/wc_fluid_{args.scene}_counters.txt",
                valu_busy=cnt.get("valu_busy"),
                valu_instructions_per_pair=cnt.get("valu_instructions_per_pair"),
                valu_flop_frac_nonpacked_peak=cnt.get("valu_flop_frac_nonpacked_peak"),
                ta_busy=cnt.get("ta_busy"), td_busy=cnt.get("td_busy"),
                l2_hit=cnt.get("l2_hit"))
    step = dict(algorithmic_bytes=step_bytes, achieved_GBps=round(step_bytes / (t / args.steps) / 1e9, 2),
                frac=round(step_bytes / (t / args.steps) / 1e9 / HBM_PEAK_GBS, 5),
                hbm_floor_ms=round(step_bytes / (HBM_PEAK_GBS * 1e9) * 1e3, 4))
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
                                    "neighbour_lists": ["walk", "global", "lds"][shape[2]],
                                    "cell_tile": shape[3]}},
        "roofline": roof,
        "step_roofline": step,
        "kernel_ms_per_step": kernels,
        "status": {"escaped": status[0], "nan": status[1]},
        "globals_after": {"front": g[0], "potential": g[1], "kinetic": g[2], "mass": g[3]},
        "all_particle_steps_per_s": sc.n * args.steps / t,
    }
    samples = []
    if not args.no_granular:
        out["granular"], s = granular()
        samples += s
    if not args.no_batch:
        out["batched"], s = batched_vs_plain()
        samples += s
    if not args.no_serve:
        from tisphi_amd.serve import serve_bench
        out["served"], s = serve_bench(log=say)
        samples += [(("served", k), scn, st, gg) for k, scn, st, gg in s]
    if not args.no_cpu:
        base, fp64_3d, errs = cpu_baseline(sc, g_check, samples, steps=args.cpu_steps, runs=args.cpu_runs)
        out["cpu_baseline"] = base
        out["gpu_over_cpu"] = value / base["value"]
        out["fp64_error"] = dict(steps=FP64_STEPS_3D, error=fp64_3d,
                                 note="GPU globals (front, potential, kinetic energy, mass) of this "
                                      "workload after the steps, relative to the FP64 oracle")
        for (leg, k), e in errs.items():
            if leg == "granular":
                out["granular"][k]["fp64_error"] = e
            elif leg == "batched":
                out["batched"].setdefault("fp64_error", {})[str(k)] = e
            else:
                out["served"]["fp64_sample"][k]["fp64_error"] = e["error"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
