"""Headless counterpart of the reference's run_simulation.py + ui_sim loop.

    python -m tisphi_amd.run_simulation --scene_file data/scenes/x.json

Same CLI flag (run_simulation.py:8-11).  The GGUI window is replaced by a
headless loop with the reference's step semantics (eng/ui_sim.py:99-242):
``stepsPerRenderUpdate`` steps per frame, ``exitAtStep`` / ``exitAtTime``
end the run (``cur_time = dt * count_step``), ``stopAtStep`` /
``stopAtTime`` / ``pauseAtStart`` have no interactive meaning and are
ignored, exports follow ``exportEveryTime`` / ``exportEveryRender`` with
``exportCSV`` / ``exportVTK`` (PNG frames need the window and are not
written).  ``--steps`` caps the run for scenes that set neither exit rule.
``--checkpoint`` writes a state snapshot at the end and ``--resume``
continues one (an engine addition; the reference has no checkpoint).
As the window loop, the run colours the particles at step 0 and after every
frame (``assign_color``, eng/ui_sim.py:94, 107, 282-286) when colorTitle > 0,
on the device: ``case.ps.pt.val`` / ``case.ps.pt.color`` hold the result.
"""
import argparse
import json
import os
import sys
import time

from . import export
from .eng.configer_builder import SimConfiger
from .eng.simulation import Simulation


def assign_color(case, given_max, given_min, fix_max, fix_min):
    """eng/ui_sim.py:282-286: value, range and colour of the frame."""
    if case.ps.color_title > 0:
        case.solver.assign_value_color()
        case.ps.v_maxmin(given_max, given_min, fix_max, fix_min)
        case.ps.set_color()


def run(scene_path, max_steps=None, out_dir=None, log=print, resume=None, checkpoint=None):
    """Run a scene (or continue the snapshot ``resume``); ``max_steps``
    counts from step 0 like exitAtStep; ``checkpoint`` names a snapshot
    written at the end."""
    if resume:
        case, count_step = Simulation.resume(resume)
        cfg = case.cfg
    else:
        cfg = SimConfiger(scene_file_path=scene_path)
        case = Simulation(cfg)
        count_step = 0
    g = cfg.get_cfg_default
    substeps = max(1, int(g("stepsPerRenderUpdate", 1)))
    exit_at_step = g("exitAtStep", 0)
    exit_at_time = g("exitAtTime", 0)
    every_time = g("exportEveryTime", 0)
    every_render = g("exportEveryRender", 0)
    want = (g("exportCSV", False) or g("exportVTK", False))
    judge_path = (every_render > 0 or every_time > 0) and want and out_dir is not None
    if judge_path:
        os.makedirs(out_dir, exist_ok=True)
        with open(os.path.join(out_dir, "_info.txt"), "w") as f:
            f.write("==== Running Info ====\nTotal particle number: {:,}\ndt={:.6f}s\n\n==== Configure Info ====\n{}"
                    .format(case.ps.particle_num[None], case.solver.dt[None], json.dumps(cfg.config)))
    if not (exit_at_step or exit_at_time or max_steps):
        raise SystemExit("scene sets neither exitAtStep nor exitAtTime: pass --steps")
    dt = case.solver.dt[None]
    start_step = count_step
    next_time = every_time
    while every_time > 0 and next_time <= dt * count_step:
        next_time += every_time  # thresholds a resumed run already passed
    colour = (g("givenMax", -1), g("givenMin", -1), g("fixMax", 0), g("fixMin", 0))
    assign_color(case, *colour)

    def do_export(stamp):
        pos, data = case.ps.dump()
        if g("exportCSV", False):
            export.export_csv(os.path.join(out_dir, f"sim.csv.{stamp}.csv"), pos, data)
        if g("exportVTK", False):
            export.export_vtu(os.path.join(out_dir, f"sim.vtk.{stamp}"), pos, data)

    if judge_path and every_time > 0 and not resume:
        do_export(export.time_stamp(0.0))
    elif judge_path and every_render > 0 and not resume:
        do_export(export.step_stamp(0))
    t0 = time.perf_counter()
    done = ((exit_at_step and count_step >= exit_at_step) or (exit_at_time and dt * count_step >= exit_at_time)
            or (max_steps and count_step >= max_steps))
    while not done:
        n = substeps
        if max_steps:
            n = min(n, max_steps - count_step)
        case.solver.steps(n)
        count_step += n
        assign_color(case, *colour)
        cur_time = dt * count_step
        if judge_path:
            if every_time > 0 and cur_time >= next_time:
                do_export(export.time_stamp(cur_time))
                next_time += every_time
            elif every_render > 0 and count_step % (every_render * substeps) == 0:
                do_export(export.step_stamp(count_step))
        done = ((exit_at_step and count_step >= exit_at_step) or (exit_at_time and cur_time >= exit_at_time)
                or (max_steps and count_step >= max_steps))
    case.solver.sync()
    wall = time.perf_counter() - t0
    glb = case.ps.globals()
    log(f"steps {count_step} ({count_step - start_step} run)  sim time {dt * count_step:.6f}s  "
        f"wall {wall:.3f}s  front {glb[0]:.6f}  PE {glb[1]:.6f}  KE {glb[2]:.6f}  M {glb[3]:.6f}")
    if checkpoint:
        case.ps.save_state(checkpoint, count_step)
    return case, count_step, glb


def main(argv=None):
    ap = argparse.ArgumentParser(description="tiSPHi (MI355X engine, headless)")
    ap.add_argument("--scene_file", default="", help="scene file")
    ap.add_argument("--steps", type=int, default=None, help="stop after this many steps")
    ap.add_argument("--out", default=None, help="export directory")
    ap.add_argument("--resume", default=None, help="continue from this state snapshot (.npz)")
    ap.add_argument("--checkpoint", default=None, help="write a state snapshot (.npz) at the end")
    a = ap.parse_args(argv)
    if not a.scene_file and not a.resume:
        ap.error("--scene_file or --resume is needed")
    run(a.scene_file, a.steps, a.out, resume=a.resume, checkpoint=a.checkpoint)
    return 0


if __name__ == "__main__":
    sys.exit(main())
