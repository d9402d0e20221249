"""Simulation composition (reference: eng/simulation.py:7-22)."""
from .configer_builder import SimConfiger  # noqa: F401  (re-export, as in the reference)
from .particle_system import ParticleSystem
from .solver_sph_dp import DPSPHSolver
from .solver_sph_muI import MUISPHSolver
from .solver_sph_wc import WCSPHSolver


class Simulation:
    def __init__(self, config) -> None:
        self.cfg = config
        self.solver_type = self.cfg.get_cfg("simulationMethod")
        if self.solver_type not in (1, 2, 3):
            raise NotImplementedError(f"Solver type {self.solver_type} has not been implemented.")
        self.ps = ParticleSystem(self.cfg)
        self.solver = self.build_solver()

    @classmethod
    def resume(cls, path):
        """(simulation, step) continued from a snapshot written by
        ``ps.save_state``: the scene is rebuilt from the JSON the snapshot
        holds, then its state is uploaded over the generated one."""
        from .. import snapshot
        meta, arrays = snapshot.load(path)
        cfg = SimConfiger(config=meta["config"])
        cfg.scene_file_path = meta.get("scene_file")
        case = cls(cfg)
        snapshot.restore(case.ps.engine, case.ps.slot, meta, arrays)
        return case, int(meta["step"])

    def build_solver(self):
        if self.solver_type == 1:
            return WCSPHSolver(self.ps)
        elif self.solver_type == 2:
            return MUISPHSolver(self.ps)
        elif self.solver_type == 3:
            return DPSPHSolver(self.ps)
        raise NotImplementedError(f"Solver type {self.solver_type} has not been implemented.")
