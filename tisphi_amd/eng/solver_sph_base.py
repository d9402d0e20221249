"""SPHBase with the reference's interface (eng/solver_sph_base.py:10-793).

``step()`` (:41-51) launches the fused device sequence for one reference
step; the individual reference kernels are exposed as single device passes
for callers that drive them one by one.
"""
import math

import numpy as np

from .particle_system import ScalarField


class SPHBase:
    def __init__(self, particle_system):
        self.ps = particle_system
        cfg = self.ps.cfg
        self.g = np.array(cfg.get_cfg("gravitation"), dtype=np.float64)
        self.flagKernel = cfg.get_cfg("kernel")
        self.flagKernelCorr = cfg.get_cfg("kernelCorrection")
        self.flagTI = cfg.get_cfg("timeIntegration")
        self.flagXSPH = cfg.get_cfg("xsph")
        self.dt_min = cfg.get_cfg("timeStepSizeMin")
        self.dt = ScalarField(lambda: self.ps.scene.dt)
        self.I = np.eye(self.ps.dim)
        self.I3 = np.eye(3)
        self.epsilon = 1e-8
        self.alert_ratio = 0.01

    @staticmethod
    def calc_dt_CFL(CFL_component, vsound, dt_min, smoothing_len):
        dt = CFL_component * smoothing_len / vsound
        return max(dt_min, dt - dt % dt_min)

    # ------------------------------------------------------ time integration
    def step(self):
        self.ps.engine.step(1)

    def steps(self, n):
        """``n`` consecutive step() calls in one submission."""
        self.ps.engine.step(int(n))

    def sync(self):
        self.ps.engine.sync()

    def _pass(self, name):
        self.ps.engine.run(name)

    def substep(self):
        if self.flagTI == 1:
            self.substep_SE()
        elif self.flagTI == 2:
            self.substep_LF()
        elif self.flagTI == 4:
            self.substep_RK()
        else:
            raise NotImplementedError(f"timeIntegration {self.flagTI}")

    def one_step(self):
        self._pass("one_step")

    def init_real2tmp(self):
        self._pass("init_tmp")

    def advect_SE(self):
        self._pass("se")

    def substep_SE(self):
        self.one_step()
        self.advect_SE()

    def advect_LF_half(self):
        self._pass("lf_half")

    def advect_LF(self):
        self._pass("lf_full")

    def substep_LF(self):
        self.one_step()
        self.advect_LF_half()
        self.one_step()
        self.advect_LF()

    def substep_RK(self):
        raise NotImplementedError("RK4 runs inside step(); its stages are not exposed one by one")

    def advect_pos(self):
        self._pass("advect_pos")

    def advect_something(self):
        self._pass("advect_something")

    def calc_kernel_corr(self):
        self.calc_CSPM_f()

    def calc_CSPM_f(self):
        self._pass("shepard")

    def enforce_boundary(self):
        self._pass("enforce_boundary")

    def solve_rigid_body(self):
        """Shape matching of the dynamic rigid blocks (solver_sph_base.py:
        472-510), one device pass (also part of every fused step)."""
        self._pass("solve_rigid")

    def init_rigid_body(self):
        """The rest centres of mass (solver_sph_base.py:467-470) are
        registered with the engine when the scene is uploaded."""

    # ----------------------------------------------------------- visuals
    def assign_value_color(self):
        """Colour value per particle (solver_sph_base.py:721-789), on the
        device (``ps.pt.val`` reads it)."""
        ps = self.ps
        ps._upload_colour_inputs()
        ps._colour_calls += 1
        ps.engine.assign_value_color(ps.slot, ps.color_title, ps.show_bdy, seed=ps._colour_calls)

    # material helpers used by the concrete solvers
    @staticmethod
    def _deg2rad(x):
        return x / 180 * math.pi
