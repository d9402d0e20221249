"""mu(I) granular SPH solver (reference: eng/solver_sph_muI.py:6-158).

Material of ``mat_soil[0]``; sound speed 24 m/s (:23-25).  one_step: soil
velocity gradient / pressure / mu(I) stress, dummy-wall extrapolation of
velocity, density and stress, then the stress acceleration with viscous
damping (:195-265); advect_something regularises the stress (:267-289).
"""
import math

from .solver_sph_base import SPHBase


class MUISPHSolver(SPHBase):
    def __init__(self, particle_system):
        super().__init__(particle_system)
        mat = self.ps.mat_soil[0]
        self.density0 = mat["density0"]
        self.coh = mat["cohesion"]
        self.fric = mat["friction"] / 180 * math.pi
        self.E = mat["EYoungMod"]
        self.poi = mat["poison"]
        self.dila = mat["dilatancy"] / 180 * math.pi
        self.eta_0 = 0.0
        self.mu = math.tan(self.fric)
        self.vsound = 24.0
        self.vsound2 = self.vsound ** 2
