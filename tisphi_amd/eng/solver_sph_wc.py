"""Weakly compressible SPH solver (reference: eng/solver_sph_wc.py:6-133).

Material of ``mat_fluid[0]``; sound speed 60 m/s and dt from the CFL rule
(:12-21).  one_step: Tait EOS, dummy-wall extrapolation, then d_density and
d_vel of the fluid (:82-127) — executed by the engine's WC kernels.
"""
from .solver_sph_base import SPHBase


class WCSPHSolver(SPHBase):
    def __init__(self, particle_system):
        super().__init__(particle_system)
        mat = self.ps.mat_fluid[0]
        self.density0 = mat["density0"]
        self.viscosity = mat["viscosity"]
        self.stiffness = mat["stiffness"]
        self.exponent = mat["exponent"]
        self.vsound = 60
