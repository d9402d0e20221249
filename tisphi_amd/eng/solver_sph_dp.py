"""Drucker-Prager elastoplastic SPH solver (reference: eng/solver_sph_dp.py).

Material of ``mat_soil[0]``; sound speed sqrt(E/rho0) (:20-22), derived
constants (:25-29), geostatic initial stress (init_stress,
solver_sph_base.py:249-260, applied when the particles are generated).
The return mapping (adapt_stress, :73-96), plastic flag (:98-109) and the
Bui (2008) stress rate (:171-208) run in FP64 on the device.
"""
import math

import numpy as np

from .solver_sph_base import SPHBase


class DPSPHSolver(SPHBase):
    def __init__(self, particle_system):
        super().__init__(particle_system)
        mat = self.ps.mat_soil[0]
        self.density0 = mat["density0"]
        self.coh = mat["cohesion"]
        self.fric = mat["friction"] / 180 * math.pi
        self.E = mat["EYoungMod"]
        self.poi = mat["poison"]
        self.dila = mat["dilatancy"] / 180 * math.pi
        self.vsound2 = self.E / self.density0
        self.vsound = math.sqrt(self.vsound2)
        self.eps_f = 1e-4
        t = math.tan(self.fric)
        self.alpha_fric = t / math.sqrt(9 + 12 * t ** 2)
        self.k_c = 3 * self.coh / math.sqrt(9 + 12 * t ** 2)
        self.G = self.E / (2 * (1 + self.poi))
        self.K = self.E / (3 * (1 - 2 * self.poi))
        self.De = np.array([[1.0 - self.poi, self.poi, 0.0, self.poi],
                            [self.poi, 1.0 - self.poi, 0.0, self.poi],
                            [0.0, 0.0, (1.0 - 2.0 * self.poi) * 0.5, 0.0],
                            [self.poi, self.poi, 0.0, 1.0 - self.poi]]) * (
            self.E / ((1.0 + self.poi) * (1.0 - 2.0 * self.poi)))
