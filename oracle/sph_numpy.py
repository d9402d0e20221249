"""Independent FP64 numpy restatement of the reference step, brute force.

TEST INFRASTRUCTURE ONLY (tests/ import it to pin the C++ oracle).

Differs from the C++ oracle on purpose in everything that does not change
the mathematics: absolute float64 positions instead of cell + offset, an
O(N^2) all-pairs neighbour search instead of the cell grid, vectorised pair
sums in arbitrary order, full 3x3 tensors.  Agreement of the two to ~1e-10
therefore pins the C++ oracle's grid, sort, pass order and per-pair
formulas against a second reading of the reference:

  step order ............ eng/solver_sph_base.py:41-51, 79-120, 228-247
  RK4 ................... eng/solver_sph_base.py:134-180
  collision boundary .... eng/solver_sph_base.py:524-601
  Shepard ............... eng/solver_sph_base.py:386-398
  Wendland / cubic ...... eng/solver_sph_base.py:298-358
  WC one_step ........... eng/solver_sph_wc.py:33-71, 82-133
  muI one_step/advect ... eng/solver_sph_muI.py:171-189, 195-289
  DP one_step/advect .... eng/solver_sph_dp.py:41-118, 171-296
  dummy tasks ........... eng/solver_sph_base.py:647-669, 713-715
  rigid bodies .......... eng/solver_sph_wc.py:68-71, 90-106, 125-126;
                          eng/solver_sph_base.py:472-518 (the fluid's pair
                          force scattered onto rigid neighbours as the
                          reference writes it; polar factor by SVD)
Parity-mode decisions D1-D3 (two-phase loops) are applied as in the C++
oracle.  Only for small scenes (N up to a few thousand).
"""
import math

import numpy as np

FLUID, SOIL, RIGID, DUMMY, REP = 1, 2, 11, -1, -2


def is_flow(t):
    return (t == FLUID) | (t == SOIL)


def is_bdy(t):
    return (t == DUMMY) | (t == REP)


class NumpySPH:
    def __init__(self, desc, pos, vel, density, mass, m_V, mtype, stress6=None, obj=None,
                 rigid=()):
        self.d = dict(desc)
        self.dim = desc["dim"]
        self.x = np.array(pos, dtype=np.float64)
        self.v = np.array(vel, dtype=np.float64)
        self.rho = np.array(density, dtype=np.float64)
        self.mass = np.array(mass, dtype=np.float64)
        self.mV = np.array(m_V, dtype=np.float64)
        self.t = np.array(mtype, dtype=np.int64)
        n = self.n = len(self.rho)
        self.sig = np.zeros((n, 3, 3))
        if stress6 is not None:
            s = np.asarray(stress6, dtype=np.float64)
            self.sig[:, 0, 0], self.sig[:, 1, 1], self.sig[:, 2, 2] = s[:, 0], s[:, 1], s[:, 2]
            self.sig[:, 0, 1] = self.sig[:, 1, 0] = s[:, 3]
            self.sig[:, 1, 2] = self.sig[:, 2, 1] = s[:, 4]
            self.sig[:, 0, 2] = self.sig[:, 2, 0] = s[:, 5]
        self.p = np.zeros(n)
        self.rhot = np.zeros(n)
        self.vt = np.zeros((n, 3))
        self.sigt = np.zeros((n, 3, 3))
        self.dsig = np.zeros((n, 3, 3))
        self.drho = np.zeros(n)
        self.dv = np.zeros((n, 3))
        self.cspm = np.ones(n)
        self.seq = np.zeros(n)
        self.seqp = np.zeros(n)
        self.dseq = np.zeros(n)
        self.dseqp = np.zeros(n)
        self.flag = np.zeros(n, dtype=np.int64)
        self.g = np.array(desc["gravity"], dtype=np.float64)
        self.h = desc["smoothing_len"]
        self.rs = desc["support_radius"]
        self.dt = desc["dt"]
        self.rho0 = desc["density0"]
        t = math.tan(desc["friction"])
        self.alpha = t / math.sqrt(9 + 12 * t * t)
        self.kc = 3 * desc["cohesion"] / math.sqrt(9 + 12 * t * t)
        self.G = desc["young"] / (2 * (1 + desc["poisson"]))
        self.K = desc["young"] / (3 * (1 - 2 * desc["poisson"]))
        self.sin_psi = math.sin(desc["dilatancy"])
        # rigid blocks: object ids, initial positions, dynamic objects' rest
        # centres (calc_rigid_rest_m, solver_sph_base.py:517-518)
        self.obj = np.zeros(n, dtype=np.int64) if obj is None else np.asarray(obj, dtype=np.int64)
        self.x0 = self.x.copy()
        self.rigid = [(r["obj"], np.asarray(r["rest_cm"], dtype=np.float64)) for r in rigid
                      if r["dynamic"]]
        self.rig_dyn = np.zeros(n, dtype=bool)
        for o, _ in self.rigid:
            self.rig_dyn |= (self.t == RIGID) & (self.obj == o)

    # ------------------------------------------------------------- kernels
    def W(self, r):
        h1 = 1.0 / self.h
        q = r * h1
        ok = (r > 1e-8) & (q <= 2.0)
        if self.d["kernel"] == 1:
            k = (7.0 / (4.0 * math.pi) if self.dim == 2 else 21.0 / (2.0 * math.pi)) * h1 ** self.dim
            return np.where(ok, k * (1.0 - 0.5 * q) ** 4 * (1.0 + 2.0 * q), 0.0)
        k = (15.0 / 7.0 / math.pi if self.dim == 2 else 3.0 / 2.0 / math.pi) * h1 ** self.dim
        w = np.where(q <= 1.0, k * (0.5 * q ** 3 - q ** 2 + 2.0 / 3.0), k / 6.0 * (2.0 - q) ** 3)
        return np.where(ok, w, 0.0)

    def gradW(self, xij, r):
        h1 = 1.0 / self.h
        q = r * h1
        ok = (r > 1e-8) & (q <= 2.0)
        rs = np.where(ok, r, 1.0)
        if self.d["kernel"] == 1:
            k = (7.0 / (4.0 * math.pi) if self.dim == 2 else 21.0 / (2.0 * math.pi)) * h1 ** self.dim
            s = k * (1.0 - 0.5 * q) ** 3 * (-5.0 * q) * h1
        else:
            k = (15.0 / 7.0 / math.pi if self.dim == 2 else 3.0 / 2.0 / math.pi) * h1 ** self.dim
            s = np.where(q <= 1.0, k * q * (1.5 * q - 2.0), k * (-0.5 * (2.0 - q) ** 2)) * h1
        return np.where(ok, s / rs, 0.0)[:, None] * xij

    # --------------------------------------------------------- neighbours
    def pairs(self, x=None):
        x = self.x if x is None else x
        d = x[:, None, :] - x[None, :, :]
        r2 = np.einsum("ijk,ijk->ij", d, d)
        m = r2 < self.rs * self.rs
        np.fill_diagonal(m, False)
        i, j = np.nonzero(m)
        xij = x[i] - x[j]
        return i, j, xij, np.sqrt(np.einsum("ij,ij->i", xij, xij))

    def _sum(self, i, vals):
        out = np.zeros((self.n,) + vals.shape[1:])
        np.add.at(out, i, vals)
        return out

    # ------------------------------------------------------------- passes
    def shepard(self):
        i, j, xij, r = self.pairs()
        s = self._sum(i, np.where(is_flow(self.t[j]), self.mV[j] * self.W(r), 0.0))
        self.cspm = np.where(s != 0.0, 1.0 / np.where(s != 0, s, 1.0), 1.0)
        if self.d.get("kernel_corr", 0) == 1:
            self.cspm_L(i, j, xij, r)
        elif self.d.get("kernel_corr", 0) == 2:
            self.Lc = np.zeros((self.n, 3, 3))  # never computed: the zero field

    def cspm_L(self, i, j, xij, r):
        """calc_CSPM_L (solver_sph_base.py:400-423), inverted (np.linalg)."""
        gw = self.gradW(xij, r)
        same = is_flow(self.t[i]) & (self.t[i] == self.t[j])
        T = np.where(same[:, None, None], (self.mV[j][:, None] * -xij)[:, :, None] * gw[:, None, :], 0.0)
        L = self._sum(i, T)
        M = np.tile(np.eye(3), (self.n, 1, 1))
        flow = is_flow(self.t)
        if self.dim == 2:
            d = np.linalg.det(L[:, :2, :2])
            ok = flow & (np.abs(d) > 1e-8)
            M[ok] = 0.0
            M[ok, :2, :2] = np.linalg.inv(L[ok, :2, :2])
        else:
            d = np.linalg.det(L)
            ok = flow & (np.abs(d) > 1e-8)
            M[ok] = np.linalg.inv(L[ok])
        self.Lc = M

    def corr(self, i, gw):
        """kernel_deriv_corr (solver_sph_base.py:375-384): with
        kernelCorrection 2 the branch is `pass`, leaving the zero vector."""
        kc = self.d.get("kernel_corr", 0)
        if kc == 0:
            return gw
        if kc == 2:
            return np.zeros_like(gw)
        return np.einsum("nab,nb->na", self.Lc[i], gw)

    def init_tmp(self):
        real = self.t > 0
        self.rhot[real] = self.rho[real]
        self.vt[real] = self.v[real]
        soil = self.t == SOIL
        self.sigt[soil] = self.sig[soil]

    def _bdy_common(self, i, j, xij, r):
        """Sums over flow neighbours of boundary (and, in WC, rigid)
        particles."""
        solid = is_bdy(self.t[i]) | (self.t[i] == RIGID)
        w = np.where(solid & is_flow(self.t[j]), self.W(r), 0.0)
        sv = self._sum(i, (self.mV[j] * w)[:, None] * self.vt[j])
        return w, sv

    def wc_one_step(self):
        fl = self.t == FLUID
        B, gam = self.d["stiffness"], self.d["exponent"]
        self.p[fl] = np.maximum(B * ((self.rhot[fl] / self.rho0) ** gam - 1.0), 0.0)
        i, j, xij, r = self.pairs()
        bd = is_bdy(self.t) | (self.t == RIGID)  # solver_sph_wc.py:90
        w, sv = self._bdy_common(i, j, xij, r)
        sp = self._sum(i, self.mV[j] * (self.p[j] + self.rhot[j] * self.g[1] * xij[:, 1]) * w)
        self.vt[bd] = 2 * self.v[bd] - sv[bd] * self.cspm[bd, None]
        self.rhot[bd] = self.rho0
        self.p[bd] = np.maximum(sp[bd] * self.cspm[bd], 0.0)
        gw = self.gradW(xij, r)
        sel = (self.t[i] == FLUID)
        u = self.vt[i] - self.vt[j]
        dd = self._sum(i, np.where(sel, self.mV[j] * np.einsum("ij,ij->i", u, self.corr(i, gw)), 0.0))
        tj = self.t[j]
        vis_ok = (tj == FLUID) | is_bdy(tj) | (tj == RIGID)
        c1 = 2 * (self.dim + 2) * self.d["viscosity"]
        mn = np.minimum(np.einsum("ij,ij->i", u, xij), 0.0)
        fac = np.where(tj == FLUID, 1.0, self.rho0 / self.rhot[i])
        vis = np.where(vis_ok, c1 * self.mV[j] * fac * mn / (r ** 2 + 0.01 * self.h ** 2), 0.0)
        pc = -self.rho0 * self.mV[j] * (self.p[i] / self.rhot[i] ** 2 + self.p[j] / self.rhot[j] ** 2)
        dv = self._sum(i, np.where(sel, vis + pc, 0.0)[:, None] * gw) + self._rep(i, j, xij, r)
        self.drho[fl] = dd[fl] * self.rhot[fl]
        self.dv[fl] = dv[fl] + self.g
        # rigid: static d_vel = 0 (:125-126); dynamic g (:105-106) minus the
        # fluid pair terms, scattered from the fluid side (:68-71)
        rig = self.t == RIGID
        self.dv[rig] = 0.0
        if self.rig_dyn.any():
            onto = sel & self.rig_dyn[j]
            back = np.zeros((self.n, 3))
            np.add.at(back, j[onto], ((vis + pc)[:, None] * gw)[onto])
            self.dv[self.rig_dyn] = self.g - back[self.rig_dyn]

    def _rep(self, i, j, xij, r):
        """Repulsive force on each particle from its repulsive neighbours
        (solver_sph_base.py:676-690; boundary 3 / 4, WC and muI only)."""
        if self.d["boundary"] not in (3, 4) or self.d["method"] == 3:
            return np.zeros((self.n, 3))
        dd = self.d["particle_diameter"]
        chi = np.where((r > 0) & (r < dd), 1.0 - r / dd, 0.0)
        gam = r / (0.75 * self.h)
        f = np.where((gam > 0) & (gam <= 2 / 3), 2 / 3,
                     np.where((gam > 2 / 3) & (gam <= 1), 2 * gam - 1.5 * gam ** 2,
                              np.where((gam > 1) & (gam < 2), 0.5 * (2 - gam) ** 2, 0.0)))
        fac = np.where(self.t[j] == REP, 0.01 * self.d["vsound"] ** 2 * chi * f / r ** 2, 0.0)
        return self._sum(i, fac[:, None] * xij)

    def _vgrad(self, i, j, xij, r, sel):
        gw = self.corr(i, self.gradW(xij, r))
        du = (self.vt[j] - self.vt[i]) * self.mV[j][:, None]
        L = self._sum(i, np.where(sel[:, None, None], du[:, :, None] * gw[:, None, :], 0.0))
        dd = self._sum(i, np.where(sel, self.mV[j] * np.einsum("ij,ij->i", self.vt[i] - self.vt[j], gw), 0.0))
        return L, dd, gw

    def _soil_bdy(self, i, j, xij, r, extrapolate_density):
        bd = is_bdy(self.t) | (self.t == RIGID)  # solver_sph_muI.py:96, dp.py:221
        w, sv = self._bdy_common(i, j, xij, r)
        sr = self._sum(i, self.mV[j] * self.rhot[j] * w)
        gx = (self.rhot[j][:, None] * self.g[None, :]) * xij
        M = self.sigt[j].copy()
        for a in range(3):
            M[:, a, a] += gx[:, a]
        ss = self._sum(i, (self.mV[j] * w)[:, None, None] * M)
        self.vt[bd] = 2 * self.v[bd] - sv[bd] * self.cspm[bd, None]
        if extrapolate_density:
            self.rhot[bd] = np.maximum(sr[bd] * self.cspm[bd], self.rho0)
        else:
            self.rhot[bd] = self.rho0
        self.sigt[bd] = ss[bd] * self.cspm[bd, None, None]

    def _soil_dv(self, i, j, xij, r, gw):
        soil = self.t == SOIL
        sel = self.t[i] == SOIL
        M = self.sigt[j] / (self.rhot[j] ** 2)[:, None, None] + self.sigt[i] / (self.rhot[i] ** 2)[:, None, None]
        tmp = (self.mV[j] * self.rhot[j])[:, None] * np.einsum("nab,nb->na", M, gw)
        dv = self._sum(i, np.where(sel[:, None], tmp, 0.0)) + self._rep(i, j, xij, r)
        damp = -5e-5 * np.sqrt(self.d["young"] / (self.rhot * self.h ** 2))
        self.dv[soil] = dv[soil] + self.g + damp[soil, None] * self.vt[soil]
        # rigid: static 0; dynamic g minus the soil terms scattered onto it
        # (solver_sph_muI.py:45-46, solver_sph_dp.py:164-165)
        self.dv[self.t == RIGID] = 0.0
        if self.rig_dyn.any():
            onto = sel & self.rig_dyn[j]
            back = np.zeros((self.n, 3))
            np.add.at(back, j[onto], tmp[onto])
            self.dv[self.rig_dyn] = self.g - back[self.rig_dyn]

    def mui_one_step(self):
        soil = self.t == SOIL
        i, j, xij, r = self.pairs()
        L, dd, gw = self._vgrad(i, j, xij, r, self.t[i] == SOIL)
        self.drho[soil] = dd[soil] * self.rhot[soil]
        c2 = self.d["vsound"] ** 2
        p = np.maximum(c2 * (self.rho - self.rho0), 0.0)
        self.p[soil] = p[soil]
        e = 0.5 * (L + np.transpose(L, (0, 2, 1)))
        en = np.sqrt(0.5 * np.sum(e * e, axis=(1, 2))) + 1e-8
        mu = math.tan(self.d["friction"])
        fac = (self.d["cohesion"] + p * mu) / en
        sigt = fac[:, None, None] * e - p[:, None, None] * np.eye(3)
        self.sigt[soil] = sigt[soil]
        tr = np.trace(e, axis1=1, axis2=2)
        dev = e - (tr / 3.0)[:, None, None] * np.eye(3)
        self.dseq[soil] = np.sqrt(np.sum(dev * dev, axis=(1, 2)) * 2 / 3)[soil]
        self._soil_bdy(i, j, xij, r, True)
        self._soil_dv(i, j, xij, r, gw)

    # Drucker-Prager (solver_sph_dp.py)
    def invariants(self, st):
        I1 = np.trace(st, axis1=1, axis2=2)
        s = st - (I1 / 3.0)[:, None, None] * np.eye(3)
        sJ2 = np.sqrt(0.5 * np.sum(s * s, axis=(1, 2)))
        return s, I1, sJ2, sJ2 + self.alpha * I1 - self.kc

    def adapt(self, st):
        st = st.copy()
        s, I1, sJ2, f = self.invariants(st)
        a = f > 1e-4
        b = a & (f > sJ2)
        st[b] -= ((I1[b] - self.kc / self.alpha) / 3.0)[:, None, None] * np.eye(3)
        s, I1, sJ2, f2 = self.invariants(st)
        with np.errstate(divide="ignore", invalid="ignore"):
            rr = (-I1 * self.alpha + self.kc) / sJ2
        new = rr[:, None, None] * s + (I1 / 3.0)[:, None, None] * np.eye(3)
        st[a] = new[a]
        return st

    def flags(self, st):
        s, I1, sJ2, f = self.invariants(st)
        fl = np.ones(len(f), dtype=np.int64)
        fl[f < -1e-4] = 0
        hi = f > 1e-4
        fl[hi & (f >= sJ2)] = 3
        fl[hi & (f < sJ2)] = 2
        return fl

    def bui2008(self, L, st):
        s, I1, sJ2, f = self.invariants(st)
        e = 0.5 * (L + np.transpose(L, (0, 2, 1)))
        w = 0.5 * (L - np.transpose(L, (0, 2, 1)))
        J = np.einsum("nik,njk->nij", st, w) + np.einsum("nkj,nik->nij", st, w)
        tr = np.trace(e, axis1=1, axis2=2)
        ed = e - (tr / 3.0)[:, None, None] * np.eye(3)
        te = 2.0 * self.G * ed + (self.K * tr)[:, None, None] * np.eye(3)
        plastic = (f >= -1e-4) & (sJ2 > 1e-8)
        sJ2s = np.where(plastic, sJ2, 1.0)
        lam = np.where(plastic, (3.0 * self.alpha * self.K * tr + (self.G / sJ2s) * np.sum(s * e, axis=(1, 2)))
                       / (27.0 * self.alpha * self.K * self.sin_psi + self.G), 0.0)
        tg = lam[:, None, None] * (9.0 * self.K * self.sin_psi * np.eye(3) + (self.G / sJ2s)[:, None, None] * s)
        ds = J + te - tg
        dq = np.sqrt(np.sum(ed * ed, axis=(1, 2)) * 2 / 3)
        gp = sJ2 + 3 * I1 * self.sin_psi
        with np.errstate(divide="ignore", invalid="ignore"):
            ep = np.where(np.abs(ds) > 1e-8, gp[:, None, None] / ds, 0.0) * lam[:, None, None]
        ept = ep - (np.trace(ep, axis1=1, axis2=2) / 3.0)[:, None, None] * np.eye(3)
        dqp = np.where(plastic, np.sqrt(np.sum(ept * ept, axis=(1, 2)) * 2 / 3), 0.0)
        return ds, dq, dqp

    def dp_one_step(self):
        soil = self.t == SOIL
        self.sigt[soil] = self.adapt(self.sigt[soil])
        i, j, xij, r = self.pairs()
        self._soil_bdy(i, j, xij, r, False)
        L, dd, gw = self._vgrad(i, j, xij, r, self.t[i] == SOIL)
        self.drho[soil] = dd[soil] * self.rhot[soil]
        ds, dq, dqp = self.bui2008(L[soil], self.sigt[soil])
        self.dsig[soil] = ds
        self.dseq[soil] = dq
        self.dseqp[soil] = dqp
        self._soil_dv(i, j, xij, r, gw)

    def one_step(self):
        m = self.d["method"]
        if m == 1:
            self.wc_one_step()
        elif m == 2:
            self.mui_one_step()
        else:
            self.dp_one_step()

    def lf_half(self):
        real = self.t > 0
        self.rhot[real] += 0.5 * self.dt * self.drho[real]
        self.mV[real] = self.mass[real] / self.rhot[real]
        self.vt[real] += 0.5 * self.dt * self.dv[real]
        if self.d["method"] == 3:
            soil = self.t == SOIL
            self.sigt[soil] += 0.5 * self.dt * self.dsig[soil]

    def lf_full(self):
        real = self.t > 0
        self.rho[real] += self.dt * self.drho[real]
        self.mV[real] = self.mass[real] / self.rho[real]
        self.v[real] += self.dt * self.dv[real]
        if self.d["method"] == 3:
            soil = self.t == SOIL
            self.sig[soil] += self.dt * self.dsig[soil]

    def advect_pos(self):
        real = self.t > 0
        if self.d["xsph"]:
            i, j, xij, r = self.pairs()
            same = (self.t[i] == self.t[j]) & (self.t[i] > 0)
            xs = self._sum(i, np.where(same, self.mV[j] * self.W(r), 0.0)[:, None] * (self.v[j] - self.v[i]))
            self.x[real] += self.dt * (self.v[real] + 0.5 * xs[real])
        else:
            self.x[real] += self.dt * self.v[real]

    def advect_something(self):
        m = self.d["method"]
        if m == 1:
            fl = self.t == FLUID
            self.rho[fl] = np.maximum(self.rho0, self.rho[fl])
            self.mV[fl] = self.mass[fl] / self.rho[fl]
            return
        soil = self.t == SOIL
        self.rho[soil] = np.maximum(self.rho0, self.rho[soil])
        self.mV[soil] = self.mass[soil] / self.rho[soil]
        if m == 2:
            self.seq[soil] += self.dt * self.dseq[soil]
            self.sig[soil] = self.sigt[soil]
            i, j, xij, r = self.pairs()
            same = (self.t[i] == self.t[j]) & (self.t[i] == SOIL)
            reg = self._sum(i, np.where(same, self.mV[j] * self.W(r), 0.0)[:, None, None] * self.sig[j])
            self.sig[soil] = reg[soil] * self.cspm[soil, None, None]
        else:
            self.flag[soil] = self.flags(self.sig[soil])
            self.sig[soil] = self.adapt(self.sig[soil])
            self.seq[soil] += self.dt * self.dseq[soil]
            self.seqp[soil] += self.dt * self.dseqp[soil]

    def rk4(self):
        """substep_RK (solver_sph_base.py:134-180): four one_step calls
        weighted 1, 2, 2, 1; each of the first three intermediate states is
        the step-start state plus half a step of the last rate (advect_RK_4
        uses 0.5 dt before every stage, the fourth included)."""
        real = self.t > 0
        soil = self.t == SOIL
        acc_rho = np.zeros(self.n)
        acc_v = np.zeros((self.n, 3))
        acc_sig = np.zeros((self.n, 3, 3))
        for k, m in enumerate((1, 2, 2, 1)):
            self.one_step()
            acc_rho[real] += self.drho[real] * m
            acc_v[real] += self.dv[real] * m
            if self.d["method"] == 3:
                acc_sig[soil] += self.dsig[soil] * m
            if k < 3:
                self.rhot[real] = 0.5 * self.dt * self.drho[real] + self.rho[real]
                self.mV[real] = self.mass[real] / self.rhot[real]
                self.vt[real] = 0.5 * self.dt * self.dv[real] + self.v[real]
                if self.d["method"] == 3:
                    self.sigt[soil] = 0.5 * self.dt * self.dsig[soil] + self.sig[soil]
        self.rho[real] += self.dt / 6.0 * acc_rho[real]
        self.mV[real] = self.mass[real] / self.rho[real]
        self.v[real] += self.dt / 6.0 * acc_v[real]
        if self.d["method"] == 3:
            self.sig[soil] += self.dt / 6.0 * acc_sig[soil]

    def enforce_boundary(self):
        """enforce_boundary_2D/3D + simulate_collisions
        (solver_sph_base.py:524-601) for the particles judge_enforce_bdy
        picks: flow particles with boundary 1, dynamic rigid ones always.
        Per axis the position is clamped to the wall minus (radius - 1e-8)
        (no top wall), the wall normals summed, and the velocity loses
        1.3 times its component along the normalised sum."""
        sel = ((self.d["boundary"] == 1) & is_flow(self.t)) | self.rig_dyn
        if not sel.any():
            return
        ds, de = np.asarray(self.d["domain_start"]), np.asarray(self.d["domain_end"])
        dist = self.d["particle_radius"] - 1e-8
        pos = self.x[sel]  # the tests read the position before any clamp
        x = pos.copy()
        nrm = np.zeros_like(x)
        axes = (0, 1, 2) if self.dim == 3 else (0, 1)
        for a in axes:
            if a != 1:  # the top wall is commented out in the reference
                hi = pos[:, a] > de[a] - dist
                nrm[hi, a] += 1.0
                x[hi, a] = de[a] - dist
            lo = pos[:, a] <= ds[a] + dist
            nrm[lo, a] -= 1.0
            x[lo, a] = ds[a] + dist
        ln = np.sqrt(np.sum(nrm * nrm, axis=1))
        hit = ln > 1e-8
        u = np.where(hit[:, None], nrm / np.where(hit, ln, 1.0)[:, None], 0.0)
        v = self.v[sel]
        v = v - (1.3 * np.sum(v * u, axis=1))[:, None] * u
        self.x[sel] = x
        self.v[sel] = v

    def step(self):
        self.shepard()
        self.init_tmp()
        ti = self.d["time_integration"]
        if ti == 1:
            self.one_step()
            self.lf_full()
        elif ti == 4:
            self.rk4()
        else:
            self.one_step()
            self.lf_half()
            self.one_step()
            self.lf_full()
        self.advect_pos()
        self.advect_something()
        self.solve_rigid()
        self.enforce_boundary()

    def solve_rigid(self):
        """solve_constraints (solver_sph_base.py:479-500): cm, A, the polar
        factor R = U V^T of A's SVD (a proper rotation), x <- cm + R q."""
        for o, cm0 in self.rigid:
            sel = (self.t == RIGID) & (self.obj == o)
            m = self.mass[sel]
            cm = (m[:, None] * self.x[sel]).sum(axis=0) / m.sum()
            q = self.x0[sel] - cm0
            p = self.x[sel] - cm
            A = np.einsum("n,na,nb->ab", self.mV[sel] * self.rho[sel], p, q)
            if self.dim == 2:
                U, _, Vt = np.linalg.svd(A[:2, :2])
                R2 = U @ Vt
                if np.linalg.det(R2) < 0:
                    U[:, 1] *= -1
                    R2 = U @ Vt
                R = np.eye(3)
                R[:2, :2] = R2
            else:
                U, _, Vt = np.linalg.svd(A)
                R = U @ Vt
                if np.linalg.det(R) < 0:
                    U[:, 2] *= -1
                    R = U @ Vt
            goal = cm + q @ R.T
            self.x[sel] = self.x[sel] + (goal - self.x[sel])
            if self.dim == 2:
                self.x[sel, 2] = 0.0

    def stress6(self, which="sig"):
        s = getattr(self, which)
        return np.stack([s[:, 0, 0], s[:, 1, 1], s[:, 2, 2], s[:, 0, 1], s[:, 1, 2], s[:, 0, 2]], axis=1)
