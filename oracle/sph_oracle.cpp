// sph_oracle.cpp — CPU oracle: a C++17/OpenMP restatement of the reference
// tiSPHi particle hot path.  TEST INFRASTRUCTURE ONLY: used by tests/, by
// __graft_entry__.smoke() as the checker and by bench.py's cpu_baseline leg.
// Nothing in the product library links or calls it.
//
// Every pass cites the reference lines it restates.  Deviations from the
// reference's literal code are the parity-mode decisions:
//   D1 WC one_step loop A in two phases (EOS for all fluid, then the dummy
//      extrapolation reads the new pressures)   solver_sph_wc.py:86-103
//   D2 muI advect_something in two phases (sigma <- sigma_tmp for all soil,
//      then regularise from those values)        solver_sph_muI.py:151-156
//   D3 XSPH from pre-advect positions            solver_sph_base.py:231-238
//   D4 stable counting sort (ascending previous index in a cell), which is
//      the serial order of the reference           particle_system.py:238-252
//   D5 stencil cells outside the grid contribute nothing
//                                                  particle_system.py:259-269
// Positions are held as an integer cell plus a cell-local offset of the
// pair-arithmetic type R, the same representation as the GPU engine.
// The Drucker-Prager stress state, return mapping and constitutive update
// run in double for both precisions (the engine's mixed mode); global
// reductions are double.
#include "oracle.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>

namespace {

enum {
  F_POS = 0, F_VEL, F_DENSITY, F_MASS, F_M_V, F_PRESSURE, F_TYPE, F_ID0,
  F_OBJ_ID, F_STRESS, F_STRESS_TMP, F_D_STRESS, F_STRAIN_EQU, F_STRAIN_EQU_P,
  F_D_STRAIN_EQU, F_D_STRAIN_EQU_P, F_FLAG_RETMAP, F_CSPM_F, F_D_DENSITY,
  F_D_VEL, F_DENSITY_TMP, F_V_TMP, F_CELL, F_OFFSET, F_CELL_START, F_CSPM_L,
  F_X0 = 28  // initial positions, indexed by id0 (generation order)
};
enum {
  P_SORT = 0, P_SHEPARD, P_INIT_TMP, P_ONE_STEP, P_LF_HALF, P_LF_FULL,
  P_ADVECT_POS, P_ADVECT_SOMETHING, P_STEP, P_SE, P_ENFORCE_BOUNDARY, P_SOLVE_RIGID
};
constexpr int FLUID = 1, SOIL = 2, RIGID = 11, DUMMY = -1, REP = -2;
constexpr int WC = 1, MUI = 2, DP = 3;

inline bool is_flow(int t) { return t == FLUID || t == SOIL; }
inline bool is_real(int t) { return t > 0; }
inline bool is_bdy(int t) { return t == DUMMY || t == REP; }

// Orthogonal factor of the polar decomposition A = R S (ti.polar_decompose
// in solve_constraints, solver_sph_base.py:490).  2D (the xy block, R_zz =
// 1): the nearest rotation, (a + d, c - b) normalised.  3D: Newton's
// iteration X <- (X + X^-T) / 2 (Higham 1986) from X = A.  A singular A
// gives I.
inline void polar_rotation(const double A[9], int dim, double R[9]) {
  for (int c = 0; c < 9; ++c) R[c] = (c % 4 == 0) ? 1.0 : 0.0;
  if (dim == 2) {
    const double s = A[0] + A[4], t = A[3] - A[1];
    const double nrm = std::sqrt(s * s + t * t);
    if (nrm == 0.0) return;
    R[0] = s / nrm;
    R[1] = -t / nrm;
    R[3] = t / nrm;
    R[4] = s / nrm;
    return;
  }
  double X[9];
  for (int c = 0; c < 9; ++c) X[c] = A[c];
  for (int it = 0; it < 100; ++it) {
    double cf[9];  // cofactors: X^-T = cof(X) / det(X)
    cf[0] = X[4] * X[8] - X[5] * X[7];
    cf[1] = X[5] * X[6] - X[3] * X[8];
    cf[2] = X[3] * X[7] - X[4] * X[6];
    cf[3] = X[2] * X[7] - X[1] * X[8];
    cf[4] = X[0] * X[8] - X[2] * X[6];
    cf[5] = X[1] * X[6] - X[0] * X[7];
    cf[6] = X[1] * X[5] - X[2] * X[4];
    cf[7] = X[2] * X[3] - X[0] * X[5];
    cf[8] = X[0] * X[4] - X[1] * X[3];
    const double det = X[0] * cf[0] + X[1] * cf[1] + X[2] * cf[2];
    if (det == 0.0) return;
    double d = 0.0;
    for (int c = 0; c < 9; ++c) {
      const double xn = 0.5 * (X[c] + cf[c] / det);
      const double e = std::fabs(xn - X[c]);
      d = e > d ? e : d;
      X[c] = xn;
    }
    if (d <= 1e-15) break;
  }
  for (int c = 0; c < 9; ++c) R[c] = X[c];
}

// Stress component order xx,yy,zz,xy,yz,zx <-> 3x3 index.
inline int sidx(int a, int b) {
  static const int m[3][3] = {{0, 3, 5}, {3, 1, 4}, {5, 4, 2}};
  return m[a][b];
}

template <class T>
T ipow_or_pow(T x, double e) {
  double ie = std::nearbyint(e);
  if (ie == e && std::fabs(e) <= 64) {
    long n = (long)std::fabs(ie);
    T r = T(1), b = x;
    while (n) {
      if (n & 1) r = r * b;
      b = b * b;
      n >>= 1;
    }
    return ie < 0 ? T(1) / r : r;
  }
  return (T)std::pow(x, (T)e);
}

template <class R>
struct Sim {
  orc_params P{};
  int dim = 3;
  int64_t n = 0;
  int gn[3] = {1, 1, 1};
  int64_t C = 0;
  double grid = 0;
  R g = 0, rs2 = 0, h1 = 0, kW = 0, kC = 0, kG = 0, kGc = 0, eps = (R)1e-8;
  R dt = 0, half_dt = 0, rho0 = 0;
  R visc_c1 = 0, visc_c2 = 0, grav[3] = {0, 0, 0};
  // repulsive-particle force (boundary 3 / 4; WC and muI)
  bool rep = false;
  R rep_k = 0, rep_d = 0, rep_h = 0;
  // DP constants (double)
  double alpha = 0, kc = 0, G = 0, K = 0, sin_psi = 0, eps_f = 1e-4;
  int64_t escaped = 0;

  std::vector<int32_t> cx, cy, cz, type, id0, obj, flag, key;
  std::vector<R> ox, oy, oz, vx, vy, vz, rho, mass, mV, pr, rhot, vtx, vty,
      vtz, cspm, drho, dvx, dvy, dvz;
  std::vector<double> sig[6], sigt[6], dsig[6], seq, seqp, dseq, dseqp;
  std::vector<R> drhoRK, dvRKx, dvRKy, dvRKz;
  std::vector<double> dsigRK[6];
  std::vector<R> xsx, xsy, xsz;
  std::vector<R> Lc[9];  // CSPM_L (row-major 3x3), kernelCorrection 1
  // dynamic rigid objects: ids, rest centres of mass; initial positions by id0
  int nrig = 0, rig_obj[4] = {0, 0, 0, 0};
  double rig_cm0[4][3] = {};
  std::vector<double> x0[3];
  std::vector<int64_t> start;  // size C+1, exclusive scan of cell counts

  bool f32_state() const { return sizeof(R) == 4; }
  // muI stress lives in R; DP stress lives in double.
  double rnd_mui(double x) const { return P.method == MUI ? (double)(R)x : x; }

  void init(const orc_params* p) {
    P = *p;
    dim = P.dim;
    n = P.n_particles;
    for (int a = 0; a < 3; ++a) gn[a] = P.grid_num[a];
    C = (int64_t)gn[0] * gn[1] * gn[2];
    grid = P.grid_size;
    g = (R)grid;
    R rs = (R)P.support_radius;
    rs2 = rs * rs;
    double h = P.smoothing_len, h1d = 1.0 / h;
    h1 = (R)h1d;
    // Wendland C2 and cubic spline normalisations (solver_sph_base.py:302,338)
    double kw = dim == 2 ? 7.0 / (4.0 * M_PI) : 21.0 / (2.0 * M_PI);
    kW = (R)(kw * std::pow(h1d, dim));
    double kcub = dim == 2 ? 15.0 / 7.0 / M_PI : 3.0 / 2.0 / M_PI;
    kC = (R)(kcub * std::pow(h1d, dim));
    // W'(r)/r factors (q h1 / r = h1^2): -5 kW h1^2 and kC h1^2
    kG = (R)(-5.0 * kw * std::pow(h1d, dim) * h1d * h1d);
    kGc = (R)(kcub * std::pow(h1d, dim) * h1d * h1d);
    dt = (R)P.dt;
    half_dt = (R)(0.5 * P.dt);
    rho0 = (R)P.density0;
    visc_c1 = (R)(2.0 * (dim + 2) * P.viscosity);
    visc_c2 = (R)(0.01 * h * h);
    for (int a = 0; a < 3; ++a) grav[a] = (R)P.gravity[a];
    rep = (P.boundary == 3 || P.boundary == 4) && P.method != 3;  // DP leaves it out
    rep_k = (R)(0.01 * P.vsound * P.vsound);
    rep_d = (R)P.particle_diameter;
    rep_h = (R)(0.75 * P.smoothing_len);
    double t = std::tan(P.friction);
    alpha = t / std::sqrt(9 + 12 * t * t);
    kc = 3 * P.cohesion / std::sqrt(9 + 12 * t * t);
    G = P.young / (2 * (1 + P.poisson));
    K = P.young / (3 * (1 - 2 * P.poisson));
    sin_psi = std::sin(P.dilatancy);
    auto zi = [&](std::vector<int32_t>& v) { v.assign(n, 0); };
    auto zr = [&](std::vector<R>& v) { v.assign(n, R(0)); };
    auto zd = [&](std::vector<double>& v) { v.assign(n, 0.0); };
    zi(cx); zi(cy); zi(cz); zi(type); zi(id0); zi(obj); zi(flag); zi(key);
    for (auto* v : {&ox, &oy, &oz, &vx, &vy, &vz, &rho, &mass, &mV, &pr, &rhot,
                    &vtx, &vty, &vtz, &cspm, &drho, &dvx, &dvy, &dvz, &drhoRK,
                    &dvRKx, &dvRKy, &dvRKz, &xsx, &xsy, &xsz})
      zr(*v);
    for (int c = 0; c < 6; ++c) { zd(sig[c]); zd(sigt[c]); zd(dsig[c]); zd(dsigRK[c]); }
    zd(seq); zd(seqp); zd(dseq); zd(dseqp);
    for (int c = 0; c < 9; ++c) Lc[c].assign(n, (c % 4 == 0) ? R(1) : R(0));
    for (int c = 0; c < 3; ++c) x0[c].assign(n, 0.0);
    start.assign(C + 1, 0);
  }

  // ---------------------------------------------------------------- kernels
  // Wendland C2 (solver_sph_base.py:334-358), cubic spline (:298-330).
  R W(R rn) const {
    R q = rn * h1;
    if (P.kernel == 1) {
      if (rn > eps && q <= R(2)) {
        R q1 = R(1) - R(0.5) * q;
        R q2 = q1 * q1;
        return kW * (q2 * q2) * (R(1) + R(2) * q);
      }
      return R(0);
    }
    if (rn > eps && q <= R(2)) {
      if (q <= R(1)) {
        R qq = q * q;
        return kC * (R(0.5) * (qq * q) - qq + R(2.0 / 3.0));
      }
      R t = R(2) - q;
      return kC / R(6) * (t * t * t);
    }
    return R(0);
  }
  // gradient = fac * x_ij, fac = W'(r)/r (solver_sph_base.py:316-330,
  // 348-358: k q1^3 (-5 q) h1 r/|r| and k q (1.5 q - 2) h1 r/|r|, with
  // q h1 / |r| = h1^2 folded into the constants kG, kGc).
  R gradW_fac(R rn) const {
    R q = rn * h1;
    if (!(rn > eps && q <= R(2))) return R(0);
    if (P.kernel == 1) {
      R q1 = R(1) - R(0.5) * q;
      return kG * ((q1 * q1) * q1);
    }
    if (q <= R(1)) return kGc * (R(1.5) * q - R(2));
    R t = R(2) - q;
    return kGc * (R(-0.5) * t * t) / q;
  }

  // --------------------------------------------------------- neighbour walk
  // for_all_neighbors (particle_system.py:259-269): 3^dim stencil in
  // ndrange((-1,2),...) order, j ascending inside each cell, i != j and
  // |x_ij| < r_s.  (scx,scy,scz) is the stencil centre cell (floor of the
  // current position); the pair vector uses each particle's stored cell.
  template <class Fn>
  void walk(int64_t i, int scx, int scy, int scz, Fn&& fn) const {
    for (int dx = -1; dx <= 1; ++dx) {
      int nx = scx + dx;
      if (nx < 0 || nx >= gn[0]) continue;
      for (int dy = -1; dy <= 1; ++dy) {
        int ny = scy + dy;
        if (ny < 0 || ny >= gn[1]) continue;
        for (int dz = -1; dz <= 1; ++dz) {
          if (dim == 2 && dz != 0) continue;
          int nz = scz + dz;
          if (nz < 0 || nz >= gn[2]) continue;
          int64_t k = ((int64_t)nx * gn[1] + ny) * gn[2] + nz;
          R bx = ox[i] + (R)(cx[i] - nx) * g;
          R by = oy[i] + (R)(cy[i] - ny) * g;
          R bz = oz[i] + (R)(cz[i] - nz) * g;
          for (int64_t j = start[k]; j < start[k + 1]; ++j) {
            if (j == i) continue;
            R x = bx - ox[j], y = by - oy[j], z = bz - oz[j];
            R r2 = std::fma(z, z, std::fma(y, y, x * x));  // same roundings as the engine
            if (r2 < rs2) fn(j, x, y, z, r2);
          }
        }
      }
    }
  }
  template <class Fn>
  void walk(int64_t i, Fn&& fn) const { walk(i, cx[i], cy[i], cz[i], fn); }

  // -------------------------------------------------------------- the grid
  // Cell of an offset along one axis: floor((double)o / grid).
  void renorm(int32_t& c, R& o) const {
    double od = (double)o;
    if (od != od) return;
    double k = std::floor(od / grid);
    double r = od - k * grid;
    if (r >= grid) { k += 1; r -= grid; }
    R on = (R)r;
    if (on < R(0)) on = R(0);
    if ((double)on >= grid) { k += 1; on = R(0); }
    c += (int32_t)k;
    o = on;
  }
  void clamp_cell(int32_t& c, R& o, int a) {
    if (c < 0 || c >= gn[a]) {
      int32_t cc = c < 0 ? 0 : gn[a] - 1;
      o = (R)((double)o + (double)(c - cc) * grid);
      c = cc;
      ++escaped;
    }
  }
  int64_t flat(int x, int y, int z) const { return ((int64_t)x * gn[1] + y) * gn[2] + z; }

  // update_grid_id + prefix sum + counting_sort (particle_system.py:228-257),
  // with a stable slot order (D4).
  void sort() {
    for (int64_t i = 0; i < n; ++i) {
      renorm(cx[i], ox[i]);
      renorm(cy[i], oy[i]);
      clamp_cell(cx[i], ox[i], 0);
      clamp_cell(cy[i], oy[i], 1);
      if (dim == 3) {
        renorm(cz[i], oz[i]);
        clamp_cell(cz[i], oz[i], 2);
      } else {
        cz[i] = 0;
        oz[i] = R(0);
      }
      key[i] = (int32_t)flat(cx[i], cy[i], cz[i]);
    }
    std::vector<int64_t> cnt(C + 1, 0);
    for (int64_t i = 0; i < n; ++i) cnt[key[i]]++;
    start[0] = 0;
    for (int64_t c = 0; c < C; ++c) start[c + 1] = start[c] + cnt[c];
    std::vector<int64_t> cur(start.begin(), start.end() - 1), perm(n);
    for (int64_t i = 0; i < n; ++i) perm[cur[key[i]]++] = i;
    permute(perm);
  }
  template <class T>
  static void apply(std::vector<T>& v, const std::vector<int64_t>& perm) {
    std::vector<T> t(v.size());
    for (size_t k = 0; k < perm.size(); ++k) t[k] = v[perm[k]];
    v.swap(t);
  }
  void permute(const std::vector<int64_t>& perm) {
    for (auto* v : {&cx, &cy, &cz, &type, &id0, &obj, &flag, &key}) apply(*v, perm);
    for (auto* v : {&ox, &oy, &oz, &vx, &vy, &vz, &rho, &mass, &mV, &pr, &rhot,
                    &vtx, &vty, &vtz, &cspm, &drho, &dvx, &dvy, &dvz, &drhoRK,
                    &dvRKx, &dvRKy, &dvRKz})
      apply(*v, perm);
    for (int c = 0; c < 6; ++c) { apply(sig[c], perm); apply(sigt[c], perm); apply(dsig[c], perm); apply(dsigRK[c], perm); }
    for (auto* v : {&seq, &seqp, &dseq, &dseqp}) apply(*v, perm);
  }

  // ------------------------------------------------ kernel correction
  // calc_CSPM_L (solver_sph_base.py:400-423): for flow particles
  // L = sum_{j same type} m_V_j (x_j - x_i) (x) grad W_ij, stored inverted
  // when |det| > eps (2D: the upper 2x2 block, the rest zero), else I;
  // other particles keep I.
  void calc_cspm_L() {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
      R M[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
      if (is_flow(type[i])) {
        R L[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        const int ti = type[i];
        walk(i, [&](int64_t j, R x, R y, R z, R r2) {
          if (type[j] != ti) return;
          R f = gradW_fac(std::sqrt(r2));
          R g[3] = {f * x, f * y, f * z};
          R c[3] = {mV[j] * (-x), mV[j] * (-y), mV[j] * (-z)};
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) L[a][b] += c[a] * g[b];
        });
        invert_corr(L, M);
      }
      for (int c = 0; c < 9; ++c) Lc[c][i] = M[c];
    }
  }
  // The inverse of calc_CSPM_L: adjugate times 1/det (row-major out).
  void invert_corr(const R L[3][3], R M[9]) const {
    if (dim == 2) {
      R det = L[0][0] * L[1][1] - L[0][1] * L[1][0];
      if (std::abs(det) > eps) {
        R id = R(1) / det;
        R o[9] = {L[1][1] * id, -L[0][1] * id, 0, -L[1][0] * id, L[0][0] * id, 0, 0, 0, 0};
        for (int c = 0; c < 9; ++c) M[c] = o[c];
      }
      return;
    }
    R c00 = L[1][1] * L[2][2] - L[1][2] * L[2][1];
    R c01 = L[1][2] * L[2][0] - L[1][0] * L[2][2];
    R c02 = L[1][0] * L[2][1] - L[1][1] * L[2][0];
    R det = L[0][0] * c00 + L[0][1] * c01 + L[0][2] * c02;
    if (std::abs(det) > eps) {
      R id = R(1) / det;
      R o[9] = {c00 * id, (L[0][2] * L[2][1] - L[0][1] * L[2][2]) * id, (L[0][1] * L[1][2] - L[0][2] * L[1][1]) * id,
                c01 * id, (L[0][0] * L[2][2] - L[0][2] * L[2][0]) * id, (L[0][2] * L[1][0] - L[0][0] * L[1][2]) * id,
                c02 * id, (L[0][1] * L[2][0] - L[0][0] * L[2][1]) * id, (L[0][0] * L[1][1] - L[0][1] * L[1][0]) * id};
      for (int c = 0; c < 9; ++c) M[c] = o[c];
    }
  }
  // kernel_deriv_corr (solver_sph_base.py:375-384): CSPM_L_i grad W with
  // kernelCorrection 1, the zero vector with 2 (its branch body is `pass`,
  // so tmp keeps its initial 0), grad W unchanged with 0.
  void corr(int64_t i, R g[3]) const {
    if (P.kernel_corr == 0) return;
    if (P.kernel_corr == 2) {
      g[0] = g[1] = g[2] = R(0);
      return;
    }
    R o[3];
    for (int a = 0; a < 3; ++a) o[a] = (Lc[3 * a][i] * g[0] + Lc[3 * a + 1][i] * g[1]) + Lc[3 * a + 2][i] * g[2];
    for (int a = 0; a < 3; ++a) g[a] = o[a];
  }

  // ------------------------------------------------------------ Shepard
  // calc_kernel_corr (solver_sph_base.py:363-368): calc_CSPM_f, then
  // calc_CSPM_L with kernelCorrection 1; with 2 nothing more (CSPM_L keeps
  // the zero it was allocated with).
  void shepard() {
    calc_cspm_f();
    if (P.kernel_corr == 1) calc_cspm_L();
    if (P.kernel_corr == 2)
      for (int c = 0; c < 9; ++c) std::fill(Lc[c].begin(), Lc[c].end(), R(0));
  }
  // calc_CSPM_f (solver_sph_base.py:386-398): 1 / sum_flow m_V_j W_ij.
  void calc_cspm_f() {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
      R s = R(0);
      walk(i, [&](int64_t j, R x, R y, R z, R r2) {
        if (is_flow(type[j])) s += mV[j] * W(std::sqrt(r2));
      });
      cspm[i] = s != R(0) ? R(1) / s : R(1);
    }
  }

  // init_real2tmp (solver_sph_base.py:67-74)
  void init_tmp() {
#pragma omp parallel for
    for (int64_t i = 0; i < n; ++i) {
      if (is_real(type[i])) {
        rhot[i] = rho[i];
        vtx[i] = vx[i]; vty[i] = vy[i]; vtz[i] = vz[i];
      }
      if (type[i] == SOIL)
        for (int c = 0; c < 6; ++c) sigt[c][i] = sig[c][i];
    }
  }

  // calc_repulsive_force (solver_sph_base.py:676-690): F = fac * x_ij.
  R rep_fac(R rn) const {
    R chi = (rn > R(0) && rn < rep_d) ? R(1) - rn / rep_d : R(0);
    R gam = rn / rep_h;
    const R two3 = (R)(2.0 / 3.0);
    R f = 0;
    if (gam > R(0) && gam <= two3) f = two3;
    else if (gam > two3 && gam <= R(1)) f = R(2) * gam - R(1.5) * (gam * gam);
    else if (gam > R(1) && gam < R(2)) {
      R t = R(2) - gam;
      f = R(0.5) * (t * t);
    }
    return rep_k * chi * f / (rn * rn);
  }
  // calc_rep_force_task (solver_sph_wc.py:74-76, solver_sph_muI.py:34-36):
  // a second sweep after the pair forces, repulsive neighbours only.
  void add_rep_force(int64_t i, R dv[3]) const {
    walk(i, [&](int64_t j, R x, R y, R z, R r2) {
      if (type[j] != REP) return;
      R f = rep_fac(std::sqrt(r2));
      dv[0] += f * x;
      dv[1] += f * y;
      dv[2] += f * z;
    });
  }

  // --------------------------------------------------------------- WCSPH
  R eos(R rt) const {  // solver_sph_wc.py:88
    R v = (R)P.stiffness * (ipow_or_pow(rt / rho0, P.exponent) - R(1));
    return v > R(0) ? v : R(0);
  }
  // Dummy extrapolation task pair (solver_sph_base.py:653-663) for WC.
  void wc_bdy(int64_t i) {
    R sv[3] = {0, 0, 0}, sp = 0;
    walk(i, [&](int64_t j, R x, R y, R z, R r2) {
      if (!is_flow(type[j])) return;
      R w = W(std::sqrt(r2));
      sv[0] += mV[j] * vtx[j] * w;
      sv[1] += mV[j] * vty[j] * w;
      sv[2] += mV[j] * vtz[j] * w;
      sp += mV[j] * (pr[j] + rhot[j] * grav[1] * y) * w;
    });
    vtx[i] = R(2) * vx[i] - sv[0] * cspm[i];
    vty[i] = R(2) * vy[i] - sv[1] * cspm[i];
    vtz[i] = R(2) * vz[i] - sv[2] * cspm[i];
    rhot[i] = rho0;
    R pv = sp * cspm[i];
    pr[i] = pv > R(0) ? pv : R(0);
  }
  // Loop B: d_density and d_vel of fluid i (solver_sph_wc.py:108-123;
  // tasks solver_sph_base.py:195-203, solver_sph_wc.py:33-71).
  void wc_fluid(int64_t i) {
    R dd = 0, dv[3] = {0, 0, 0};
    const R pri = pr[i] / (rhot[i] * rhot[i]);
    const R rr = rho0 / rhot[i];
    walk(i, [&](int64_t j, R x, R y, R z, R r2) {
      R rn = std::sqrt(r2);
      R f = gradW_fac(rn);
      R gx = f * x, gy = f * y, gz = f * z;
      R ux = vtx[i] - vtx[j], uy = vty[i] - vty[j], uz = vtz[i] - vtz[j];
      R cg[3] = {gx, gy, gz};  // d_density takes the corrected gradient
      corr(i, cg);
      dd += (mV[j] * ux) * cg[0] + (mV[j] * uy) * cg[1] + (mV[j] * uz) * cg[2];
      R vis = 0;
      int tj = type[j];
      if (tj == FLUID || is_bdy(tj) || tj == RIGID) {
        R vx_ = ux * x + uy * y + uz * z;
        R mn = vx_ < R(0) ? vx_ : R(0);
        R a = visc_c1 * mV[j];
        if (tj != FLUID) a = a * rr;
        vis = a * mn / (r2 + visc_c2);
      }
      R pc = (-rho0 * mV[j]) * (pri + pr[j] / (rhot[j] * rhot[j]));
      R sc = vis + pc;  // viscosity + pressure force, both along grad W
      dv[0] += sc * gx;
      dv[1] += sc * gy;
      dv[2] += sc * gz;
    });
    if (rep) add_rep_force(i, dv);
    drho[i] = dd * rhot[i];
    dvx[i] = dv[0] + grav[0];
    dvy[i] = dv[1] + grav[1];
    dvz[i] = dv[2] + grav[2];
  }
  void wc_one_step() {  // solver_sph_wc.py:82-127 (D1)
#pragma omp parallel for
    for (int64_t i = 0; i < n; ++i)
      if (type[i] == FLUID) pr[i] = eos(rhot[i]);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i)
      if (is_bdy(type[i]) || type[i] == RIGID) wc_bdy(i);  // :90-103
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
      if (type[i] == FLUID) wc_fluid(i);
      else if (type[i] == RIGID) wc_rigid(i);
    }
  }
  // Rigid particle i (solver_sph_wc.py:105-106, 68-71, 125-126): static ->
  // d_vel 0; dynamic -> g minus the calc_d_vel_task term of every fluid
  // neighbour j, formed from j's side (x_ji = x_j - x_i, j's density_tmp in
  // the wall viscosity).  The reference subtracts these in loop B with
  // atomic updates in no fixed order; here i gathers them in its own
  // neighbour order.
  int add_rigid(int o, const double cm[3]) {  // calc_rigid_rest_m, :517-518
    int k = 0;
    while (k < nrig && rig_obj[k] != o) ++k;
    if (k == 4) return -1;
    rig_obj[k] = o;
    for (int a = 0; a < 3; ++a) rig_cm0[k][a] = cm[a];
    if (k == nrig) ++nrig;
    return 0;
  }
  int rigid_slot(int64_t i) const {
    if (type[i] != RIGID) return -1;
    for (int k = 0; k < nrig; ++k)
      if (rig_obj[k] == obj[i]) return k;
    return -1;
  }
  void wc_rigid(int64_t i) {
    drho[i] = 0;
    if (rigid_slot(i) < 0) {
      dvx[i] = dvy[i] = dvz[i] = R(0);
      return;
    }
    R dv[3] = {grav[0], grav[1], grav[2]};
    const R mVi = mV[i];
    const R pri = pr[i] / (rhot[i] * rhot[i]);
    walk(i, [&](int64_t j, R x, R y, R z, R r2) {
      if (type[j] != FLUID) return;
      const R xj = -x, yj = -y, zj = -z;
      R rn = std::sqrt(r2);
      R f = gradW_fac(rn);
      R gx = f * xj, gy = f * yj, gz = f * zj;
      R ux = vtx[j] - vtx[i], uy = vty[j] - vty[i], uz = vtz[j] - vtz[i];
      R vx_ = ux * xj + uy * yj + uz * zj;
      R mn = vx_ < R(0) ? vx_ : R(0);
      R rr = rho0 / rhot[j];
      R vis = ((visc_c1 * mVi) * rr) * mn / (r2 + visc_c2);
      R pc = (-rho0 * mVi) * (pr[j] / (rhot[j] * rhot[j]) + pri);
      R sc = vis + pc;
      dv[0] -= sc * gx;
      dv[1] -= sc * gy;
      dv[2] -= sc * gz;
    });
    dvx[i] = dv[0];
    dvy[i] = dv[1];
    dvz[i] = dv[2];
  }
  // solve_rigid_body / solve_constraints (solver_sph_base.py:472-510), in
  // FP64: cm = sum m x / sum m; A = sum m_V rho (x - cm) (x0 - cm0)^T;
  // R = polar rotation of A; x += cm + R (x0 - cm0) - x.
  void solve_rigid() {
    for (int k = 0; k < nrig; ++k) {
      double m = 0, mx[3] = {0, 0, 0};
      for (int64_t i = 0; i < n; ++i) {
        if (rigid_slot(i) != k) continue;
        m += (double)mass[i];
        for (int a = 0; a < 3; ++a) mx[a] += (double)mass[i] * abs_or0(a, i);
      }
      double cm[3] = {mx[0] / m, mx[1] / m, mx[2] / m}, A[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      for (int64_t i = 0; i < n; ++i) {
        if (rigid_slot(i) != k) continue;
        const double w = (double)mV[i] * (double)rho[i];
        double q[3], p[3];
        for (int a = 0; a < 3; ++a) {
          q[a] = x0[a][id0[i]] - rig_cm0[k][a];
          p[a] = abs_or0(a, i) - cm[a];
        }
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) A[3 * a + b] += w * p[a] * q[b];
      }
      double Rm[9];
      polar_rotation(A, dim, Rm);
      for (int64_t i = 0; i < n; ++i) {
        if (rigid_slot(i) != k) continue;
        double q[3];
        for (int a = 0; a < 3; ++a) q[a] = x0[a][id0[i]] - rig_cm0[k][a];
        for (int a = 0; a < dim; ++a) {
          const double goal = cm[a] + (Rm[3 * a] * q[0] + Rm[3 * a + 1] * q[1] + Rm[3 * a + 2] * q[2]);
          const double x = abs_pos(a, i);
          set_abs(a, i, x + (goal - x) * 1.0);
        }
      }
    }
  }

  // ------------------------------------------------------------------ muI
  // Soil loop 1 (solver_sph_muI.py:200-225): velocity gradient, d_density,
  // pressure, mu(I) stress, equivalent strain rate.
  void mui_soil1(int64_t i) {
    R L[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, dd = 0;
    walk(i, [&](int64_t j, R x, R y, R z, R r2) {
      R f = gradW_fac(std::sqrt(r2));
      R gw[3] = {f * x, f * y, f * z};
      corr(i, gw);
      R u[3] = {mV[j] * (vtx[j] - vtx[i]), mV[j] * (vty[j] - vty[i]), mV[j] * (vtz[j] - vtz[i])};
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) L[a][b] += u[a] * gw[b];
      dd += (mV[j] * (vtx[i] - vtx[j])) * gw[0] + (mV[j] * (vty[i] - vty[j])) * gw[1] +
            (mV[j] * (vtz[i] - vtz[j])) * gw[2];
    });
    drho[i] = dd * rhot[i];
    R p = (R)(P.vsound * P.vsound) * (rho[i] - rho0);
    p = p > R(0) ? p : R(0);
    pr[i] = p;
    R e[3][3];
    R ss = 0;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        e[a][b] = R(0.5) * (L[a][b] + L[b][a]);
      }
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) ss += e[a][b] * e[a][b];
    R en = std::sqrt(R(0.5) * ss) + eps;
    R fac = R(0) + ((R)P.cohesion + p * (R)std::tan(P.friction)) / en;
    for (int a = 0; a < 3; ++a)
      for (int b = a; b < 3; ++b) {
        R v = fac * e[a][b] - (a == b ? p : R(0));
        sigt[sidx(a, b)][i] = (double)v;
      }
    R tr3 = (e[0][0] + e[1][1] + e[2][2]) / R(3);
    R sd = 0;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        R d = e[a][b] - (a == b ? tr3 : R(0));
        sd += d * d;
      }
    dseq[i] = (double)(R)std::sqrt(sd * R(2) / R(3));
  }
  // Dummy loop (muI :228-242, DP :220-231; tasks solver_sph_base.py:647-669).
  void soil_bdy(int64_t i, bool extrapolate_density) {
    R sv[3] = {0, 0, 0}, sr = 0, ss[6] = {0, 0, 0, 0, 0, 0};
    walk(i, [&](int64_t j, R x, R y, R z, R r2) {
      if (!is_flow(type[j])) return;
      R w = W(std::sqrt(r2));
      sv[0] += mV[j] * vtx[j] * w;
      sv[1] += mV[j] * vty[j] * w;
      sv[2] += mV[j] * vtz[j] * w;
      sr += mV[j] * rhot[j] * w;
      R xr[3] = {x, y, z};
      for (int a = 0; a < 3; ++a) {
        R m = (R)sigt[a][j] + (rhot[j] * grav[a]) * xr[a];
        ss[a] += mV[j] * m * w;
      }
      for (int c = 3; c < 6; ++c) ss[c] += mV[j] * (R)sigt[c][j] * w;
    });
    vtx[i] = R(2) * vx[i] - sv[0] * cspm[i];
    vty[i] = R(2) * vy[i] - sv[1] * cspm[i];
    vtz[i] = R(2) * vz[i] - sv[2] * cspm[i];
    if (extrapolate_density) {
      R r = sr * cspm[i];
      rhot[i] = r > rho0 ? r : rho0;
    } else {
      rhot[i] = rho0;
    }
    for (int c = 0; c < 6; ++c) sigt[c][i] = rnd_mui((double)(ss[c] * cspm[i]));
  }
  // Soil loop 3 (muI :248-261; DP :260-270): acceleration from stress plus
  // viscous damping (solver_sph_base.py:713-715).
  void soil_dvel_sum(int64_t i, R dv[3]) const {
    dv[0] = dv[1] = dv[2] = R(0);
    R ri2 = rhot[i] * rhot[i];
    R si[6];
    for (int c = 0; c < 6; ++c) si[c] = (R)sigt[c][i] / ri2;
    walk(i, [&](int64_t j, R x, R y, R z, R r2) {
      R f = gradW_fac(std::sqrt(r2));
      R gw[3] = {f * x, f * y, f * z};
      corr(i, gw);
      R rj2 = rhot[j] * rhot[j];
      R s = mV[j] * rhot[j];
      for (int a = 0; a < 3; ++a) {
        R acc = 0;
        for (int b = 0; b < 3; ++b) {
          int c = sidx(a, b);
          acc += (s * ((R)sigt[c][j] / rj2 + si[c])) * gw[b];
        }
        dv[a] += acc;
      }
    });
  }
  void soil_dvel_finish(int64_t i, const R dv[3]) {
    R h2 = (R)(P.smoothing_len * P.smoothing_len);
    R damp = R(-5e-5) * std::sqrt((R)P.young / (rhot[i] * h2));
    dvx[i] = dv[0] + grav[0] + damp * vtx[i];
    dvy[i] = dv[1] + grav[1] + damp * vty[i];
    dvz[i] = dv[2] + grav[2] + damp * vtz[i];
  }
  void mui_one_step() {  // solver_sph_muI.py:195-265
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i)
      if (type[i] == SOIL) mui_soil1(i);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i)
      if (is_bdy(type[i]) || type[i] == RIGID) soil_bdy(i, true);  // :96-110
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i)
      if (type[i] == SOIL) {
        R dv[3];
        soil_dvel_sum(i, dv);
        if (rep) add_rep_force(i, dv);
        soil_dvel_finish(i, dv);
      } else if (type[i] == RIGID) {
        soil_rigid(i);
      }
  }
  // Rigid particle i of a soil scene (solver_sph_muI.py:38-46, 111-112,
  // 130-132; solver_sph_dp.py:158-165, 233-234, 272-274): static -> d_vel
  // 0; dynamic -> g minus the calc_d_vel_from_stress_task term of every soil
  // neighbour j, formed from j's side (x_ji, j's corrected gradient), in
  // i's neighbour order (the reference subtracts them with atomics).
  void soil_rigid(int64_t i) {
    if (rigid_slot(i) < 0) {
      dvx[i] = dvy[i] = dvz[i] = R(0);
      return;
    }
    R dv[3] = {grav[0], grav[1], grav[2]};
    const R ri2 = rhot[i] * rhot[i];
    const R s = mV[i] * rhot[i];
    walk(i, [&](int64_t j, R x, R y, R z, R r2) {
      if (type[j] != SOIL) return;
      R f = gradW_fac(std::sqrt(r2));
      R gw[3] = {f * -x, f * -y, f * -z};
      corr(j, gw);
      R rj2 = rhot[j] * rhot[j];
      for (int a = 0; a < 3; ++a) {
        R acc = 0;
        for (int b = 0; b < 3; ++b) {
          int c = sidx(a, b);
          acc += (s * ((R)sigt[c][i] / ri2 + (R)sigt[c][j] / rj2)) * gw[b];
        }
        dv[a] -= acc;
      }
    });
    dvx[i] = dv[0];
    dvy[i] = dv[1];
    dvz[i] = dv[2];
  }
  // advect_something_func (muI :267-289) with D2.
  void mui_advect_something() {
#pragma omp parallel for
    for (int64_t i = 0; i < n; ++i) {
      if (type[i] != SOIL) continue;
      rho[i] = rho[i] > rho0 ? rho[i] : rho0;
      mV[i] = mass[i] / rho[i];
      seq[i] = (double)((R)seq[i] + dt * (R)dseq[i]);
      for (int c = 0; c < 6; ++c) sig[c][i] = sigt[c][i];
    }
    std::vector<double> out[6];
    for (int c = 0; c < 6; ++c) out[c] = sig[c];
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
      if (type[i] != SOIL) continue;
      int32_t sx = cx[i], sy = cy[i], sz = cz[i];
      R tx = ox[i], ty = oy[i], tz = oz[i];
      renorm(sx, tx);
      renorm(sy, ty);
      if (dim == 3) renorm(sz, tz);
      R s[6] = {0, 0, 0, 0, 0, 0};
      walk(i, sx, sy, sz, [&](int64_t j, R x, R y, R z, R r2) {
        if (type[j] != type[i]) return;
        R w = W(std::sqrt(r2));
        for (int c = 0; c < 6; ++c) s[c] += mV[j] * (R)sig[c][j] * w;
      });
      for (int c = 0; c < 6; ++c) out[c][i] = (double)(s[c] * cspm[i]);
    }
    for (int c = 0; c < 6; ++c) sig[c].swap(out[c]);
  }

  // ------------------------------------------------------------------- DP
  // Invariants (solver_sph_dp.py:41-67).
  struct Inv { double s[6], I1, sJ2, f; };
  Inv invariants(const double st[6]) const {
    Inv v;
    v.I1 = st[0] + st[1] + st[2];
    double m = v.I1 / 3.0;
    for (int c = 0; c < 6; ++c) v.s[c] = st[c] - (c < 3 ? m : 0.0);
    double ss = 0;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) ss += v.s[sidx(a, b)] * v.s[sidx(a, b)];
    v.sJ2 = std::sqrt(0.5 * ss);
    v.f = v.sJ2 + alpha * v.I1 - kc;
    return v;
  }
  // adapt_stress / adapt_1 / adapt_2 (solver_sph_dp.py:73-96).
  void adapt(double st[6]) const {
    Inv v = invariants(st);
    if (v.f > eps_f) {
      if (v.f > v.sJ2) {
        double t = (v.I1 - kc / alpha) / 3.0;
        for (int c = 0; c < 3; ++c) st[c] -= t;
      }
      v = invariants(st);
      double r = (-v.I1 * alpha + kc) / v.sJ2;
      for (int c = 0; c < 6; ++c) st[c] = r * v.s[c] + (c < 3 ? v.I1 / 3.0 : 0.0);
    }
  }
  int upd_flag(const double st[6]) const {  // solver_sph_dp.py:98-109
    Inv v = invariants(st);
    if (v.f < -eps_f) return 0;
    if (v.f > eps_f) return v.f >= v.sJ2 ? 3 : 2;
    return 1;
  }
  // calc_d_stress_Bui2008 (solver_sph_dp.py:171-208), double precision.
  void bui2008(int64_t i, const double L[3][3]) {
    double st[6];
    for (int c = 0; c < 6; ++c) st[c] = sigt[c][i];
    Inv v = invariants(st);
    double e[3][3], w[3][3], S[3][3], s3[3][3];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        e[a][b] = 0.5 * (L[a][b] + L[b][a]);
        w[a][b] = 0.5 * (L[a][b] - L[b][a]);
        S[a][b] = st[sidx(a, b)];
        s3[a][b] = v.s[sidx(a, b)];
      }
    double J[3][3];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        J[a][b] = S[a][0] * w[b][0] + S[a][1] * w[b][1] + S[a][2] * w[b][2] +
                  S[0][b] * w[a][0] + S[1][b] * w[a][1] + S[2][b] * w[a][2];
    double tr = e[0][0] + e[1][1] + e[2][2];
    double ed[3][3];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) ed[a][b] = e[a][b] - (a == b ? tr / 3.0 : 0.0);
    bool plastic = v.f >= -eps_f && v.sJ2 > 1e-8;
    double lam = 0;
    double gt[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    if (plastic) {
      double se = 0;
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) se += s3[a][b] * e[a][b];
      lam = (3.0 * alpha * K * tr + (G / v.sJ2) * se) / (27.0 * alpha * K * sin_psi + G);
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
          gt[a][b] = lam * ((a == b ? 9.0 * K * sin_psi : 0.0) + G / v.sJ2 * s3[a][b]);
    }
    double ds[3][3];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        ds[a][b] = J[a][b] + (2.0 * G * ed[a][b] + (a == b ? K * tr : 0.0)) - gt[a][b];
    for (int a = 0; a < 3; ++a)
      for (int b = a; b < 3; ++b) dsig[sidx(a, b)][i] = ds[a][b];
    double sd = 0;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) sd += ed[a][b] * ed[a][b];
    dseq[i] = std::sqrt(sd * 2.0 / 3.0);
    double dp = 0;
    if (plastic) {
      double gp = v.sJ2 + 3.0 * v.I1 * sin_psi;  // pti.g_p is never written: 0
      double ep[3][3];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
          ep[a][b] = (std::fabs(ds[a][b]) > 1e-8 ? gp / ds[a][b] : 0.0) * lam;
      double t3 = (ep[0][0] + ep[1][1] + ep[2][2]) / 3.0;
      double q = 0;
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          double d = ep[a][b] - (a == b ? t3 : 0.0);
          q += d * d;
        }
      dp = std::sqrt(q * 2.0 / 3.0);
    }
    dseqp[i] = dp;
  }
  void dp_soil(int64_t i) {  // solver_sph_dp.py:237-270
    R L[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, dd = 0;
    walk(i, [&](int64_t j, R x, R y, R z, R r2) {
      R f = gradW_fac(std::sqrt(r2));
      R gw[3] = {f * x, f * y, f * z};
      corr(i, gw);
      R u[3] = {mV[j] * (vtx[j] - vtx[i]), mV[j] * (vty[j] - vty[i]), mV[j] * (vtz[j] - vtz[i])};
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) L[a][b] += u[a] * gw[b];
      dd += (mV[j] * (vtx[i] - vtx[j])) * gw[0] + (mV[j] * (vty[i] - vty[j])) * gw[1] +
            (mV[j] * (vtz[i] - vtz[j])) * gw[2];
    });
    drho[i] = dd * rhot[i];
    double Ld[3][3];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) Ld[a][b] = (double)L[a][b];
    bui2008(i, Ld);
    R dv[3];
    soil_dvel_sum(i, dv);
    soil_dvel_finish(i, dv);
  }
  void dp_one_step() {  // solver_sph_dp.py:210-274
#pragma omp parallel for
    for (int64_t i = 0; i < n; ++i)
      if (type[i] == SOIL) {
        double st[6];
        for (int c = 0; c < 6; ++c) st[c] = sigt[c][i];
        adapt(st);
        for (int c = 0; c < 6; ++c) sigt[c][i] = st[c];
      }
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i)
      if (is_bdy(type[i]) || type[i] == RIGID) soil_bdy(i, false);  // :220-231
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
      if (type[i] == SOIL) dp_soil(i);
      else if (type[i] == RIGID) soil_rigid(i);
    }
  }
  void dp_advect_something() {  // solver_sph_dp.py:276-296
#pragma omp parallel for
    for (int64_t i = 0; i < n; ++i) {
      if (type[i] != SOIL) continue;
      rho[i] = rho[i] > rho0 ? rho[i] : rho0;
      mV[i] = mass[i] / rho[i];
      double st[6];
      for (int c = 0; c < 6; ++c) st[c] = sig[c][i];
      flag[i] = upd_flag(st);
      adapt(st);
      for (int c = 0; c < 6; ++c) sig[c][i] = st[c];
      seq[i] += P.dt * dseq[i];
      seqp[i] += P.dt * dseqp[i];
    }
  }

  // ----------------------------------------------------------- integrators
  void one_step() {
    if (P.method == WC) wc_one_step();
    else if (P.method == MUI) mui_one_step();
    else dp_one_step();
  }
  bool soil_stress_rate() const { return P.method == DP; }
  void lf_half() {  // advect_LF_half (solver_sph_base.py:96-104)
#pragma omp parallel for
    for (int64_t i = 0; i < n; ++i) {
      if (is_real(type[i])) {
        rhot[i] += half_dt * drho[i];
        mV[i] = mass[i] / rhot[i];
        vtx[i] += half_dt * dvx[i];
        vty[i] += half_dt * dvy[i];
        vtz[i] += half_dt * dvz[i];
      }
      if (type[i] == SOIL && soil_stress_rate())
        for (int c = 0; c < 6; ++c) sigt[c][i] += 0.5 * P.dt * dsig[c][i];
    }
  }
  void lf_full() {  // advect_LF (:106-114); advect_SE (:79-87) is the same
#pragma omp parallel for
    for (int64_t i = 0; i < n; ++i) {
      if (is_real(type[i])) {
        rho[i] += dt * drho[i];
        mV[i] = mass[i] / rho[i];
        vx[i] += dt * dvx[i];
        vy[i] += dt * dvy[i];
        vz[i] += dt * dvz[i];
      }
      if (type[i] == SOIL && soil_stress_rate())
        for (int c = 0; c < 6; ++c) sig[c][i] += P.dt * dsig[c][i];
    }
  }
  // RK4 (solver_sph_base.py:134-180)
  void rk_init() {
    for (int64_t i = 0; i < n; ++i) {
      if (is_real(type[i])) { drhoRK[i] = 0; dvRKx[i] = dvRKy[i] = dvRKz[i] = 0; }
      if (type[i] == SOIL) for (int c = 0; c < 6; ++c) dsigRK[c][i] = 0;
    }
  }
  void rk_update(int m) {
    R mr = (R)m;
    for (int64_t i = 0; i < n; ++i) {
      if (is_real(type[i])) {
        drhoRK[i] += drho[i] * mr;
        dvRKx[i] += dvx[i] * mr; dvRKy[i] += dvy[i] * mr; dvRKz[i] += dvz[i] * mr;
      }
      if (type[i] == SOIL) for (int c = 0; c < 6; ++c) dsigRK[c][i] += dsig[c][i] * m;
    }
  }
  void rk_advect4() {
    for (int64_t i = 0; i < n; ++i) {
      if (is_real(type[i])) {
        rhot[i] = half_dt * drho[i] + rho[i];
        mV[i] = mass[i] / rhot[i];
        vtx[i] = half_dt * dvx[i] + vx[i];
        vty[i] = half_dt * dvy[i] + vy[i];
        vtz[i] = half_dt * dvz[i] + vz[i];
      }
      if (type[i] == SOIL && soil_stress_rate())
        for (int c = 0; c < 6; ++c) sigt[c][i] = 0.5 * P.dt * dsig[c][i] + sig[c][i];
    }
  }
  void rk_advect() {
    R d6 = (R)(P.dt / 6.0);
    for (int64_t i = 0; i < n; ++i) {
      if (is_real(type[i])) {
        rho[i] += d6 * drhoRK[i];
        mV[i] = mass[i] / rho[i];
        vx[i] += d6 * dvRKx[i]; vy[i] += d6 * dvRKy[i]; vz[i] += d6 * dvRKz[i];
      }
      if (type[i] == SOIL && soil_stress_rate())
        for (int c = 0; c < 6; ++c) sig[c][i] += P.dt / 6.0 * dsigRK[c][i];
    }
  }
  void substep() {
    if (P.time_integration == 1) {
      one_step();
      lf_full();
    } else if (P.time_integration == 2) {
      one_step();
      lf_half();
      one_step();
      lf_full();
    } else {
      rk_init();
      const int m[4] = {1, 2, 2, 1};
      for (int k = 0; k < 4; ++k) {
        one_step();
        rk_update(m[k]);
        if (k < 3) rk_advect4();
      }
      rk_advect();
    }
  }
  // advect_pos (solver_sph_base.py:228-238) with D3.
  void advect_pos() {
    if (P.xsph) {
#pragma omp parallel for schedule(dynamic, 256)
      for (int64_t i = 0; i < n; ++i) {
        R s[3] = {0, 0, 0};
        if (is_real(type[i])) {
          walk(i, [&](int64_t j, R x, R y, R z, R r2) {
            if (type[j] != type[i]) return;
            R w = W(std::sqrt(r2));
            s[0] += mV[j] * (vx[j] - vx[i]) * w;
            s[1] += mV[j] * (vy[j] - vy[i]) * w;
            s[2] += mV[j] * (vz[j] - vz[i]) * w;
          });
        }
        xsx[i] = s[0]; xsy[i] = s[1]; xsz[i] = s[2];
      }
    }
#pragma omp parallel for
    for (int64_t i = 0; i < n; ++i) {
      if (!is_real(type[i])) continue;
      if (P.xsph) {
        ox[i] += dt * (vx[i] + R(0.5) * xsx[i]);
        oy[i] += dt * (vy[i] + R(0.5) * xsy[i]);
        oz[i] += dt * (vz[i] + R(0.5) * xsz[i]);
      } else {
        ox[i] += dt * vx[i];
        oy[i] += dt * vy[i];
        oz[i] += dt * vz[i];
      }
      if (dim == 2) oz[i] = R(0);
    }
  }
  void advect_something() {
    if (P.method == WC) {  // solver_sph_wc.py:129-133 -> chk_density
#pragma omp parallel for
      for (int64_t i = 0; i < n; ++i)
        if (type[i] == FLUID) {
          rho[i] = rho[i] > rho0 ? rho[i] : rho0;
          mV[i] = mass[i] / rho[i];
        }
    } else if (P.method == MUI) {
      mui_advect_something();
    } else {
      dp_advect_something();
    }
  }
  double abs_pos(int a, int64_t i) const {
    const int32_t c = a == 0 ? cx[i] : a == 1 ? cy[i] : cz[i];
    const R o = a == 0 ? ox[i] : a == 1 ? oy[i] : oz[i];
    return P.vdomain_start[a] + (double)c * grid + (double)o;
  }
  double abs_or0(int a, int64_t i) const { return (a < 2 || dim == 3) ? abs_pos(a, i) : 0.0; }
  void set_abs(int a, int64_t i, double x) {
    int32_t& c = a == 0 ? cx[i] : a == 1 ? cy[i] : cz[i];
    R& o = a == 0 ? ox[i] : a == 1 ? oy[i] : oz[i];
    o = (R)(x - P.vdomain_start[a] - (double)c * grid);
  }
  // enforce_boundary_2D/3D (solver_sph_base.py:540-601) for the particles
  // of judge_enforce_bdy (:532-538): flow particles with boundary == 1,
  // dynamic rigid particles always.
  void enforce_boundary() {
    if (P.boundary != 1 && nrig == 0) return;
    double dist = P.particle_radius - 1e-8;
    for (int64_t i = 0; i < n; ++i) {
      if (!((P.boundary == 1 && is_flow(type[i])) || rigid_slot(i) >= 0)) continue;
      double pos[3] = {abs_pos(0, i), abs_pos(1, i), dim == 3 ? abs_pos(2, i) : 0.0};
      double nrm[3] = {0, 0, 0};
      if (pos[0] > P.domain_end[0] - dist) { nrm[0] += 1; set_abs(0, i, P.domain_end[0] - dist); }
      if (pos[0] <= P.domain_start[0] + dist) { nrm[0] -= 1; set_abs(0, i, P.domain_start[0] + dist); }
      if (pos[1] <= P.domain_start[1] + dist) { nrm[1] -= 1; set_abs(1, i, P.domain_start[1] + dist); }
      if (dim == 3) {
        if (pos[2] > P.domain_end[2] - dist) { nrm[2] += 1; set_abs(2, i, P.domain_end[2] - dist); }
        if (pos[2] <= P.domain_start[2] + dist) { nrm[2] -= 1; set_abs(2, i, P.domain_start[2] + dist); }
      }
      double len = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
      if (len > 1e-8) {
        R u[3] = {(R)(nrm[0] / len), (R)(nrm[1] / len), (R)(nrm[2] / len)};
        R vd = vx[i] * u[0] + vy[i] * u[1] + vz[i] * u[2];
        R k = R(1.3) * vd;
        vx[i] -= k * u[0];
        vy[i] -= k * u[1];
        vz[i] -= k * u[2];
      }
    }
  }
  void step() {  // SPHBase.step (solver_sph_base.py:41-51)
    sort();
    shepard();
    init_tmp();
    substep();
    advect_pos();
    advect_something();
    solve_rigid();
    enforce_boundary();
  }

  void globals(double out[4]) const {
    double front = -std::numeric_limits<double>::infinity(), pe = 0, ke = 0, m = 0;
    for (int64_t i = 0; i < n; ++i) {
      if (!is_flow(type[i])) continue;
      double x = abs_pos(0, i), y = abs_pos(1, i), z = dim == 3 ? abs_pos(2, i) : 0.0;
      double mi = (double)mass[i];
      front = std::max(front, x);
      pe += mi * -(P.gravity[0] * x + P.gravity[1] * y + P.gravity[2] * z);
      double v2 = (double)vx[i] * vx[i] + (double)vy[i] * vy[i] + (double)vz[i] * vz[i];
      ke += 0.5 * mi * v2;
      m += mi;
    }
    out[0] = front; out[1] = pe; out[2] = ke; out[3] = m;
  }

  // ------------------------------------------------------ field transfer
  int set(int f, const void* data) {
    const double* d = (const double*)data;
    const int32_t* iv = (const int32_t*)data;
    auto s1 = [&](std::vector<R>& v) { for (int64_t i = 0; i < n; ++i) v[i] = (R)d[i]; };
    auto s3 = [&](std::vector<R>& a, std::vector<R>& b, std::vector<R>& c) {
      for (int64_t i = 0; i < n; ++i) { a[i] = (R)d[3 * i]; b[i] = (R)d[3 * i + 1]; c[i] = (R)d[3 * i + 2]; }
    };
    auto s6 = [&](std::vector<double>* t) {
      for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 6; ++c) t[c][i] = rnd_mui(d[6 * i + c]);
    };
    switch (f) {
      case F_X0:
        for (int64_t i = 0; i < n; ++i)
          for (int c = 0; c < 3; ++c) x0[c][i] = d[3 * i + c];
        return 0;
      case F_POS:
        for (int64_t i = 0; i < n; ++i) {
          int32_t* cc[3] = {&cx[i], &cy[i], &cz[i]};
          R* oo[3] = {&ox[i], &oy[i], &oz[i]};
          for (int a = 0; a < 3; ++a) {
            if (a == 2 && dim == 2) { *cc[a] = 0; *oo[a] = R(0); continue; }
            double rel = d[3 * i + a] - P.vdomain_start[a];
            int32_t c = (int32_t)(rel / grid);  // pos_to_index (particle_system.py:216-218)
            *cc[a] = c;
            R o = (R)(rel - (double)c * grid);
            *oo[a] = o;
          }
        }
        return 0;
      case F_VEL: s3(vx, vy, vz); return 0;
      case F_DENSITY: s1(rho); return 0;
      case F_MASS: s1(mass); return 0;
      case F_M_V: s1(mV); return 0;
      case F_PRESSURE: s1(pr); return 0;
      case F_TYPE: for (int64_t i = 0; i < n; ++i) type[i] = iv[i]; return 0;
      case F_ID0: for (int64_t i = 0; i < n; ++i) id0[i] = iv[i]; return 0;
      case F_OBJ_ID: for (int64_t i = 0; i < n; ++i) obj[i] = iv[i]; return 0;
      case F_STRESS: s6(sig); return 0;
      case F_STRESS_TMP: s6(sigt); return 0;
      case F_D_STRESS: for (int64_t i = 0; i < n; ++i) for (int c = 0; c < 6; ++c) dsig[c][i] = d[6 * i + c]; return 0;
      case F_STRAIN_EQU: for (int64_t i = 0; i < n; ++i) seq[i] = rnd_mui(d[i]); return 0;
      case F_STRAIN_EQU_P: for (int64_t i = 0; i < n; ++i) seqp[i] = d[i]; return 0;
      case F_D_STRAIN_EQU: for (int64_t i = 0; i < n; ++i) dseq[i] = rnd_mui(d[i]); return 0;
      case F_D_STRAIN_EQU_P: for (int64_t i = 0; i < n; ++i) dseqp[i] = d[i]; return 0;
      case F_FLAG_RETMAP: for (int64_t i = 0; i < n; ++i) flag[i] = iv[i]; return 0;
      case F_CSPM_F: s1(cspm); return 0;
      case F_D_DENSITY: s1(drho); return 0;
      case F_D_VEL: s3(dvx, dvy, dvz); return 0;
      case F_DENSITY_TMP: s1(rhot); return 0;
      case F_V_TMP: s3(vtx, vty, vtz); return 0;
      case F_CELL:
        for (int64_t i = 0; i < n; ++i) {
          int64_t k = iv[i];
          cz[i] = (int32_t)(k % gn[2]);
          cy[i] = (int32_t)((k / gn[2]) % gn[1]);
          cx[i] = (int32_t)(k / ((int64_t)gn[2] * gn[1]));
        }
        return 0;
      case F_OFFSET: s3(ox, oy, oz); return 0;
    }
    return -1;
  }
  int get(int f, void* data) const {
    double* d = (double*)data;
    int32_t* iv = (int32_t*)data;
    auto g1 = [&](const std::vector<R>& v) { for (int64_t i = 0; i < n; ++i) d[i] = (double)v[i]; };
    auto g3 = [&](const std::vector<R>& a, const std::vector<R>& b, const std::vector<R>& c) {
      for (int64_t i = 0; i < n; ++i) { d[3 * i] = a[i]; d[3 * i + 1] = b[i]; d[3 * i + 2] = c[i]; }
    };
    auto g6 = [&](const std::vector<double>* t) {
      for (int64_t i = 0; i < n; ++i) for (int c = 0; c < 6; ++c) d[6 * i + c] = t[c][i];
    };
    switch (f) {
      case F_X0:
        for (int64_t i = 0; i < n; ++i)
          for (int c = 0; c < 3; ++c) d[3 * i + c] = x0[c][i];
        return 0;
      case F_POS:
        for (int64_t i = 0; i < n; ++i)
          for (int a = 0; a < 3; ++a) d[3 * i + a] = (a == 2 && dim == 2) ? 0.0 : abs_pos(a, i);
        return 0;
      case F_VEL: g3(vx, vy, vz); return 0;
      case F_DENSITY: g1(rho); return 0;
      case F_MASS: g1(mass); return 0;
      case F_M_V: g1(mV); return 0;
      case F_PRESSURE: g1(pr); return 0;
      case F_TYPE: for (int64_t i = 0; i < n; ++i) iv[i] = type[i]; return 0;
      case F_ID0: for (int64_t i = 0; i < n; ++i) iv[i] = id0[i]; return 0;
      case F_OBJ_ID: for (int64_t i = 0; i < n; ++i) iv[i] = obj[i]; return 0;
      case F_STRESS: g6(sig); return 0;
      case F_STRESS_TMP: g6(sigt); return 0;
      case F_D_STRESS: g6(dsig); return 0;
      case F_STRAIN_EQU: for (int64_t i = 0; i < n; ++i) d[i] = seq[i]; return 0;
      case F_STRAIN_EQU_P: for (int64_t i = 0; i < n; ++i) d[i] = seqp[i]; return 0;
      case F_D_STRAIN_EQU: for (int64_t i = 0; i < n; ++i) d[i] = dseq[i]; return 0;
      case F_D_STRAIN_EQU_P: for (int64_t i = 0; i < n; ++i) d[i] = dseqp[i]; return 0;
      case F_FLAG_RETMAP: for (int64_t i = 0; i < n; ++i) iv[i] = flag[i]; return 0;
      case F_CSPM_F: g1(cspm); return 0;
      case F_D_DENSITY: g1(drho); return 0;
      case F_D_VEL: g3(dvx, dvy, dvz); return 0;
      case F_DENSITY_TMP: g1(rhot); return 0;
      case F_V_TMP: g3(vtx, vty, vtz); return 0;
      case F_CELL: for (int64_t i = 0; i < n; ++i) iv[i] = (int32_t)flat(cx[i], cy[i], cz[i]); return 0;
      case F_OFFSET: g3(ox, oy, oz); return 0;
      case F_CELL_START: for (int64_t c = 0; c <= C; ++c) iv[c] = (int32_t)start[c]; return 0;
      case F_CSPM_L:
        for (int64_t i = 0; i < n; ++i)
          for (int c = 0; c < 9; ++c) d[9 * i + c] = (double)Lc[c][i];
        return 0;
    }
    return -1;
  }

  int pass(int id) {
    switch (id) {
      case P_SORT: sort(); return 0;
      case P_SHEPARD: shepard(); return 0;
      case P_INIT_TMP: init_tmp(); return 0;
      case P_ONE_STEP: one_step(); return 0;
      case P_LF_HALF: lf_half(); return 0;
      case P_LF_FULL: lf_full(); return 0;
      case P_SE: lf_full(); return 0;
      case P_ADVECT_POS: advect_pos(); return 0;
      case P_ADVECT_SOMETHING: advect_something(); return 0;
      case P_ENFORCE_BOUNDARY: enforce_boundary(); return 0;
      case P_SOLVE_RIGID: solve_rigid(); return 0;
      case P_STEP: step(); return 0;
    }
    return -1;
  }
  int64_t neighbors(int64_t i, int64_t* out, int64_t cap) const {
    int64_t k = 0;
    walk(i, [&](int64_t j, R, R, R, R) {
      if (k < cap) out[k] = j;
      ++k;
    });
    return k;
  }
  // Every particle's neighbours (for_all_neighbors order) in one call:
  // counts[i], then id0 of each neighbour and its stencil cell code
  // (dx+1) << 4 | (dy+1) << 2 | (dz+1) (d = neighbour cell - own cell),
  // concatenated in particle order.  Returns the total; fills ids / codes
  // only when they hold it.
  int64_t neighbor_csr(int64_t* counts, int32_t* ids, int8_t* codes, int64_t cap) const {
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = 0; i < n; ++i) {
      int64_t k = 0;
      walk(i, [&](int64_t, R, R, R, R) { ++k; });
      counts[i] = k;
    }
    std::vector<int64_t> off(n + 1, 0);
    for (int64_t i = 0; i < n; ++i) off[i + 1] = off[i] + counts[i];
    if (off[n] > cap || !ids || !codes) return off[n];
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = 0; i < n; ++i) {
      int64_t k = off[i];
      walk(i, [&](int64_t j, R, R, R, R) {
        ids[k] = id0[j];
        codes[k] = (int8_t)(((cx[j] - cx[i] + 1) << 4) | ((cy[j] - cy[i] + 1) << 2) | (cz[j] - cz[i] + 1));
        ++k;
      });
    }
    return off[n];
  }
};

struct Handle {
  int prec;
  std::unique_ptr<Sim<float>> f;
  std::unique_ptr<Sim<double>> d;
};

}  // namespace

#define DISPATCH(h, expr)                                   \
  do {                                                      \
    Handle* H = (Handle*)(h);                               \
    if (H->prec == 32) { auto& S = *H->f; return expr; }    \
    auto& S = *H->d;                                        \
    return expr;                                            \
  } while (0)

extern "C" {
void* orc_create(const orc_params* p, int precision) {
  auto* h = new Handle();
  h->prec = precision;
  if (precision == 32) {
    h->f.reset(new Sim<float>());
    h->f->init(p);
  } else {
    h->d.reset(new Sim<double>());
    h->d->init(p);
  }
  return h;
}
void orc_destroy(void* h) { delete (Handle*)h; }
int orc_set(void* h, int field, const void* data) { DISPATCH(h, S.set(field, data)); }
int orc_get(void* h, int field, void* data) { DISPATCH(h, S.get(field, data)); }
int orc_pass(void* h, int pass_id) { DISPATCH(h, S.pass(pass_id)); }
int orc_step(void* h, int nsteps) {
  for (int k = 0; k < nsteps; ++k) orc_pass(h, P_STEP);
  return 0;
}
int orc_set_rigid_object(void* h, int obj, const double cm[3]) { DISPATCH(h, S.add_rigid(obj, cm)); }
int orc_polar(const double* A, int dim, double* R) { return polar_rotation(A, dim, R), 0; }
int orc_globals(void* h, double out[4]) { DISPATCH(h, (S.globals(out), 0)); }
int64_t orc_neighbors(void* h, int64_t i, int64_t* out, int64_t cap) { DISPATCH(h, S.neighbors(i, out, cap)); }
int64_t orc_neighbor_csr(void* h, int64_t* counts, int32_t* ids, int8_t* codes, int64_t cap) {
  DISPATCH(h, S.neighbor_csr(counts, ids, codes, cap));
}
int64_t orc_escaped(void* h) { DISPATCH(h, S.escaped); }
// ---- function-level hooks for closed-form tests -------------------------
int orc_kernel_eval(const orc_params* p, int precision, const double* r, int64_t n, double* W,
                    double* gradfac) {
  orc_params q = *p;
  q.n_particles = 0;
  if (precision == 32) {
    Sim<float> S;
    S.init(&q);
    for (int64_t k = 0; k < n; ++k) {
      W[k] = S.W((float)r[k]);
      gradfac[k] = S.gradW_fac((float)r[k]);
    }
  } else {
    Sim<double> S;
    S.init(&q);
    for (int64_t k = 0; k < n; ++k) {
      W[k] = S.W(r[k]);
      gradfac[k] = S.gradW_fac(r[k]);
    }
  }
  return 0;
}
// flags (upd_flag_DP of the input) then adapt_stress in place.
int orc_dp_adapt(const orc_params* p, double* st6, int64_t n, int32_t* flags) {
  orc_params q = *p;
  q.n_particles = 0;
  Sim<double> S;
  S.init(&q);
  for (int64_t k = 0; k < n; ++k) {
    flags[k] = S.upd_flag(st6 + 6 * k);
    S.adapt(st6 + 6 * k);
  }
  return 0;
}
// Bui (2008) stress rate for given velocity gradients L (row-major 3x3).
int orc_dp_rate(const orc_params* p, const double* L9, const double* st6, int64_t n, double* ds6,
                double* dseq, double* dseqp) {
  orc_params q = *p;
  q.n_particles = 1;
  Sim<double> S;
  S.init(&q);
  for (int64_t k = 0; k < n; ++k) {
    double L[3][3];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) L[a][b] = L9[9 * k + 3 * a + b];
    for (int c = 0; c < 6; ++c) S.sigt[c][0] = st6[6 * k + c];
    S.bui2008(0, L);
    for (int c = 0; c < 6; ++c) ds6[6 * k + c] = S.dsig[c][0];
    dseq[k] = S.dseq[0];
    dseqp[k] = S.dseqp[0];
  }
  return 0;
}
void orc_set_threads(int n) { omp_set_num_threads(n); }
int orc_get_threads(void) { return omp_get_max_threads(); }
}
