// tsph_device.h — device-side data layout and per-pair arithmetic of the
// MI355X SPH engine (gfx950, wave64).
//
// Data layout in HBM (structure of arrays, one entry per particle slot of
// the engine's particle capacity; every scene owns a contiguous range):
//   cell   int32   scene-local flattened cell id  (ix*gny + iy)*gnz + iz
//   pos    float4  cell-local offset x,y,z (metres) and m_V
//   vel    float4  velocity and density
//   vtmp   float4  v_tmp and density_tmp (leap-frog half state)
//   mass   float, type int8, id0/obj int32, stress (6 comps, component-major)
// Positions are an integer cell plus an FP32 offset from the cell corner:
// the pair vector x_i - x_j = (c_i - c_j)*grid + (o_i - o_j) keeps ~1e-9 m
// resolution anywhere in the domain.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tsph {

constexpr int kFluid = 1, kSoil = 2, kRigid = 11, kDummy = -1, kRep = -2;
constexpr int kWC = 1, kMUI = 2, kDP = 3;
constexpr int kWave = 64;

__device__ __forceinline__ bool is_flow(int t) { return t == kFluid || t == kSoil; }
__device__ __forceinline__ bool is_real(int t) { return t > 0; }
__device__ __forceinline__ bool is_bdy(int t) { return t == kDummy || t == kRep; }
// Particles the WC pair passes treat as walls (dummy extrapolation, flow
// neighbours only): boundary and rigid (solver_sph_wc.py:90).
__device__ __forceinline__ bool is_solid(int t) { return is_bdy(t) || t == kRigid; }
constexpr int kMaxRigid = 4;  // dynamic rigid objects per scene

// Stress components xx,yy,zz,xy,yz,zx.
__device__ __forceinline__ constexpr int sidx(int a, int b) {
  return a == b ? a : ((a + b) == 1 ? 3 : ((a + b) == 3 ? 4 : 5));
}

// Per-scene parameters (host fills from tsph_scene_desc; float copies are
// the float64 values rounded once, like the reference's compile-time
// constants folded into float32 code).
struct SceneDev {
  int32_t active, dim, method, boundary, kernel, ti, xsph, iexp;
  int32_t gn[3];
  int32_t tile_first, n_tiles;
  int32_t pbase, n, cbase, ncell;
  double grid, vds[3], dstart[3], dend[3], radius, dt_d;
  float g, rs2, h1, kW, kC, eps;
  float kG, kGc;  // gradient factors: -5 kW h1^2 (Wendland C2), kC h1^2 (cubic)
  float rs2c;     // stencil-cell culling threshold 1.0001 rs2
  int kcorr;      // kernelCorrection: 1 CSPM, 2 MLS switch (zero corrected gradient)
  int rep;        // add the repulsive-particle force (boundary 3 / 4; WC, muI)
  float rep_k, rep_d, rep_h;  // 0.01 c^2, particle diameter, 0.75 h
  float dt, half_dt, dt6, rho0, visc_c1, visc_c2, grav[3];
  float stiffness, exponent, vsound2, coh, mu, young, h2;
  double alpha, kc, G, K, sin_psi, eps_f;
  // dynamic rigid objects (isDynamic blocks): object ids and rest centres
  // of mass (calc_rigid_rest_m, solver_sph_base.py:467-470, 517-518)
  int32_t nrig, rig_obj[kMaxRigid];
  double rig_cm0[kMaxRigid][3];
};

// Index of particle p's object among the scene's dynamic rigid objects, -1
// for a static rigid or any other particle.
__device__ __forceinline__ int rigid_slot(const SceneDev& S, int type, int obj) {
  if (type != kRigid) return -1;
  // constant indices only: a runtime index into the scene record would put
  // the whole record on the stack of every kernel that calls this
  int r = -1;
#pragma unroll
  for (int k = kMaxRigid - 1; k >= 0; --k)
    if (k < S.nrig && S.rig_obj[k] == obj) r = k;
  return r;
}

// Particle state that survives a step (double-buffered; the counting sort
// moves it from one buffer to the other).
struct PState {
  int32_t* cell;
  float4* pos;
  float4* vel;
  float4* vtmp;
  float* mass;
  int8_t* type;
  int32_t* id0;
  int32_t* obj;
  float* sig_f;    // 6 x cap, muI stress
  double* sig_d;   // 6 x cap, DP stress
  double* seq;     // strain_equ
  double* seqp;    // strain_equ_p
  int32_t* flag;   // flag_retmap
};

// Per-step work arrays (single-buffered).
struct PWork {
  int32_t* key;      // global cell key (pre-sort order)
  int32_t* slot;     // arrival slot inside the cell (histogram atomics)
  int32_t* tmp;      // scan position -> particle
  int32_t* sid;      // particle -> scene slot (static per scene range)
  float* p;          // pressure
  float* Lc;          // inverse CSPM_L, row-major 3x3, entry c at [c*cap + p]
  float4* vp;        // WC pair record of a particle: v_tmp, p / density_tmp^2;
                     // sign bit of .w set for a boundary particle
  float* cspm;       // Shepard factor
  float* drho;       // d_density
  float4* dv;        // d_vel
  float* sigt_f;     // 6 x cap, stress_tmp in float (muI; DP mirror)
  double* sigt_d;    // 6 x cap, DP stress_tmp
  double* dsig_d;    // 6 x cap, DP d_stress
  double* dseq;      // d_strain_equ
  double* dseqp;     // d_strain_equ_p
  float4* xs;        // XSPH sums
  float* drho_rk;    // RK4 accumulators
  float4* dv_rk;
  double* dsig_rk;
  int32_t* count;    // cell histogram (cell capacity)
  uint8_t* cflow;    // 1 if the cell holds a flow (fluid/soil) particle
  int32_t* start;    // exclusive scan (cell capacity + 1)
  int32_t* bsum;     // scan block sums
  unsigned long long* status;  // per slot: escaped, NaN
  double* gpart;     // globals partials per tile (4 each)
  double* gout;      // globals per slot (4 each)
  uint32_t* nl;      // neighbour lists, blocked per 64-particle tile (nl_base / nl_at)
  int32_t* nlc;      // neighbour count per particle (> nl_cap: walk instead)
  // colour feed (allocated on first use)
  double* val;       // colour value per particle (current order)
  float4* rgb;       // colour per particle (current order)
  double* x0;        // 3 x cap initial position, indexed by scene base + id0
  float4* c0;        // initial colour, indexed by scene base + id0
  double* crange;    // per slot: vmax, vmin of the last v_maxmin
  // shape matching of dynamic rigid objects (allocated with the first one)
  double* rigpart;   // per tile and object: 16 partial sums
  double* rigout;    // per slot and object: centre of mass (3), rotation (9)
};

struct KArgs {
  const SceneDev* __restrict__ scenes;
  const int2* __restrict__ tiles;  // (slot, first particle) per 64-particle tile
  PState cur, alt;
  PWork w;
  int32_t cap;     // particle capacity
  int32_t ccap;    // cell capacity
  int32_t stride;  // component stride of 6-vectors (== cap)
  int32_t nl_cap;  // neighbour-list capacity per particle
};

// ---------------------------------------------------------- pair distance
// |x_ij| from r2 >= 0: v_sqrt_f32 (within an ulp), then the neighbouring
// float whose square is on the right side of r2 (the residual sign decides,
// with exact fused multiply-adds).  Equal bit for bit to the correctly
// rounded sqrtf for every r2 >= 2^-96 (checked over all those floats by
// tsph_check_sqrt); below 2^-96 it returns a value under 1e-8, which the
// kernels below map to 0 exactly as they map sqrtf's (r > 1e-8 test).  It
// drops sqrtf's rescaling of tiny arguments and its special-value fixup:
// 7 of its 16 instructions per pair.
__device__ __forceinline__ float pair_dist(float r2) {
  const float s = __builtin_amdgcn_sqrtf(r2);
  const float sd = __int_as_float(__float_as_int(s) - 1);
  const float su = __int_as_float(__float_as_int(s) + 1);
  const float rd = fmaf(-sd, s, r2);
  const float ru = fmaf(-su, s, r2);
  float r = rd <= 0.0f ? sd : s;
  r = ru > 0.0f ? su : r;
  return r;
}

// ------------------------------------------------------------- kernels W
// Wendland C2 (solver_sph_base.py:334-358) and cubic spline (:298-330).
__device__ __forceinline__ float kernel_W(const SceneDev& S, float rn) {
  float q = rn * S.h1;
  if (!(rn > S.eps && q <= 2.0f)) return 0.0f;
  if (S.kernel == 1) {
    float q1 = 1.0f - 0.5f * q;
    float q2 = q1 * q1;
    return S.kW * (q2 * q2) * (1.0f + 2.0f * q);
  }
  if (q <= 1.0f) {
    float qq = q * q;
    return S.kC * (0.5f * (qq * q) - qq + (float)(2.0 / 3.0));
  }
  float t = 2.0f - q;
  return S.kC / 6.0f * (t * t * t);
}
// grad W = fac * x_ij with fac = W'(r)/r.  With q = r h1, q h1 / r = h1^2,
// so W'(r)/r = -5 kW h1^2 q1^3 (Wendland C2), kC h1^2 (1.5 q - 2) (cubic,
// q <= 1) and -0.5 kC h1^2 (2 - q)^2 / q (cubic, q > 1).
__device__ __forceinline__ float kernel_gradfac(const SceneDev& S, float rn) {
  float q = rn * S.h1;
  if (!(rn > S.eps && q <= 2.0f)) return 0.0f;
  if (S.kernel == 1) {
    float q1 = 1.0f - 0.5f * q;
    return S.kG * ((q1 * q1) * q1);
  }
  if (q <= 1.0f) return S.kGc * (1.5f * q - 2.0f);
  float t = 2.0f - q;
  return S.kGc * (-0.5f * t * t) / q;
}

// The same two functions with the kernel type fixed at compile time and no
// branch (both cubic pieces are formed and one is selected): identical
// values, but a pair loop built on them is straight-line code, so the
// loads of several neighbours can be in flight together.
template <int KT>
__device__ __forceinline__ float kernel_W_t(const SceneDev& S, float rn) {
  const float q = rn * S.h1;
  float w;
  if (KT == 1) {
    const float q1 = 1.0f - 0.5f * q;
    const float q2 = q1 * q1;
    w = S.kW * (q2 * q2) * (1.0f + 2.0f * q);
  } else {
    const float qq = q * q;
    const float wi = S.kC * (0.5f * (qq * q) - qq + (float)(2.0 / 3.0));
    const float t = 2.0f - q;
    const float wo = S.kC / 6.0f * (t * t * t);
    w = q <= 1.0f ? wi : wo;
  }
  return (rn > S.eps && q <= 2.0f) ? w : 0.0f;
}
template <int KT>
__device__ __forceinline__ float kernel_gradfac_t(const SceneDev& S, float rn) {
  const float q = rn * S.h1;
  float f;
  if (KT == 1) {
    const float q1 = 1.0f - 0.5f * q;
    f = S.kG * ((q1 * q1) * q1);
  } else {
    const float fi = S.kGc * (1.5f * q - 2.0f);
    const float t = 2.0f - q;
    const float fo = S.kGc * (-0.5f * t * t) / q;
    f = q <= 1.0f ? fi : fo;
  }
  return (rn > S.eps && q <= 2.0f) ? f : 0.0f;
}

// kernel_deriv_corr (solver_sph_base.py:375-384) with kernelCorrection 1:
// the gradient times the particle's stored inverse CSPM_L.
__device__ __forceinline__ void corr_load(const float* __restrict__ Lc, int stride, int p,
                                          float m[9]) {
#pragma unroll
  for (int c = 0; c < 9; ++c) m[c] = Lc[(size_t)c * stride + p];
}
__device__ __forceinline__ void corr_apply(const float m[9], float g[3]) {
  const float o0 = (m[0] * g[0] + m[1] * g[1]) + m[2] * g[2];
  const float o1 = (m[3] * g[0] + m[4] * g[1]) + m[5] * g[2];
  const float o2 = (m[6] * g[0] + m[7] * g[1]) + m[8] * g[2];
  g[0] = o0;
  g[1] = o1;
  g[2] = o2;
}
// Inverse of calc_CSPM_L (solver_sph_base.py:410-423): adjugate times 1/det
// when |det| > eps (2D: the upper 2x2 block, the rest zero); M keeps I
// otherwise.
template <int DIM>
__device__ __forceinline__ void corr_invert(const float L[3][3], float eps, float M[9]) {
  if (DIM == 2) {
    const float det = L[0][0] * L[1][1] - L[0][1] * L[1][0];
    if (fabsf(det) > eps) {
      const float id = 1.0f / det;
      M[0] = L[1][1] * id;
      M[1] = -L[0][1] * id;
      M[2] = 0.0f;
      M[3] = -L[1][0] * id;
      M[4] = L[0][0] * id;
      M[5] = 0.0f;
      M[6] = 0.0f;
      M[7] = 0.0f;
      M[8] = 0.0f;
    }
    return;
  }
  const float c00 = L[1][1] * L[2][2] - L[1][2] * L[2][1];
  const float c01 = L[1][2] * L[2][0] - L[1][0] * L[2][2];
  const float c02 = L[1][0] * L[2][1] - L[1][1] * L[2][0];
  const float det = L[0][0] * c00 + L[0][1] * c01 + L[0][2] * c02;
  if (fabsf(det) > eps) {
    const float id = 1.0f / det;
    M[0] = c00 * id;
    M[1] = (L[0][2] * L[2][1] - L[0][1] * L[2][2]) * id;
    M[2] = (L[0][1] * L[1][2] - L[0][2] * L[1][1]) * id;
    M[3] = c01 * id;
    M[4] = (L[0][0] * L[2][2] - L[0][2] * L[2][0]) * id;
    M[5] = (L[0][2] * L[1][0] - L[0][0] * L[1][2]) * id;
    M[6] = c02 * id;
    M[7] = (L[0][1] * L[2][0] - L[0][0] * L[2][1]) * id;
    M[8] = (L[0][0] * L[1][1] - L[0][1] * L[1][0]) * id;
  }
}

// Repulsive-particle force calc_repulsive_force (solver_sph_base.py:676-690):
// F = fac * x_ij with fac = 0.01 c^2 chi(r) f(r / (0.75 h)) / r^2.
__device__ __forceinline__ float rep_fac(const SceneDev& S, float rn) {
  const float chi = (rn > 0.0f && rn < S.rep_d) ? 1.0f - rn / S.rep_d : 0.0f;
  const float gam = rn / S.rep_h;
  const float two3 = (float)(2.0 / 3.0);
  float f = 0.0f;
  if (gam > 0.0f && gam <= two3) {
    f = two3;
  } else if (gam > two3 && gam <= 1.0f) {
    f = 2.0f * gam - 1.5f * (gam * gam);
  } else if (gam > 1.0f && gam < 2.0f) {
    const float t = 2.0f - gam;
    f = 0.5f * (t * t);
  }
  return S.rep_k * chi * f / (rn * rn);
}

template <class T>
__device__ __forceinline__ T ipow_dev(T x, int n) {
  T r = T(1), b = x;
  int m = n < 0 ? -n : n;
  while (m) {
    if (m & 1) r = r * b;
    b = b * b;
    m >>= 1;
  }
  return n < 0 ? T(1) / r : r;
}

// Tait equation of state (solver_sph_wc.py:88).
__device__ __forceinline__ float wc_eos(const SceneDev& S, float rt) {
  float x = rt / S.rho0;
  float xp = S.iexp != INT32_MIN ? ipow_dev(x, S.iexp) : powf(x, S.exponent);
  float v = S.stiffness * (xp - 1.0f);
  return v > 0.0f ? v : 0.0f;
}

// Cell of an offset along one axis: floor((double)o/grid) with the
// offset renormalised into [0, grid) (same rule as the oracle).
__device__ __forceinline__ void renorm(int32_t& c, float& o, double grid) {
  double od = (double)o;
  if (od != od) return;  // NaN offset: keep the cell (the NaN is reported by the step)
  double k = floor(od / grid);
  double r = od - k * grid;
  if (r >= grid) {
    k += 1.0;
    r -= grid;
  }
  float on = (float)r;
  if (on < 0.0f) on = 0.0f;
  if ((double)on >= grid) {
    k += 1.0;
    on = 0.0f;
  }
  c += (int32_t)k;
  o = on;
}

__device__ __forceinline__ void decode_cell(const SceneDev& S, int32_t c, int& x, int& y, int& z) {
  int yz = S.gn[1] * S.gn[2];
  x = c / yz;
  int r = c - x * yz;
  y = r / S.gn[2];
  z = r - y * S.gn[2];
}

// The neighbour walk of for_all_neighbors (particle_system.py:259-269):
// the 3^dim stencil in ndrange((-1,2),...) order around the centre cell
// (scx,scy,scz), particles j ascending in each cell, j != i and
// |x_ij|^2 < r_s^2.  Cells outside the grid are skipped.  (cix,ciy,ciz) is
// the cell the offset `oi` is measured from (equal to the centre except
// after advect_pos, before the next hash).
// fn(j, x, y, z, r2, m_V_j, code) with code = (dx+1)<<4 | (dy+1)<<2 | (dz+1)
// the stencil cell of j.
//
// CULL (only with the centre = i's own cell and offsets in [0, g), i.e.
// right after the hash): skip stencil cells whose box lies farther than
// r_s from i.  A candidate there is at least the per-axis gap away
// (offsets are >= 0 and < g), so its r2 exceeds rs2 by more than the
// rounding of either sum; the threshold rs2c = 1.0001 rs2 keeps that
// margin.  Visited pairs and their order are unchanged.
template <int DIM, bool CULL = false, class Fn>
__device__ __forceinline__ void walk(const SceneDev& S, const int32_t* __restrict__ start,
                                     const float4* __restrict__ pos, int32_t shift, int i,
                                     int scx, int scy, int scz, int cix, int ciy, int ciz,
                                     float oix, float oiy, float oiz, Fn&& fn) {
  const int gy = S.gn[1], gz = S.gn[2];
#pragma unroll 1
  for (int dx = -1; dx <= 1; ++dx) {
    const int nx = scx + dx;
    if (nx < 0 || nx >= S.gn[0]) continue;
    const float bx = oix + (float)(cix - nx) * S.g;
    const float ax = dx < 0 ? oix : (dx > 0 ? S.g - oix : 0.0f);
#pragma unroll 1
    for (int dy = -1; dy <= 1; ++dy) {
      const int ny = scy + dy;
      if (ny < 0 || ny >= gy) continue;
      const float by = oiy + (float)(ciy - ny) * S.g;
      const float ay = dy < 0 ? oiy : (dy > 0 ? S.g - oiy : 0.0f);
      const int32_t rowkey = S.cbase + (nx * gy + ny) * gz;
#pragma unroll 1
      for (int dz = (DIM == 3 ? -1 : 0); dz <= (DIM == 3 ? 1 : 0); ++dz) {
        const int nz = scz + dz;
        if (DIM == 3 && (nz < 0 || nz >= gz)) continue;
        if (CULL) {
          const float az = DIM == 3 ? (dz < 0 ? oiz : (dz > 0 ? S.g - oiz : 0.0f)) : 0.0f;
          if (fmaf(az, az, fmaf(ay, ay, ax * ax)) > S.rs2c) continue;
        }
        const float bz = DIM == 3 ? oiz + (float)(ciz - nz) * S.g : 0.0f;
        const int32_t key = rowkey + nz;
        const int code = ((dx + 1) << 4) | ((dy + 1) << 2) | (dz + 1);
        const int32_t jb = start[key] + shift;
        const int32_t je = start[key + 1] + shift;
        auto cand = [&](int32_t j, const float4& pj) {
          const float x = bx - pj.x;
          const float y = by - pj.y;
          const float z = DIM == 3 ? bz - pj.z : 0.0f;
          const float r2 = fmaf(z, z, fmaf(y, y, x * x));
          if (r2 < S.rs2 && j != i) fn(j, x, y, z, r2, pj.w, code);
        };
        int32_t j = jb;
        // four candidates' loads in flight before the first test
        for (; j + 4 <= je; j += 4) {
          const float4 p0 = pos[j], p1 = pos[j + 1], p2 = pos[j + 2], p3 = pos[j + 3];
          cand(j, p0);
          cand(j + 1, p1);
          cand(j + 2, p2);
          cand(j + 3, p3);
        }
        for (; j < je; ++j) cand(j, pos[j]);
      }
    }
  }
}

// Neighbour-list entry: particle index << 6 | stencil code (2 bits per axis).
constexpr int kNlShift = 6;
__device__ __forceinline__ uint32_t nl_entry(int j, int code) {
  return ((uint32_t)j << kNlShift) | (uint32_t)code;
}
// Storage of the lists, blocked per 64-particle tile and, inside a tile,
// in runs of 16 entries per particle: entry k of particle p sits at
// nl_base(p) + nl_at(k) with nl_base = tile * 64 * nl_cap + lane * 16 and
// nl_at(k) = (k / 16) * 1024 + k % 16.  A wave's lists are one contiguous
// run of 64 * nl_cap entries whatever the engine capacity (its traffic
// stays within a page or two); each lane appends into its own 64-byte run
// (four entries per aligned 16-byte store), and the replay loads a lane's
// entries k..k+3 with one 16-byte load.  (Runs of 4, which make the
// replay's loads one contiguous kilobyte per wave, measured ~4 % slower per
// step.)  nl_cap is a multiple of 16.
__host__ __device__ __forceinline__ size_t nl_base(int p, int nl_cap) {
  return (size_t)(p >> 6) * ((size_t)nl_cap << 6) + ((size_t)(p & 63) << 4);
}
__host__ __device__ __forceinline__ size_t nl_at(int k) {
  return ((size_t)(k >> 4) << 10) + (size_t)(k & 15);
}

// Per-thread table of the list replay (one 9 KiB LDS array per kernel that
// replays lists, shared by all its replays: a thread only ever touches its
// own column).
constexpr int kMaxBlock = 256;  // largest threads per block of a pair kernel
__device__ __forceinline__ float* shift_table() {
  __shared__ float tab[9 * kMaxBlock];
  return tab;
}

template <int DIM, class Fn>
__device__ __forceinline__ void for_list(const SceneDev& S, const uint32_t* __restrict__ nl,
                                         int n, int nl_cap, const float4* __restrict__ pos, int i,
                                         float oix, float oiy, float oiz, Fn&& fn) {
  const uint32_t* __restrict__ li = nl + nl_base(i, nl_cap);
  // shifted centres oi + (c_i - c_j) g for c_j - c_i = -1, 0, +1 (exact
  // products, rounded once: the same values the walk forms), kept in LDS as
  // row 3 * axis + code digit, column = thread, so that a pair looks its
  // three up by address (conflict-free: every lane reads its own column)
  // instead of by four compares and selects per axis
  float* __restrict__ tab = shift_table();
  const int t = (int)threadIdx.x;
  tab[0 * kMaxBlock + t] = oix + S.g;
  tab[1 * kMaxBlock + t] = oix;
  tab[2 * kMaxBlock + t] = oix + (-1.0f) * S.g;
  tab[3 * kMaxBlock + t] = oiy + S.g;
  tab[4 * kMaxBlock + t] = oiy;
  tab[5 * kMaxBlock + t] = oiy + (-1.0f) * S.g;
  if (DIM == 3) {
    tab[6 * kMaxBlock + t] = oiz + S.g;
    tab[7 * kMaxBlock + t] = oiz;
    tab[8 * kMaxBlock + t] = oiz + (-1.0f) * S.g;
  }
  auto pair = [&](uint32_t e, const float4& pj) {
    const int j = (int)(e >> kNlShift);
    const int code = (int)(e & 63u);
    const float x = tab[(int)((e >> 4) & 3u) * kMaxBlock + t] - pj.x;
    const float y = tab[(3 + (int)((e >> 2) & 3u)) * kMaxBlock + t] - pj.y;
    const float z = DIM == 3 ? tab[(6 + (int)(e & 3u)) * kMaxBlock + t] - pj.z : 0.0f;
    const float r2 = fmaf(z, z, fmaf(y, y, x * x));
    fn(j, x, y, z, r2, pj.w, code);
  };
  int k = 0;
  // Four entries at a time: one 16-byte load (a lane's entries k..k+3 are
  // contiguous), then the four neighbours' loads together, then the four
  // pairs in list order.
  for (; k + 4 <= n; k += 4) {
    const uint4 e4 = *reinterpret_cast<const uint4*>(li + nl_at(k));
    const float4 p0 = pos[e4.x >> kNlShift], p1 = pos[e4.y >> kNlShift];
    const float4 p2 = pos[e4.z >> kNlShift], p3 = pos[e4.w >> kNlShift];
    pair(e4.x, p0);
    pair(e4.y, p1);
    pair(e4.z, p2);
    pair(e4.w, p3);
  }
  for (; k < n; ++k) {
    const uint32_t e = li[nl_at(k)];
    pair(e, pos[e >> kNlShift]);
  }
}

}  // namespace tsph
