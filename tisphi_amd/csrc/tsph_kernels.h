// tsph_kernels.h — HIP kernels of the SPH step (gfx950).
//
// Launch model: every particle kernel runs over a fixed grid of 64-particle
// tiles; tiles[w] = (scene slot, first particle) for wave w, or slot -1 for
// an unused tile.  A wave therefore always belongs to one scene, and the
// scene parameters are read with wave-uniform (scalar) loads.  The grid is
// sized from the engine capacity, so one captured hipGraph serves any set
// of resident scenes.
#pragma once
#include <type_traits>

#include "tsph_device.h"

namespace tsph {

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  A grid that the chip holds at once (about 2048 threads per CU) is
// remapped so that every XCD walks one contiguous stretch of tiles:
// neighbouring particles (and their shared neighbour data) stay in one L2.
// A larger grid is remapped into kXcdStretches stretches per XCD, stretch c
// going to XCD c % 8: locality within a stretch, and an even share of the
// work per XCD when the work is not uniform along the tile range.  (One
// stretch per XCD left the 3D dam break's last two XCDs with 4 % of the
// fluid work: its water sits in the first 70 % of the particle order, and
// the busiest XCD set the pass time at 1.33x the even share.)  Both maps
// are bijections of [0, gridDim.x).  Stretched: block b on XCD x = b % 8 at
// position i = b / 8 takes logical block ((i / S) * 8 + x) * S + i % S for
// b below the largest multiple of 8 S, and b itself above it.
constexpr int kXcds = 8;
constexpr int kXcdStretches = 32;  // measured: 4 / 8 / 16 / 32 / 64 -> 4.03 / 3.99 / 3.96 / 3.94 / 3.95 ms per 3D step
constexpr int kOneRoundThreads = 256 * 2048;
__device__ __forceinline__ int xcd_block() {
  const int nb = gridDim.x, b = blockIdx.x;
  if (nb * (int)blockDim.x <= kOneRoundThreads) {
    const int q = nb / kXcds, r = nb % kXcds, x = b % kXcds;
    return x * q + (x < r ? x : r) + b / kXcds;
  }
  const int S = max(1, nb / (kXcds * kXcdStretches));  // blocks per stretch
  const int nmain = nb - nb % (kXcds * S);
  if (b >= nmain) return b;
  const int x = b % kXcds, i = b / kXcds;
  const int j = i / S, o = i - j * S;
  return (j * kXcds + x) * S + o;
}
__device__ __forceinline__ int wave_id() {
  return __builtin_amdgcn_readfirstlane((int)((xcd_block() * blockDim.x + threadIdx.x) >> 6));
}

// Thread -> (slot, particle).  False for idle lanes.
__device__ __forceinline__ bool tile_thread(const KArgs& a, int& s, int& p) {
  const int w = wave_id();
  const int2 t = a.tiles[w];
  if (t.x < 0) return false;
  s = __builtin_amdgcn_readfirstlane(t.x);
  p = t.y + (int)(threadIdx.x & 63);
  const SceneDev S = a.scenes[s];
  return p < S.pbase + S.n;
}

__device__ __forceinline__ int32_t scene_shift(const KArgs& a, const SceneDev& S) {
  return S.pbase - a.w.start[S.cbase];
}

// ===================================================================== sort
// update_grid_id (particle_system.py:228-236): renormalise the offset,
// flattened cell id, histogram with an arrival slot per particle.
template <int DIM>
__device__ __forceinline__ int32_t hash_one(const KArgs& a, int s, int p);
template <int DIM>
__global__ __launch_bounds__(256) void k_hash(KArgs a) {
  int s, p;
  const bool live = tile_thread(a, s, p);
  // Runs of equal keys among consecutive lanes share one histogram atomic:
  // the run's first lane adds the run length and hands out its slots.
  // The live lanes of a wave are a prefix (a tile is one scene's run of
  // particles); a wave with none leaves before any cross-lane operation.
  if (!__any(live)) return;
  int32_t key = -1;
  if (live) {
    key = hash_one<DIM>(a, s, p);
  }
  const int lane = (int)(threadIdx.x & 63);
  const int prev = __shfl_up(key, 1, 64);
  const bool head = live && (lane == 0 || prev != key);
  const unsigned long long heads = __ballot(head);
  const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1);
  const int first = 63 - __clzll(heads & upto);
  const unsigned long long above = heads & ~upto;
  const unsigned long long livemask = __ballot(live);
  const int nlive = __popcll(livemask);
  const int next = above ? (__ffsll((long long)above) - 1) : nlive;
  int base = 0;
  if (head) base = atomicAdd(&a.w.count[key], next - lane);
  base = __shfl(base, first, 64);
  if (live) a.w.slot[p] = base + (lane - first);
}

template <int DIM>
__device__ __forceinline__ int32_t hash_one(const KArgs& a, int s, int p) {
  const SceneDev S = a.scenes[s];
  int cx, cy, cz;
  decode_cell(S, a.cur.cell[p], cx, cy, cz);
  float4 P = a.cur.pos[p];
  renorm(cx, P.x, S.grid);
  renorm(cy, P.y, S.grid);
  if (DIM == 3) {
    renorm(cz, P.z, S.grid);
  } else {
    cz = 0;
    P.z = 0.0f;
  }
  int esc = 0;
  int* cc[3] = {&cx, &cy, &cz};
  float* oo[3] = {&P.x, &P.y, &P.z};
#pragma unroll
  for (int ax = 0; ax < DIM; ++ax) {
    int c = *cc[ax];
    if (c < 0 || c >= S.gn[ax]) {
      int cl = c < 0 ? 0 : S.gn[ax] - 1;
      *oo[ax] = (float)((double)*oo[ax] + (double)(c - cl) * S.grid);
      *cc[ax] = cl;
      esc = 1;
    }
  }
  if (esc) atomicAdd(&a.w.status[2 * s], 1ull);
  const int32_t c = (cx * S.gn[1] + cy) * S.gn[2] + cz;
  a.cur.cell[p] = c;
  a.cur.pos[p] = P;
  const int32_t key = S.cbase + c;
  a.w.key[p] = key;
  return key;
}

// Exclusive scan of the cell histogram (the inclusive prefix sum of
// particle_system.py:256, stored shifted by one cell).  Three phases over
// blocks of kScanBlock*kScanItems cells; the down-sweep also clears the
// histogram for the next step.
constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    int u = __shfl_up(v, d, 64);
    if (lane >= d) v += u;
  }
  return v;
}

// Block-wide exclusive scan of one int per thread; returns the total.
__device__ __forceinline__ int block_excl_scan(int v, int& total) {
  __shared__ int wsum[kScanBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int inc = wave_incl_scan(v);
  if (lane == 63) wsum[wid] = inc;
  __syncthreads();
  int off = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < kScanBlock / 64; ++k) {
    if (k < wid) off += wsum[k];
    tot += wsum[k];
  }
  __syncthreads();
  total = tot;
  return off + inc - v;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(const int32_t* __restrict__ count,
                                                            int32_t* __restrict__ bsum, int n) {
  const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    int i = base + k;
    if (i < n) s += count[i];
  }
  int total;
  block_excl_scan(s, total);
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_bsum(int32_t* __restrict__ bsum, int nb,
                                                          int32_t* __restrict__ start, int n) {
  // One block: exclusive scan of nb <= kScanTile block sums, in place.
  int loc[kScanItems];
  const int base = threadIdx.x * kScanItems;
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    int i = base + k;
    loc[k] = i < nb ? bsum[i] : 0;
    s += loc[k];
  }
  int total;
  int off = block_excl_scan(s, total);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    int i = base + k;
    if (i < nb) bsum[i] = off;
    off += loc[k];
  }
  if (threadIdx.x == 0) start[n] = total;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_down(int32_t* __restrict__ count,
                                                          const int32_t* __restrict__ bsum,
                                                          int32_t* __restrict__ start,
                                                          uint8_t* __restrict__ cflow, int n) {
  const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  int loc[kScanItems];
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    int i = base + k;
    loc[k] = i < n ? count[i] : 0;
    s += loc[k];
  }
  int total;
  int off = block_excl_scan(s, total) + bsum[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    int i = base + k;
    if (i < n) {
      start[i] = off;
      count[i] = 0;
      cflow[i] = 0;
    }
    off += loc[k];
  }
}

__global__ __launch_bounds__(256) void k_scatter(KArgs a) {
  int s, p;
  if (!tile_thread(a, s, p)) return;
  const int32_t key = a.w.key[p];
  a.w.tmp[a.w.start[key] + a.w.slot[p]] = p;
}

// counting_sort (particle_system.py:238-252) with a deterministic slot:
// rank among the particles of the same cell by previous index (the serial
// reference order), then move the persistent state to the other buffer.
// The slots of a workgroup lie in the cells of its first and last slot;
// those cells' arrival slots are staged in LDS once, so the rank count of
// every thread reads LDS (lanes of one cell read the same words) instead
// of the texture path (the global-memory count kept the TD unit ~94 % busy).
constexpr int kRankStage = 1024;
__device__ __forceinline__ float wc_eos(const SceneDev& S, float rt);
__global__ __launch_bounds__(256) void k_rank_move(KArgs a, int move_vtmp, int move_mui,
                                                   int move_dp, int fuse_init) {
  __shared__ int32_t win[kRankStage];
  __shared__ int32_t wlo, whi;
  const int total = a.w.start[a.ccap];
  const int k0 = blockIdx.x * blockDim.x;
  if (k0 >= total) return;  // workgroup-uniform
  const int k = k0 + (int)threadIdx.x;
  const bool live = k < total;
  int p = 0, b = 0, e = 0;
  int32_t key = 0;
  if (live) {
    p = a.w.tmp[k];
    key = a.w.key[p];
    b = a.w.start[key];
    e = a.w.start[key + 1];
  }
  const int klast = min(k0 + (int)blockDim.x, total) - 1;
  if (k == k0) wlo = b;
  if (k == klast) whi = e;
  __syncthreads();
  const int lo = wlo, nw = whi - lo;
  const bool staged = nw <= kRankStage;  // workgroup-uniform
  if (staged) {
    for (int q = (int)threadIdx.x; q < nw; q += (int)blockDim.x) win[q] = a.w.tmp[lo + q];
    __syncthreads();
  }
  if (!live) return;
  int rank = 0;
  if (staged) {
    for (int q = b; q < e; ++q) rank += (win[q - lo] < p) ? 1 : 0;
  } else {
    for (int q = b; q < e; ++q) rank += (a.w.tmp[q] < p) ? 1 : 0;
  }
  const int s = a.w.sid[p];
  const SceneDev S = a.scenes[s];
  const int d = S.pbase + (b - a.w.start[S.cbase]) + rank;
  a.alt.cell[d] = key - S.cbase;
  if (is_flow(a.cur.type[p])) a.w.cflow[key] = 1;  // same value from every writer
  a.alt.pos[d] = a.cur.pos[p];
  a.alt.vel[d] = a.cur.vel[p];
  a.alt.mass[d] = a.cur.mass[p];
  a.alt.type[d] = a.cur.type[p];
  a.alt.id0[d] = a.cur.id0[p];
  a.alt.obj[d] = a.cur.obj[p];
  if (move_vtmp) a.alt.vtmp[d] = a.cur.vtmp[p];
  const int st = a.stride;
  if (move_mui) {
#pragma unroll
    for (int c = 0; c < 6; ++c) a.alt.sig_f[c * st + d] = a.cur.sig_f[c * st + p];
    a.alt.seq[d] = a.cur.seq[p];
  }
  if (move_dp) {
#pragma unroll
    for (int c = 0; c < 6; ++c) a.alt.sig_d[c * st + d] = a.cur.sig_d[c * st + p];
    a.alt.seq[d] = a.cur.seq[p];
    a.alt.seqp[d] = a.cur.seqp[p];
    a.alt.flag[d] = a.cur.flag[p];
  }
  if (fuse_init) {
    // init_real2tmp (solver_sph_base.py:67-74) at the new position, and the
    // Tait pressure the first one_step starts with (solver_sph_wc.py:88).
    const int t = a.cur.type[p];
    if (is_real(t)) {
      const float4 v = a.cur.vel[p];
      a.alt.vtmp[d] = v;
      if (S.method == kWC && t == kFluid) {
        const float pv = wc_eos(S, v.w);
        a.w.p[d] = pv;
        a.w.vp[d] = make_float4(v.x, v.y, v.z, pv / (v.w * v.w));
      }
    }
    if (t == kSoil && S.method == kMUI) {
#pragma unroll
      for (int c = 0; c < 6; ++c) a.w.sigt_f[c * st + d] = a.cur.sig_f[c * st + p];
    } else if (t == kSoil && S.method == kDP) {
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        const double v = a.cur.sig_d[c * st + p];
        a.w.sigt_d[c * st + d] = v;
        a.w.sigt_f[c * st + d] = (float)v;
      }
    }
  }
}

// ============================================================ pair helpers
struct Centre {
  int cx, cy, cz;
  float4 P;
  int32_t shift;
};
__device__ __forceinline__ Centre centre_of(const KArgs& a, const SceneDev& S, int p) {
  Centre c;
  decode_cell(S, a.cur.cell[p], c.cx, c.cy, c.cz);
  c.P = a.cur.pos[p];
  c.shift = scene_shift(a, S);
  return c;
}
// Walk the stencil around particle p's own (renormalised) cell.
template <int DIM, class Fn>
__device__ __forceinline__ void walk_c(const SceneDev& S, const KArgs& a, int p, const Centre& C,
                                       Fn&& fn) {
  walk<DIM>(S, a.w.start, a.cur.pos, C.shift, p, C.cx, C.cy, C.cz, C.cx, C.cy, C.cz, C.P.x, C.P.y,
            C.P.z, fn);
}

// Neighbours of p: the stored list when it is valid for this step, else the
// stencil walk.  Both visit the same pairs in the same order.
template <int DIM, bool NL, class Fn>
__device__ __forceinline__ void neighbours(const SceneDev& S, const KArgs& a, int p,
                                           const Centre& C, Fn&& fn) {
  if (NL) {
    const int n = a.w.nlc[p];
    if (n <= a.nl_cap) {
      for_list<DIM>(S, a.w.nl, n, a.nl_cap, a.cur.pos, p, C.P.x, C.P.y, C.P.z, fn);
      return;
    }
  }
  walk_c<DIM>(S, a, p, C, fn);
}

// Appends of one particle's list, four entries per 16-byte store (a
// lane's entries k..k+3, k % 4 == 0, are contiguous and 16-byte aligned):
// a quarter of the write transactions of one store per entry.
struct ListAppender {
  uint32_t* __restrict__ li;
  int cap, n = 0;
  // the last three entries, oldest first (a shift register: no per-lane
  // compare picks the slot)
  uint32_t b0 = 0, b1 = 0, b2 = 0;
  __device__ __forceinline__ void add(uint32_t e) {
    // entries n-3..n go to nl_at(n - 3) = (n >> 4) * 1024 + (n & 12): a
    // 32-bit byte offset (lists stay far below 4 GiB per particle)
    if ((n & 3) == 3 && n < cap) {
      const uint32_t off = (((uint32_t)n >> 4) << 12) | (((uint32_t)n << 2) & 48u);
      *reinterpret_cast<uint4*>(reinterpret_cast<char*>(li) + off) = make_uint4(b0, b1, b2, e);
    }
    b0 = b1;
    b1 = b2;
    b2 = e;
    ++n;
  }
  // the last 1-3 entries
  __device__ __forceinline__ int finish() {
    const int r = n & 3;
    if (n <= cap && r > 0) {
      const int k = n - r;
      if (r == 3) li[nl_at(k)] = b0;
      if (r >= 2) li[nl_at(n - 2)] = b1;
      li[nl_at(n - 1)] = b2;
    }
    return n;
  }
};

// neighbours() with the scene's kernel type as a compile-time tag: fn gets
// KTag<1> (Wendland C2) or KTag<0> (cubic) first, so its kernel values
// come from kernel_W_t / kernel_gradfac_t (straight-line code).
template <int K>
struct KTag {
  static constexpr int v = K;
};
template <int DIM, bool NL, class Fn>
__device__ __forceinline__ void neighbours_k(const SceneDev& S, const KArgs& a, int p,
                                             const Centre& C, Fn&& fn) {
  if (S.kernel == 1)
    neighbours<DIM, NL>(S, a, p, C, [&](int j, float x, float y, float z, float r2, float mVj,
                                        int code) { fn(KTag<1>{}, j, x, y, z, r2, mVj, code); });
  else
    neighbours<DIM, NL>(S, a, p, C, [&](int j, float x, float y, float z, float r2, float mVj,
                                        int code) { fn(KTag<0>{}, j, x, y, z, r2, mVj, code); });
}

// Whether any of the 3^DIM stencil cells around (cx, cy, cz) holds a flow
// particle (per-cell flags set by the last sort).  Stops at the first one:
// cheaper than loading all flags together (measured).
template <int DIM>
__device__ __forceinline__ bool flow_in_reach(const SceneDev& S, const uint8_t* __restrict__ cflow,
                                              int cx, int cy, int cz) {
  const int gx = S.gn[0], gy = S.gn[1], gz = S.gn[2];
  for (int dx = -1; dx <= 1; ++dx) {
    const int nx = cx + dx;
    if (nx < 0 || nx >= gx) continue;
    for (int dy = -1; dy <= 1; ++dy) {
      const int ny = cy + dy;
      if (ny < 0 || ny >= gy) continue;
      for (int dz = (DIM == 3 ? -1 : 0); dz <= (DIM == 3 ? 1 : 0); ++dz) {
        const int nz = cz + dz;
        if (nz < 0 || nz >= gz) continue;
        if (cflow[S.cbase + (nx * gy + ny) * gz + nz]) return true;
      }
    }
  }
  return false;
}

// Neighbour-list build: one stencil walk per particle per step, recording
// (j << 6 | stencil code) for every neighbour; boundary particles keep only
// their flow neighbours (the only ones their passes read).  The candidate
// loop does nothing else, so the rarely taken append is the only divergent
// work in it.
template <int DIM>
__global__ __launch_bounds__(256) void k_nl_build(KArgs a, int use_cflow, int cull) {
  int s, p;
  if (!tile_thread(a, s, p)) return;
  const SceneDev S = a.scenes[s];
  const Centre C = centre_of(a, S, p);
  const int8_t* __restrict__ type = a.cur.type;
  // Boundary particles record flow neighbours only; rigid ones flow and
  // rigid neighbours (the XSPH sum of a rigid particle runs over rigid
  // ones); flow particles record all.
  const bool bdy = is_bdy(type[p]), rig = type[p] == kRigid;
  const int cap = a.nl_cap;
  uint32_t* __restrict__ li = a.w.nl + nl_base(p, cap);
  // A boundary particle only records flow neighbours: none can exist if
  // no stencil cell holds a flow particle.
  if (bdy && use_cflow && !flow_in_reach<DIM>(S, a.w.cflow, C.cx, C.cy, C.cz)) {
    a.w.nlc[p] = 0;
    return;
  }
  ListAppender ap{li, cap};
  auto rec = [&](int j, float x, float y, float z, float r2, float mVj, int code) {
    if ((bdy && !is_flow(type[j])) || (rig && !(is_flow(type[j]) || type[j] == kRigid))) return;
    ap.add(nl_entry(j, code));
  };
  if (cull)
    walk<DIM, true>(S, a.w.start, a.cur.pos, C.shift, p, C.cx, C.cy, C.cz, C.cx, C.cy, C.cz, C.P.x,
                    C.P.y, C.P.z, rec);
  else
    walk_c<DIM>(S, a, p, C, rec);
  a.w.nlc[p] = ap.finish();
}

// ------------------------------------------------ LDS-staged list build
// The same list build with the candidates read from LDS.  One wave per
// workgroup owns one tile.  Cell keys run (row, fast) with the fast axis z
// in 3D and y in 2D; a tile's particles usually share one row, so for each
// row offset (dx, dy in 3D; dx in 2D) the neighbour cells of all its lanes
// are consecutive keys and their particles one contiguous range.  The wave
// copies that range (coalesced, at most STAGE particles: the cell tile) into
// LDS and every lane walks its own three fast-axis cells there, in the order
// of walk() (same pairs, same order, same values).  A tile spanning two
// rows, or a larger range, reads global memory instead.  STAGE (128, 256 or
// 512 candidates: 2, 4 or 8 KiB of LDS per wave) trades how often a window
// fits against how many waves a CU holds; the engine's cell-tile setting
// picks it.
constexpr int kStageDefault = 256;

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int u = __shfl_xor(v, d, 64);
    v = u < v ? u : v;
  }
  return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int u = __shfl_xor(v, d, 64);
    v = u > v ? u : v;
  }
  return v;
}

template <int DIM, int STAGE>
__global__ __launch_bounds__(64) void k_nl_build_lds(KArgs a, int use_cflow, int cull) {
  __shared__ float4 sp[STAGE];
  const int w = __builtin_amdgcn_readfirstlane((int)xcd_block());
  const int2 t = a.tiles[w];
  if (t.x < 0) return;  // unused tile: the whole workgroup leaves
  const int s = __builtin_amdgcn_readfirstlane(t.x);
  const int lane = (int)threadIdx.x;
  const int p = t.y + lane;
  const SceneDev S = a.scenes[s];
  const int8_t* __restrict__ type = a.cur.type;
  const float4* __restrict__ pos = a.cur.pos;
  const int32_t* __restrict__ start = a.w.start;
  const int cap = a.nl_cap;
  const int32_t shift = scene_shift(a, S);
  const int gx = S.gn[0], gy = S.gn[1], gz = S.gn[2];
  bool act = p < S.pbase + S.n;
  int cx = 0, cy = 0, cz = 0, key = 0;
  float4 P = make_float4(0.f, 0.f, 0.f, 0.f);
  bool bdy = false, rig = false;
  if (act) {
    const int32_t c = a.cur.cell[p];
    decode_cell(S, c, cx, cy, cz);
    key = S.cbase + c;
    P = pos[p];
    bdy = is_bdy(type[p]);
    rig = type[p] == kRigid;
  }
  if (act && bdy && use_cflow && !flow_in_reach<DIM>(S, a.w.cflow, cx, cy, cz)) {
    // as in k_nl_build: no flow cell in reach, no neighbours to record
    a.w.nlc[p] = 0;
    act = false;
  }
  uint32_t* __restrict__ li = a.w.nl + nl_base(act ? p : t.y, cap);
  ListAppender ap{li, cap};
  auto rec = [&](int j, float x, float y, float z, float r2, float mVj, int code) {
    if ((bdy && !is_flow(type[j])) || (rig && !(is_flow(type[j]) || type[j] == kRigid))) return;
    ap.add(nl_entry(j, code));
  };
  const float wmin = bdy ? 2.0f : (rig ? 1.0f : 0.0f);  // staged-record filter
  const int row = DIM == 3 ? cx * gy + cy : cx;
  const int rmin = wave_min(act ? row : INT32_MAX), rmax = wave_max(act ? row : INT32_MIN);
  if (rmin > rmax) return;  // no lane left to walk (wave-uniform)
  if (rmin != rmax) {
    // two rows: each lane walks global memory
    if (act) {
      if (cull) walk<DIM, true>(S, start, pos, shift, p, cx, cy, cz, cx, cy, cz, P.x, P.y, P.z, rec);
      else walk<DIM>(S, start, pos, shift, p, cx, cy, cz, cx, cy, cz, P.x, P.y, P.z, rec);
      a.w.nlc[p] = ap.finish();
    }
    return;
  }
  // One row: rx (and ry in 3D) are the same for every active lane.  Walk
  // order dx, dy, dz as in walk(); the row offsets (dx, dy in 3D; dx in 2D)
  // are wave-uniform, and for each the three fast-axis neighbour cells of
  // all lanes are one contiguous particle range, staged once.
  const int rx = DIM == 3 ? rmin / gy : rmin;
  const int ry = DIM == 3 ? rmin - rx * gy : 0;
  const int f = DIM == 3 ? cz : cy;  // fast coordinate of the lane's cell
  const int gf = DIM == 3 ? gz : gy;
  const int flo = f > 0 ? -1 : 0, fhi = f < gf - 1 ? 1 : 0;
#pragma unroll 1
  for (int dx = -1; dx <= 1; ++dx) {
    if (rx + dx < 0 || rx + dx >= gx) continue;
    const float bx = P.x + (float)(-dx) * S.g;
    const float ax = dx < 0 ? P.x : (dx > 0 ? S.g - P.x : 0.0f);
#pragma unroll 1
    for (int dy0 = (DIM == 3 ? -1 : 0); dy0 <= (DIM == 3 ? 1 : 0); ++dy0) {
      if (DIM == 3 && (ry + dy0 < 0 || ry + dy0 >= gy)) continue;
      // key of the lane's neighbour cell at fast offset 0 in this row offset
      const int rkey = key + (DIM == 3 ? (dx * gy + dy0) * gz : dx * gy);
      const int klo = wave_min(act ? rkey + flo : INT32_MAX);
      const int khi = wave_max(act ? rkey + fhi : INT32_MIN);
      const int32_t jlo = start[klo] + shift;
      const int32_t cnt = start[khi + 1] + shift - jlo;
      const bool staged = cnt <= STAGE;  // wave-uniform
      if (staged) {
        __syncthreads();  // the previous row offset's reads are done
        // w = 2 for a flow particle (the only neighbours a boundary lane
        // records), 1 for a rigid one (a rigid lane records both), 0
        // otherwise: the lane's filter is one compare, w >= wmin, and needs
        // no global load per hit
        for (int q = lane; q < cnt; q += 64) {
          float4 v = pos[jlo + q];
          const int tq = type[jlo + q];
          v.w = is_flow(tq) ? 2.0f : (tq == kRigid ? 1.0f : 0.0f);
          sp[q] = v;
        }
        __syncthreads();
      }
#pragma unroll 1
      for (int df = -1; df <= 1; ++df) {
        const int dy = DIM == 3 ? dy0 : df;
        const int dz = DIM == 3 ? df : 0;
        bool ok = act && f + df >= 0 && f + df < gf;
        const float by = P.y + (float)(-dy) * S.g;
        const float ay = dy < 0 ? P.y : (dy > 0 ? S.g - P.y : 0.0f);
        if (ok && cull) {  // as in walk()
          const float az = DIM == 3 ? (dz < 0 ? P.z : (dz > 0 ? S.g - P.z : 0.0f)) : 0.0f;
          ok = !(fmaf(az, az, fmaf(ay, ay, ax * ax)) > S.rs2c);
        }
        if (!ok) continue;
        const float bz = DIM == 3 ? P.z + (float)(-dz) * S.g : 0.0f;
        const int nkey = rkey + df;
        const int code = ((dx + 1) << 4) | ((dy + 1) << 2) | (dz + 1);
        const int32_t jb = start[nkey] + shift, je = start[nkey + 1] + shift;
        auto cand = [&](int32_t j, const float4& pj) {
          const float x = bx - pj.x;
          const float y = by - pj.y;
          const float z = DIM == 3 ? bz - pj.z : 0.0f;
          const float r2 = fmaf(z, z, fmaf(y, y, x * x));
          if (r2 < S.rs2 && j != p) rec(j, x, y, z, r2, pj.w, code);
        };
        if (staged) {
          // One induction variable, the LDS byte offset, kept opaque so
          // the compiler derives no second one from the own-cell test; the
          // particle index is formed only on a hit.  p is a candidate of the
          // own cell only (wave-uniform offset 0), so only that cell's loop
          // carries the test, and the record is read whole (one 16-byte LDS
          // read) so the flow flag costs no second read on the hit path:
          // 10 vector instructions per rejected candidate instead of 12.
          // The entry of the staged record at byte offset o (a multiple of
          // 16) is nl_entry(jlo + o / 16, code) = (o << 2) + kbase: one
          // shift-add on the hit path.
          const uint32_t oend = (uint32_t)(je - jlo) * 16u;
          const uint32_t kbase = nl_entry(jlo, code);
          const char* __restrict__ sb = reinterpret_cast<const char*>(sp);
          auto scan = [&](uint32_t skip, auto own) {
#pragma unroll 1
            for (uint32_t o = (uint32_t)(jb - jlo) * 16u; o < oend; o += 16u) {
              asm volatile("" : "+v"(o));
              float4 pj = *reinterpret_cast<const float4*>(sb + o);
              asm volatile("" : "+v"(pj.w));  // one 16-byte read per candidate
              const float x = bx - pj.x;
              const float y = by - pj.y;
              const float z = DIM == 3 ? bz - pj.z : 0.0f;
              const float r2 = fmaf(z, z, fmaf(y, y, x * x));
              if (r2 < S.rs2 && (!decltype(own)::value || o != skip) && pj.w >= wmin)
                ap.add((o << 2) + kbase);
            }
          };
          if (dx == 0 && dy == 0 && dz == 0)
            scan((uint32_t)(p - jlo) * 16u, std::true_type{});
          else
            scan(0u, std::false_type{});
        } else {
          for (int32_t j = jb; j < je; ++j) cand(j, pos[j]);
        }
      }
    }
  }
  if (act) a.w.nlc[p] = ap.finish();
}

// calc_CSPM_f (solver_sph_base.py:386-398) over the recorded lists.
template <int DIM>
__global__ __launch_bounds__(256) void k_shepard_list(KArgs a, int skip_fluid) {
  int s, p;
  if (!tile_thread(a, s, p)) return;
  if (skip_fluid && a.cur.type[p] == kFluid) return;  // formed by the fluid pass
  const SceneDev S = a.scenes[s];
  const Centre C = centre_of(a, S, p);
  const int8_t* __restrict__ type = a.cur.type;
  float sum = 0.0f;
  neighbours_k<DIM, true>(S, a, p, C, [&](auto kt, int j, float x, float y, float z, float r2,
                                                  float mVj, int) {
    // straight-line: a non-flow neighbour adds +0 (the sum's bits stay)
    const bool fl = is_flow(type[j]);
    const float w = kernel_W_t<decltype(kt)::v>(S, pair_dist(r2));
    sum += fl ? mVj * w : 0.0f;
  });
  a.w.cspm[p] = sum != 0.0f ? 1.0f / sum : 1.0f;
}

// calc_CSPM_L (solver_sph_base.py:400-423) for scenes with
// kernelCorrection 1: L = sum_{j same type} m_V_j (x_j - x_i) (x) grad W_ij
// over a flow particle's neighbours, stored inverted; every other particle
// stores I.  With kernelCorrection 2 (the MLS switch) calc_kernel_corr
// computes no matrix (:363-368) and kernel_deriv_corr's branch for it is
// empty, so the corrected gradient is the zero vector (:375-384): the
// stored matrix is 0 (the untouched field), and the passes that apply it
// get a zero gradient.
template <int DIM, bool NL>
__global__ __launch_bounds__(256) void k_cspm_L(KArgs a) {
  int s, p;
  if (!tile_thread(a, s, p)) return;
  const SceneDev S = a.scenes[s];
  if (S.kcorr == 0) return;
  const float one = S.kcorr == 1 ? 1.f : 0.f;
  float M[9] = {one, 0.f, 0.f, 0.f, one, 0.f, 0.f, 0.f, one};
  const int8_t* __restrict__ type = a.cur.type;
  const int ti = type[p];
  if (S.kcorr == 1 && is_flow(ti)) {
    const Centre C = centre_of(a, S, p);
    float L[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
    neighbours_k<DIM, NL>(S, a, p, C, [&](auto kt, int j, float x, float y, float z, float r2,
                                          float mVj, int) {
      if (type[j] != ti) return;
      const float f = kernel_gradfac_t<decltype(kt)::v>(S, pair_dist(r2));
      const float g[3] = {f * x, f * y, f * z};
      const float c[3] = {mVj * (-x), mVj * (-y), mVj * (-z)};
#pragma unroll
      for (int aa = 0; aa < 3; ++aa)
#pragma unroll
        for (int bb = 0; bb < 3; ++bb) L[aa][bb] += c[aa] * g[bb];
    });
    corr_invert<DIM>(L, S.eps, M);
  }
#pragma unroll
  for (int c = 0; c < 9; ++c) a.w.Lc[(size_t)c * a.stride + p] = M[c];
}

// =================================================================== Shepard
// calc_CSPM_f (solver_sph_base.py:386-398)
template <int DIM>
__global__ __launch_bounds__(256) void k_shepard(KArgs a, int skip_fluid) {
  int s, p;
  if (!tile_thread(a, s, p)) return;
  if (skip_fluid && a.cur.type[p] == kFluid) return;  // formed by the fluid pass
  const SceneDev S = a.scenes[s];
  const Centre C = centre_of(a, S, p);
  const int8_t* __restrict__ type = a.cur.type;
  float sum = 0.0f;
  walk_c<DIM>(S, a, p, C, [&](int j, float x, float y, float z, float r2, float mVj, int) {
    if (is_flow(type[j])) sum += mVj * kernel_W(S, pair_dist(r2));
  });
  a.w.cspm[p] = sum != 0.0f ? 1.0f / sum : 1.0f;
}

// init_real2tmp (solver_sph_base.py:67-74)
__global__ __launch_bounds__(256) void k_init_tmp(KArgs a) {
  int s, p;
  if (!tile_thread(a, s, p)) return;
  const SceneDev S = a.scenes[s];
  const int t = a.cur.type[p];
  if (is_real(t)) a.cur.vtmp[p] = a.cur.vel[p];
  if (t == kSoil) {
    const int st = a.stride;
    if (S.method == kMUI) {
#pragma unroll
      for (int c = 0; c < 6; ++c) a.w.sigt_f[c * st + p] = a.cur.sig_f[c * st + p];
    } else if (S.method == kDP) {
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        const double v = a.cur.sig_d[c * st + p];
        a.w.sigt_d[c * st + p] = v;
        a.w.sigt_f[c * st + p] = (float)v;
      }
    }
  }
}

// ===================================================================== WCSPH
__global__ __launch_bounds__(256) void k_wc_eos(KArgs a) {  // solver_sph_wc.py:86-88
  int s, p;
  if (!tile_thread(a, s, p)) return;
  if (a.cur.type[p] == kFluid) {
    const float4 vt = a.cur.vtmp[p];
    const float pv = wc_eos(a.scenes[s], vt.w);
    a.w.p[p] = pv;
    a.w.vp[p] = make_float4(vt.x, vt.y, vt.z, pv / (vt.w * vt.w));
  }
}

// Dummy extrapolation (solver_sph_wc.py:90-103; solver_sph_base.py:653-663).
// With SHEP the same walk first forms the particle's Shepard sum
// (calc_CSPM_f, solver_sph_base.py:386-398: flow neighbours, list order),
// which the fused step then leaves out of the Shepard pass: bitwise equal.
template <int DIM, bool NL, bool SHEP>
__global__ __launch_bounds__(256) void k_wc_bdy(KArgs a) {
  int s, p;
  if (!tile_thread(a, s, p)) return;
  if (!is_solid(a.cur.type[p])) return;  // boundary and rigid (solver_sph_wc.py:90-103)
  const SceneDev S = a.scenes[s];
  const int8_t* __restrict__ type = a.cur.type;
  const float4* __restrict__ vtmp = a.cur.vtmp;
  const float* __restrict__ pr = a.w.p;
  const float gy = S.grav[1];
  // Most dummy particles have no flow particle in reach (an empty list):
  // their sums stay zero, so they skip the centre and the walk and only
  // form the extrapolation below (the same values the empty walk gives).
  const float4 v = a.cur.vel[p];
  const float cs_in = SHEP ? 1.0f : a.w.cspm[p];
  float svx = 0.f, svy = 0.f, svz = 0.f, sp = 0.f, shep = 0.f;
  if (!NL || a.w.nlc[p] != 0) {
    const Centre C = centre_of(a, S, p);
    // Straight-line pair terms (a non-flow neighbour adds +0, which leaves
    // every sum's bits unchanged: the sums start at +0 and never become -0
    // in round-to-nearest), so the loads of a batch of neighbours are
    // issued together instead of one branch at a time.
    neighbours_k<DIM, NL>(S, a, p, C, [&](auto kt, int j, float x, float y, float z, float r2,
                                                    float mVj, int) {
      const bool fl = is_flow(type[j]);
      const float4 vj = vtmp[j];
      const float prj = pr[j];
      const float w = kernel_W_t<decltype(kt)::v>(S, pair_dist(r2));
      if (SHEP) shep += fl ? mVj * w : 0.0f;
      svx += fl ? mVj * vj.x * w : 0.0f;
      svy += fl ? mVj * vj.y * w : 0.0f;
      svz += fl ? mVj * vj.z * w : 0.0f;
      sp += fl ? mVj * (prj + vj.w * gy * y) * w : 0.0f;
    });
  }
  float cs;
  if (SHEP) {
    cs = shep != 0.0f ? 1.0f / shep : 1.0f;
    a.w.cspm[p] = cs;
  } else {
    cs = cs_in;
  }
  const float4 vt = make_float4(2.0f * v.x - svx * cs, 2.0f * v.y - svy * cs,
                                2.0f * v.z - svz * cs, S.rho0);
  a.cur.vtmp[p] = vt;
  const float pv = sp * cs;
  const float pc = pv > 0.0f ? pv : 0.0f;
  a.w.p[p] = pc;
  // sign bit marks the boundary particle for the fluid pass
  a.w.vp[p] = make_float4(vt.x, vt.y, vt.z,
                          __int_as_float(__float_as_int(pc / (S.rho0 * S.rho0)) | (int)0x80000000));
}

// Fluid loop B (solver_sph_wc.py:108-123; tasks solver_sph_base.py:195-203,
// solver_sph_wc.py:33-71): d_density and d_vel in one walk.  With SHEP
// the same walk also forms the particle's Shepard sum (calc_CSPM_f,
// solver_sph_base.py:386-398: flow neighbours, list order), which the fused
// step then leaves out of the Shepard pass: in WC only boundary particles
// read CSPM_f within the step, so fluid particles get it here, bitwise equal.
// The pair terms of one fluid particle, one neighbour at a time in list
// order: d_density (solver_sph_base.py:195-203), viscosity and pressure
// force (solver_sph_wc.py:33-71) and, with SHEP, the Shepard sum.  q is the
// neighbour's pair record (v_tmp, p / rho_tmp^2, sign bit set for a wall).
template <int KT, bool SHEP>
struct WcPairSum {
  float vix, viy, viz, pri, c1, c2, rho0, rr;
  float dd = 0.f, dvx = 0.f, dvy = 0.f, dvz = 0.f, shep = 0.f;
  __device__ __forceinline__ void add(const SceneDev& S, float x, float y, float z, float r2, float mVj,
                                      const float4& q) {
    const float rn = pair_dist(r2);
    const float f = kernel_gradfac_t<KT>(S, rn);
    const float gx = f * x, gy = f * y, gz = f * z;
    const bool bj = __float_as_int(q.w) < 0;  // boundary neighbour
    const float ux = vix - q.x, uy = viy - q.y, uz = viz - q.z;
    dd += (mVj * ux) * gx + (mVj * uy) * gy + (mVj * uz) * gz;
    if (SHEP) shep += bj ? 0.0f : mVj * kernel_W_t<KT>(S, rn);
    // viscosity_force: fluid and boundary neighbours (solver_sph_wc.py:33-45)
    const float vx_ = ux * x + uy * y + uz * z;
    const float mn = vx_ < 0.0f ? vx_ : 0.0f;
    const float a0 = c1 * mVj;
    const float aa = bj ? a0 * rr : a0;
    const float vis = aa * mn / (r2 + c2);
    const float pc = (-rho0 * mVj) * (pri + fabsf(q.w));
    const float sc = vis + pc;
    dvx += sc * gx;
    dvy += sc * gy;
    dvz += sc * gz;
  }
  __device__ __forceinline__ void store(const KArgs& a, const SceneDev& S, int p, float rho_i) const {
    a.w.drho[p] = dd * rho_i;
    a.w.dv[p] = make_float4(dvx + S.grav[0], dvy + S.grav[1], dvz + S.grav[2], 0.0f);
    if (SHEP) a.w.cspm[p] = shep != 0.0f ? 1.0f / shep : 1.0f;
  }
};

template <int DIM, bool NL, bool SHEP, int KT, bool KC>
__device__ __forceinline__ void wc_fluid_one(const KArgs& a, const SceneDev& S, int p) {
  const Centre C = centre_of(a, S, p);
  // A WC scene holds fluid, boundary and rigid particles only (checked at
  // upload), and rigid ones are walls to the fluid: a neighbour's pair
  // record (written by the EOS or the boundary pass, sign bit set for a
  // wall) says which it is, so the pass reads position + m_V and
  // v_tmp + p/rho_tmp^2 per neighbour and nothing else.
  const float4* __restrict__ vp = a.w.vp;
  const float4 vi = a.cur.vtmp[p];
  const float pri = vp[p].w;
  const float c1 = S.visc_c1, c2 = S.visc_c2, rho0 = S.rho0;
  const float rr = rho0 / vi.w;  // rho0 / rho_i of the boundary viscosity
  if (!KC) {
    WcPairSum<KT, SHEP> acc{vi.x, vi.y, vi.z, pri, c1, c2, rho0, rr};
    neighbours<DIM, NL>(S, a, p, C, [&](int j, float x, float y, float z, float r2, float mVj, int) {
      acc.add(S, x, y, z, r2, mVj, vp[j]);
    });
    if (S.rep) {
      // calc_rep_force_task (solver_sph_wc.py:74-76, 118-121): a second
      // sweep adding the repulsive particles' force after the pair forces
      const int8_t* __restrict__ type = a.cur.type;
      neighbours<DIM, NL>(S, a, p, C, [&](int j, float x, float y, float z, float r2, float, int) {
        if (type[j] != kRep) return;
        const float f = rep_fac(S, sqrtf(r2));
        acc.dvx += f * x;
        acc.dvy += f * y;
        acc.dvz += f * z;
      });
    }
    acc.store(a, S, p, vi.w);
    return;
  }
  float M[9];
  corr_load(a.w.Lc, a.stride, p, M);
  float dd = 0.f, dvx = 0.f, dvy = 0.f, dvz = 0.f, shep = 0.f;
  neighbours<DIM, NL>(S, a, p, C, [&](int j, float x, float y, float z, float r2, float mVj, int) {
    const float rn = pair_dist(r2);
    const float f = kernel_gradfac_t<KT>(S, rn);
    const float gx = f * x, gy = f * y, gz = f * z;
    const float4 q = vp[j];
    const bool bj = __float_as_int(q.w) < 0;  // boundary neighbour
    const float ux = vi.x - q.x, uy = vi.y - q.y, uz = vi.z - q.z;
    {  // d_density takes the corrected gradient, the forces do not
      float cg[3] = {gx, gy, gz};
      corr_apply(M, cg);
      dd += (mVj * ux) * cg[0] + (mVj * uy) * cg[1] + (mVj * uz) * cg[2];
    }
    if (SHEP) shep += bj ? 0.0f : mVj * kernel_W_t<KT>(S, rn);
    // viscosity_force: fluid and boundary neighbours (solver_sph_wc.py:33-45)
    const float vx_ = ux * x + uy * y + uz * z;
    const float mn = vx_ < 0.0f ? vx_ : 0.0f;
    const float a0 = c1 * mVj;
    const float aa = bj ? a0 * rr : a0;
    const float vis = aa * mn / (r2 + c2);
    const float pc = (-rho0 * mVj) * (pri + fabsf(q.w));
    const float sc = vis + pc;
    dvx += sc * gx;
    dvy += sc * gy;
    dvz += sc * gz;
  });
  if (S.rep) {
    // calc_rep_force_task (solver_sph_wc.py:74-76, 118-121): a second sweep
    // adding the repulsive particles' force after the pair forces
    const int8_t* __restrict__ type = a.cur.type;
    neighbours<DIM, NL>(S, a, p, C, [&](int j, float x, float y, float z, float r2, float, int) {
      if (type[j] != kRep) return;
      const float f = rep_fac(S, sqrtf(r2));
      dvx += f * x;
      dvy += f * y;
      dvz += f * z;
    });
  }
  a.w.drho[p] = dd * vi.w;
  a.w.dv[p] = make_float4(dvx + S.grav[0], dvy + S.grav[1], dvz + S.grav[2], 0.0f);
  if (SHEP) a.w.cspm[p] = shep != 0.0f ? 1.0f / shep : 1.0f;
}

// d_vel of a dynamic rigid particle j: g (loop A, solver_sph_wc.py:105-106)
// minus the pair force of every fluid neighbour i (loop B, :68-71, where
// fluid i subtracts its calc_d_vel_task term from j).  The reference
// scatters with atomic updates in no fixed order; here j gathers the same
// terms, formed from i's side (x_ij = x_i - x_j, i's density_tmp in the
// wall viscosity), in its own neighbour order.
template <int DIM, bool NL, int KT>
__device__ __forceinline__ float4 wc_rigid_force(const KArgs& a, const SceneDev& S, int p) {
  const Centre C = centre_of(a, S, p);
  const int8_t* __restrict__ type = a.cur.type;
  const float4* __restrict__ vtmp = a.cur.vtmp;
  const float4* __restrict__ vp = a.w.vp;
  const float4 vj = vtmp[p];
  const float mVj = C.P.w;
  const float prj = fabsf(vp[p].w);  // p_j / rho_j^2 of the wall record
  const float c1 = S.visc_c1, c2 = S.visc_c2, rho0 = S.rho0;
  float dvx = S.grav[0], dvy = S.grav[1], dvz = S.grav[2];
  neighbours<DIM, NL>(S, a, p, C, [&](int i, float x, float y, float z, float r2, float, int) {
    if (type[i] != kFluid) return;
    const float xi = -x, yi = -y, zi = -z;  // x_i - x_j
    const float rn = pair_dist(r2);
    const float f = kernel_gradfac_t<KT>(S, rn);
    const float gx = f * xi, gy = f * yi, gz = f * zi;
    const float4 vi = vtmp[i];
    const float ux = vi.x - vj.x, uy = vi.y - vj.y, uz = vi.z - vj.z;
    const float vx_ = ux * xi + uy * yi + uz * zi;
    const float mn = vx_ < 0.0f ? vx_ : 0.0f;
    const float rr = rho0 / vi.w;
    const float vis = ((c1 * mVj) * rr) * mn / (r2 + c2);
    const float pc = (-rho0 * mVj) * (vp[i].w + prj);
    const float sc = vis + pc;
    dvx -= sc * gx;
    dvy -= sc * gy;
    dvz -= sc * gz;
  });
  return make_float4(dvx, dvy, dvz, 0.0f);
}

template <int DIM, bool NL, bool SHEP>
__global__ __launch_bounds__(256) void k_wc_fluid(KArgs a) {
  int s, p;
  if (!tile_thread(a, s, p)) return;
  const int tp = a.cur.type[p];
  if (tp != kFluid) {
    // Rates of other particles are never written by the reference: zero;
    // a static rigid particle's d_vel is zero too (solver_sph_wc.py:125-126).
    float4 dv = make_float4(0.f, 0.f, 0.f, 0.f);
    if (tp == kRigid) {
      const SceneDev S = a.scenes[s];
      if (rigid_slot(S, tp, a.cur.obj[p]) >= 0) {
        if (S.kernel == 1) dv = wc_rigid_force<DIM, NL, 1>(a, S, p);
        else dv = wc_rigid_force<DIM, NL, 0>(a, S, p);
      }
    }
    a.w.drho[p] = 0.0f;
    a.w.dv[p] = dv;
    return;
  }
  const SceneDev S = a.scenes[s];
  if (S.kcorr != 0) {
    if (S.kernel == 1) wc_fluid_one<DIM, NL, SHEP, 1, true>(a, S, p);
    else wc_fluid_one<DIM, NL, SHEP, 0, true>(a, S, p);
  } else {
    if (S.kernel == 1) wc_fluid_one<DIM, NL, SHEP, 1, false>(a, S, p);
    else wc_fluid_one<DIM, NL, SHEP, 0, false>(a, S, p);
  }
}

// ======================================================================= muI
// Soil loop 1 (solver_sph_muI.py:200-225).
template <int DIM, bool NL>
__global__ __launch_bounds__(256) void k_mui_soil1(KArgs a) {
  int s, p;
  if (!tile_thread(a, s, p)) return;
  if (a.cur.type[p] != kSoil) {
    a.w.drho[p] = 0.0f;
    a.w.p[p] = 0.0f;
    a.w.dseq[p] = 0.0;
    return;
  }
  const SceneDev S = a.scenes[s];
  const Centre C = centre_of(a, S, p);
  const float4* __restrict__ vtmp = a.cur.vtmp;
  const float4 vi = vtmp[p];
  const bool kc = S.kcorr != 0;
  float M[9];
  if (kc) corr_load(a.w.Lc, a.stride, p, M);
  float L[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
  float dd = 0.f;
  neighbours_k<DIM, NL>(S, a, p, C, [&](auto kt, int j, float x, float y, float z, float r2,
                                                  float mVj, int) {
    const float f = kernel_gradfac_t<decltype(kt)::v>(S, pair_dist(r2));
    float gw[3] = {f * x, f * y, f * z};
    if (kc) corr_apply(M, gw);
    const float4 vj = vtmp[j];
    const float u[3] = {mVj * (vj.x - vi.x), mVj * (vj.y - vi.y), mVj * (vj.z - vi.z)};
#pragma unroll
    for (int aa = 0; aa < 3; ++aa)
#pragma unroll
      for (int bb = 0; bb < 3; ++bb) L[aa][bb] += u[aa] * gw[bb];
    dd += (mVj * (vi.x - vj.x)) * gw[0] + (mVj * (vi.y - vj.y)) * gw[1] +
          (mVj * (vi.z - vj.z)) * gw[2];
  });
  a.w.drho[p] = dd * vi.w;
  float pp = S.vsound2 * (a.cur.vel[p].w - S.rho0);
  pp = pp > 0.0f ? pp : 0.0f;
  a.w.p[p] = pp;
  float e[3][3];
  float ss = 0.f;
#pragma unroll
  for (int aa = 0; aa < 3; ++aa)
#pragma unroll
    for (int bb = 0; bb < 3; ++bb) e[aa][bb] = 0.5f * (L[aa][bb] + L[bb][aa]);
#pragma unroll
  for (int aa = 0; aa < 3; ++aa)
#pragma unroll
    for (int bb = 0; bb < 3; ++bb) ss += e[aa][bb] * e[aa][bb];
  const float en = sqrtf(0.5f * ss) + S.eps;
  const float fac = 0.0f + (S.coh + pp * S.mu) / en;
  const int st = a.stride;
#pragma unroll
  for (int aa = 0; aa < 3; ++aa)
#pragma unroll
    for (int bb = aa; bb < 3; ++bb)
      a.w.sigt_f[sidx(aa, bb) * st + p] = fac * e[aa][bb] - (aa == bb ? pp : 0.0f);
  const float tr3 = (e[0][0] + e[1][1] + e[2][2]) / 3.0f;
  float sd = 0.f;
#pragma unroll
  for (int aa = 0; aa < 3; ++aa)
#pragma unroll
    for (int bb = 0; bb < 3; ++bb) {
      const float dlt = e[aa][bb] - (aa == bb ? tr3 : 0.0f);
      sd += dlt * dlt;
    }
  a.w.dseq[p] = (double)sqrtf(sd * 2.0f / 3.0f);
}

// Dummy loop of muI (:228-242) and DP (:220-231).
template <int DIM, bool EXTRAP_DENSITY, bool NL>
__global__ __launch_bounds__(256) void k_soil_bdy(KArgs a) {
  int s, p;
  if (!tile_thread(a, s, p)) return;
  if (!is_solid(a.cur.type[p])) return;  // boundary and rigid (muI :96, DP :221)
  const SceneDev S = a.scenes[s];
  const int8_t* __restrict__ type = a.cur.type;
  const float4* __restrict__ vtmp = a.cur.vtmp;
  const float* __restrict__ sgt = a.w.sigt_f;
  const int st = a.stride;
  float sv[3] = {0.f, 0.f, 0.f}, sr = 0.f, ssg[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  // as in k_wc_bdy: an empty list skips the walk (the sums stay +0), and
  // the pair terms are straight-line (a non-flow neighbour adds +0)
  if (!NL || a.w.nlc[p] != 0) {
    const Centre C = centre_of(a, S, p);
    neighbours_k<DIM, NL>(S, a, p, C, [&](auto kt, int j, float x, float y, float z, float r2,
                                                    float mVj, int) {
      const bool fl = is_flow(type[j]);
      const float w = kernel_W_t<decltype(kt)::v>(S, pair_dist(r2));
      const float4 vj = vtmp[j];
      float sj[6];
#pragma unroll
      for (int c = 0; c < 6; ++c) sj[c] = sgt[c * st + j];
      sv[0] += fl ? mVj * vj.x * w : 0.0f;
      sv[1] += fl ? mVj * vj.y * w : 0.0f;
      sv[2] += fl ? mVj * vj.z * w : 0.0f;
      sr += fl ? mVj * vj.w * w : 0.0f;
      const float xr[3] = {x, y, z};
#pragma unroll
      for (int c = 0; c < 3; ++c) ssg[c] += fl ? mVj * (sj[c] + (vj.w * S.grav[c]) * xr[c]) * w : 0.0f;
#pragma unroll
      for (int c = 3; c < 6; ++c) ssg[c] += fl ? mVj * sj[c] * w : 0.0f;
    });
  }
  const float cs = a.w.cspm[p];
  const float4 v = a.cur.vel[p];
  float rt = S.rho0;
  if (EXTRAP_DENSITY) {
    const float r = sr * cs;
    rt = r > S.rho0 ? r : S.rho0;
  }
  a.cur.vtmp[p] =
      make_float4(2.0f * v.x - sv[0] * cs, 2.0f * v.y - sv[1] * cs, 2.0f * v.z - sv[2] * cs, rt);
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const float val = ssg[c] * cs;
    a.w.sigt_f[c * st + p] = val;
    if (!EXTRAP_DENSITY) a.w.sigt_d[c * st + p] = (double)val;
  }
}

// Acceleration from stress (muI :248-261; DP :260-270) + viscous damping
// (solver_sph_base.py:713-715).
template <int DIM>
__device__ __forceinline__ void soil_dv_pair(const SceneDev& S, const float* __restrict__ sgt,
                                             int st, int j, const float si[6], float mVj,
                                             float rtj, const float gw[3], float dv[3]) {
  const float rj2 = rtj * rtj;
  const float sc = mVj * rtj;
  float sj[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) sj[c] = sgt[c * st + j];
#pragma unroll
  for (int aa = 0; aa < 3; ++aa) {
    float acc = 0.f;
#pragma unroll
    for (int bb = 0; bb < 3; ++bb) {
      const int c = sidx(aa, bb);
      acc += (sc * (sj[c] / rj2 + si[c])) * gw[bb];
    }
    dv[aa] += acc;
  }
}

// d_vel of a dynamic rigid particle p of a soil scene: g (muI :111-112, DP
// :233-234) minus the calc_d_vel_from_stress_task term of every soil
// neighbour i (muI :38-46, DP :158-165), formed from i's side (x_ip, i's
// corrected gradient) and gathered in p's neighbour order (the reference
// subtracts them with atomic updates in no fixed order).
template <int DIM, bool NL, int KT>
__device__ __forceinline__ float4 soil_rigid_force(const KArgs& a, const SceneDev& S, int p) {
  const Centre C = centre_of(a, S, p);
  const int8_t* __restrict__ type = a.cur.type;
  const float4* __restrict__ vtmp = a.cur.vtmp;
  const float* __restrict__ sgt = a.w.sigt_f;
  const int st = a.stride;
  const float rtp = vtmp[p].w, mVp = C.P.w;
  const bool kc = S.kcorr != 0;
  float dv[3] = {S.grav[0], S.grav[1], S.grav[2]};
  neighbours<DIM, NL>(S, a, p, C, [&](int i, float x, float y, float z, float r2, float, int) {
    if (type[i] != kSoil) return;
    const float f = kernel_gradfac_t<KT>(S, pair_dist(r2));
    float gw[3] = {f * -x, f * -y, f * -z};
    if (kc) {
      float M[9];
      corr_load(a.w.Lc, st, i, M);
      corr_apply(M, gw);
    }
    const float ri2 = vtmp[i].w * vtmp[i].w;
    float si[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) si[c] = sgt[c * st + i] / ri2;
    float t[3] = {0.f, 0.f, 0.f};
    soil_dv_pair<DIM>(S, sgt, st, p, si, mVp, rtp, gw, t);
    dv[0] -= t[0];
    dv[1] -= t[1];
    dv[2] -= t[2];
  });
  return make_float4(dv[0], dv[1], dv[2], 0.0f);
}
// The non-soil branch of the soil force passes: zero rates, or the
// gathered force of a dynamic rigid particle (a static one keeps 0,
// muI :130-132, DP :272-274).
template <int DIM, bool NL>
__device__ __forceinline__ float4 soil_nonsoil_dv(const KArgs& a, int s, int p) {
  const int tp = a.cur.type[p];
  if (tp != kRigid) return make_float4(0.f, 0.f, 0.f, 0.f);
  const SceneDev S = a.scenes[s];
  if (rigid_slot(S, tp, a.cur.obj[p]) < 0) return make_float4(0.f, 0.f, 0.f, 0.f);
  return S.kernel == 1 ? soil_rigid_force<DIM, NL, 1>(a, S, p) : soil_rigid_force<DIM, NL, 0>(a, S, p);
}

__device__ __forceinline__ float4 soil_dv_finish(const SceneDev& S, const float dv[3], float4 vi) {
  const float damp = -5e-5f * sqrtf(S.young / (vi.w * S.h2));
  return make_float4(dv[0] + S.grav[0] + damp * vi.x, dv[1] + S.grav[1] + damp * vi.y,
                     dv[2] + S.grav[2] + damp * vi.z, 0.0f);
}

template <int DIM, bool NL>
__global__ __launch_bounds__(256) void k_mui_soil3(KArgs a) {
  int s, p;
  if (!tile_thread(a, s, p)) return;
  if (a.cur.type[p] != kSoil) {
    a.w.dv[p] = soil_nonsoil_dv<DIM, NL>(a, s, p);
    return;
  }
  const SceneDev S = a.scenes[s];
  const Centre C = centre_of(a, S, p);
  const float4* __restrict__ vtmp = a.cur.vtmp;
  const float* __restrict__ sgt = a.w.sigt_f;
  const int st = a.stride;
  const float4 vi = vtmp[p];
  const float ri2 = vi.w * vi.w;
  float si[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) si[c] = sgt[c * st + p] / ri2;
  float dv[3] = {0.f, 0.f, 0.f};
  const bool kc = S.kcorr != 0;
  float M[9];
  if (kc) corr_load(a.w.Lc, a.stride, p, M);
  neighbours_k<DIM, NL>(S, a, p, C, [&](auto kt, int j, float x, float y, float z, float r2,
                                                  float mVj, int) {
    const float f = kernel_gradfac_t<decltype(kt)::v>(S, pair_dist(r2));
    float gw[3] = {f * x, f * y, f * z};
    if (kc) corr_apply(M, gw);
    soil_dv_pair<DIM>(S, sgt, st, j, si, mVj, vtmp[j].w, gw, dv);
  });
  if (S.rep) {
    // calc_rep_force_task (solver_sph_muI.py:34-36, 121-123), after the
    // stress forces and before gravity and damping
    const int8_t* __restrict__ type = a.cur.type;
    neighbours<DIM, NL>(S, a, p, C, [&](int j, float x, float y, float z, float r2, float, int) {
      if (type[j] != kRep) return;
      const float f = rep_fac(S, sqrtf(r2));
      dv[0] += f * x;
      dv[1] += f * y;
      dv[2] += f * z;
    });
  }
  a.w.dv[p] = soil_dv_finish(S, dv, vi);
}

// advect_something_func, muI (:267-289), phase 1 of D2.
__global__ __launch_bounds__(256) void k_mui_adv1(KArgs a) {
  int s, p;
  if (!tile_thread(a, s, p)) return;
  if (a.cur.type[p] != kSoil) return;
  const SceneDev S = a.scenes[s];
  float4 v = a.cur.vel[p];
  v.w = v.w > S.rho0 ? v.w : S.rho0;
  a.cur.vel[p] = v;
  a.cur.pos[p].w = a.cur.mass[p] / v.w;
  a.cur.seq[p] = (double)((float)a.cur.seq[p] + S.dt * (float)a.w.dseq[p]);
  const int st = a.stride;
#pragma unroll
  for (int c = 0; c < 6; ++c) a.cur.sig_f[c * st + p] = a.w.sigt_f[c * st + p];
}

// Phase 2 of D2: stress regularisation from the post-advect positions; the
// stencil centre is the cell of the new position, the pair vectors use the
// stored (pre-advect) cells.  The result is parked in the other state
// buffer's stress, which is dead until the next counting sort refills it.
template <int DIM>
__global__ __launch_bounds__(256) void k_mui_regu(KArgs a) {
  int s, p;
  if (!tile_thread(a, s, p)) return;
  if (a.cur.type[p] != kSoil) return;
  const SceneDev S = a.scenes[s];
  const Centre C = centre_of(a, S, p);
  int sx = C.cx, sy = C.cy, sz = C.cz;
  float tx = C.P.x, ty = C.P.y, tz = C.P.z;
  renorm(sx, tx, S.grid);
  renorm(sy, ty, S.grid);
  if (DIM == 3) renorm(sz, tz, S.grid);
  const int8_t* __restrict__ type = a.cur.type;
  const float* __restrict__ sg = a.cur.sig_f;
  const int st = a.stride;
  const int ti = type[p];
  float acc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  walk<DIM>(S, a.w.start, a.cur.pos, C.shift, p, sx, sy, sz, C.cx, C.cy, C.cz, C.P.x, C.P.y,
            C.P.z, [&](int j, float x, float y, float z, float r2, float mVj, int) {
              if (type[j] != ti) return;
              const float w = kernel_W(S, pair_dist(r2));
#pragma unroll
              for (int c = 0; c < 6; ++c) acc[c] += mVj * sg[c * st + j] * w;
            });
  const float cs = a.w.cspm[p];
#pragma unroll
  for (int c = 0; c < 6; ++c) a.alt.sig_f[c * st + p] = acc[c] * cs;
}

__global__ __launch_bounds__(256) void k_mui_adv2(KArgs a) {
  int s, p;
  if (!tile_thread(a, s, p)) return;
  if (a.cur.type[p] != kSoil) return;
  const int st = a.stride;
#pragma unroll
  for (int c = 0; c < 6; ++c) a.cur.sig_f[c * st + p] = a.alt.sig_f[c * st + p];
}

// ======================================================================== DP
struct Inv {
  double s[6], I1, sJ2, f;
};
__device__ __forceinline__ Inv dp_invariants(const SceneDev& S, const double st[6]) {
  Inv v;
  v.I1 = st[0] + st[1] + st[2];
  const double m = v.I1 / 3.0;
#pragma unroll
  for (int c = 0; c < 6; ++c) v.s[c] = st[c] - (c < 3 ? m : 0.0);
  double ss = 0.0;
#pragma unroll
  for (int aa = 0; aa < 3; ++aa)
#pragma unroll
    for (int bb = 0; bb < 3; ++bb) ss += v.s[sidx(aa, bb)] * v.s[sidx(aa, bb)];
  v.sJ2 = sqrt(0.5 * ss);
  v.f = v.sJ2 + S.alpha * v.I1 - S.kc;
  return v;
}
// adapt_stress (solver_sph_dp.py:73-96), FP64.
__device__ __forceinline__ void dp_adapt(const SceneDev& S, double st[6]) {
  Inv v = dp_invariants(S, st);
  if (v.f > S.eps_f) {
    if (v.f > v.sJ2) {
      const double t = (v.I1 - S.kc / S.alpha) / 3.0;
#pragma unroll
      for (int c = 0; c < 3; ++c) st[c] -= t;
    }
    v = dp_invariants(S, st);
    const double r = (-v.I1 * S.alpha + S.kc) / v.sJ2;
#pragma unroll
    for (int c = 0; c < 6; ++c) st[c] = r * v.s[c] + (c < 3 ? v.I1 / 3.0 : 0.0);
  }
}
__device__ __forceinline__ int dp_flag(const SceneDev& S, const double st[6]) {
  const Inv v = dp_invariants(S, st);
  if (v.f < -S.eps_f) return 0;
  if (v.f > S.eps_f) return v.f >= v.sJ2 ? 3 : 2;
  return 1;
}

__global__ __launch_bounds__(256) void k_dp_adapt(KArgs a) {  // one_step loop 1
  int s, p;
  if (!tile_thread(a, s, p)) return;
  if (a.cur.type[p] != kSoil) return;
  const SceneDev S = a.scenes[s];
  const int st = a.stride;
  double t[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) t[c] = a.w.sigt_d[c * st + p];
  dp_adapt(S, t);
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    a.w.sigt_d[c * st + p] = t[c];
    a.w.sigt_f[c * st + p] = (float)t[c];
  }
}

// calc_d_stress_Bui2008 (solver_sph_dp.py:171-208), FP64.
__device__ void dp_bui2008(const SceneDev& S, const double L[3][3], const double st[6],
                           double ds_out[6], double& dseq, double& dseqp) {
  const Inv v = dp_invariants(S, st);
  double e[3][3], w[3][3], Sg[3][3], s3[3][3];
#pragma unroll
  for (int aa = 0; aa < 3; ++aa)
#pragma unroll
    for (int bb = 0; bb < 3; ++bb) {
      e[aa][bb] = 0.5 * (L[aa][bb] + L[bb][aa]);
      w[aa][bb] = 0.5 * (L[aa][bb] - L[bb][aa]);
      Sg[aa][bb] = st[sidx(aa, bb)];
      s3[aa][bb] = v.s[sidx(aa, bb)];
    }
  const double tr = e[0][0] + e[1][1] + e[2][2];
  double ed[3][3];
#pragma unroll
  for (int aa = 0; aa < 3; ++aa)
#pragma unroll
    for (int bb = 0; bb < 3; ++bb) ed[aa][bb] = e[aa][bb] - (aa == bb ? tr / 3.0 : 0.0);
  const bool plastic = v.f >= -S.eps_f && v.sJ2 > 1e-8;
  double lam = 0.0;
  if (plastic) {
    double se = 0.0;
#pragma unroll
    for (int aa = 0; aa < 3; ++aa)
#pragma unroll
      for (int bb = 0; bb < 3; ++bb) se += s3[aa][bb] * e[aa][bb];
    lam = (3.0 * S.alpha * S.K * tr + (S.G / v.sJ2) * se) /
          (27.0 * S.alpha * S.K * S.sin_psi + S.G);
  }
  double ds[3][3];
#pragma unroll
  for (int aa = 0; aa < 3; ++aa)
#pragma unroll
    for (int bb = 0; bb < 3; ++bb) {
      const double J = Sg[aa][0] * w[bb][0] + Sg[aa][1] * w[bb][1] + Sg[aa][2] * w[bb][2] +
                       Sg[0][bb] * w[aa][0] + Sg[1][bb] * w[aa][1] + Sg[2][bb] * w[aa][2];
      const double gt =
          plastic ? lam * ((aa == bb ? 9.0 * S.K * S.sin_psi : 0.0) + S.G / v.sJ2 * s3[aa][bb])
                  : 0.0;
      ds[aa][bb] = J + (2.0 * S.G * ed[aa][bb] + (aa == bb ? S.K * tr : 0.0)) - gt;
    }
#pragma unroll
  for (int aa = 0; aa < 3; ++aa)
#pragma unroll
    for (int bb = aa; bb < 3; ++bb) ds_out[sidx(aa, bb)] = ds[aa][bb];
  double sd = 0.0;
#pragma unroll
  for (int aa = 0; aa < 3; ++aa)
#pragma unroll
    for (int bb = 0; bb < 3; ++bb) sd += ed[aa][bb] * ed[aa][bb];
  dseq = sqrt(sd * 2.0 / 3.0);
  dseqp = 0.0;
  if (plastic) {
    const double gp = v.sJ2 + 3.0 * v.I1 * S.sin_psi;
    double ep[3][3];
#pragma unroll
    for (int aa = 0; aa < 3; ++aa)
#pragma unroll
      for (int bb = 0; bb < 3; ++bb)
        ep[aa][bb] = (fabs(ds[aa][bb]) > 1e-8 ? gp / ds[aa][bb] : 0.0) * lam;
    const double t3 = (ep[0][0] + ep[1][1] + ep[2][2]) / 3.0;
    double q = 0.0;
#pragma unroll
    for (int aa = 0; aa < 3; ++aa)
#pragma unroll
      for (int bb = 0; bb < 3; ++bb) {
        const double d = ep[aa][bb] - (aa == bb ? t3 : 0.0);
        q += d * d;
      }
    dseqp = sqrt(q * 2.0 / 3.0);
  }
}

// DP soil loop (solver_sph_dp.py:237-270): velocity gradient, d_density
// and the stress acceleration in one walk, then the FP64 stress rate.
template <int DIM, bool NL>
__global__ __launch_bounds__(256) void k_dp_soil(KArgs a) {
  int s, p;
  if (!tile_thread(a, s, p)) return;
  if (a.cur.type[p] != kSoil) {
    const int st = a.stride;
    a.w.drho[p] = 0.0f;
    a.w.dv[p] = soil_nonsoil_dv<DIM, NL>(a, s, p);
#pragma unroll
    for (int c = 0; c < 6; ++c) a.w.dsig_d[c * st + p] = 0.0;
    a.w.dseq[p] = 0.0;
    a.w.dseqp[p] = 0.0;
    return;
  }
  const SceneDev S = a.scenes[s];
  const Centre C = centre_of(a, S, p);
  const float4* __restrict__ vtmp = a.cur.vtmp;
  const float* __restrict__ sgt = a.w.sigt_f;
  const int st = a.stride;
  const float4 vi = vtmp[p];
  const float ri2 = vi.w * vi.w;
  float si[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) si[c] = sgt[c * st + p] / ri2;
  float L[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
  float dd = 0.f, dv[3] = {0.f, 0.f, 0.f};
  const bool kc = S.kcorr != 0;
  float M[9];
  if (kc) corr_load(a.w.Lc, a.stride, p, M);
  neighbours_k<DIM, NL>(S, a, p, C, [&](auto kt, int j, float x, float y, float z, float r2,
                                                  float mVj, int) {
    const float f = kernel_gradfac_t<decltype(kt)::v>(S, pair_dist(r2));
    float gw[3] = {f * x, f * y, f * z};
    if (kc) corr_apply(M, gw);
    const float4 vj = vtmp[j];
    const float u[3] = {mVj * (vj.x - vi.x), mVj * (vj.y - vi.y), mVj * (vj.z - vi.z)};
#pragma unroll
    for (int aa = 0; aa < 3; ++aa)
#pragma unroll
      for (int bb = 0; bb < 3; ++bb) L[aa][bb] += u[aa] * gw[bb];
    dd += (mVj * (vi.x - vj.x)) * gw[0] + (mVj * (vi.y - vj.y)) * gw[1] +
          (mVj * (vi.z - vj.z)) * gw[2];
    soil_dv_pair<DIM>(S, sgt, st, j, si, mVj, vj.w, gw, dv);
  });
  a.w.drho[p] = dd * vi.w;
  double Ld[3][3], sd[6], ds[6], dq, dqp;
#pragma unroll
  for (int aa = 0; aa < 3; ++aa)
#pragma unroll
    for (int bb = 0; bb < 3; ++bb) Ld[aa][bb] = (double)L[aa][bb];
#pragma unroll
  for (int c = 0; c < 6; ++c) sd[c] = a.w.sigt_d[c * st + p];
  dp_bui2008(S, Ld, sd, ds, dq, dqp);
#pragma unroll
  for (int c = 0; c < 6; ++c) a.w.dsig_d[c * st + p] = ds[c];
  a.w.dseq[p] = dq;
  a.w.dseqp[p] = dqp;
  a.w.dv[p] = soil_dv_finish(S, dv, vi);
}

// advect_something_func, DP (:276-296).
__global__ __launch_bounds__(256) void k_dp_adv(KArgs a) {
  int s, p;
  if (!tile_thread(a, s, p)) return;
  if (a.cur.type[p] != kSoil) return;
  const SceneDev S = a.scenes[s];
  float4 v = a.cur.vel[p];
  v.w = v.w > S.rho0 ? v.w : S.rho0;
  a.cur.vel[p] = v;
  a.cur.pos[p].w = a.cur.mass[p] / v.w;
  const int st = a.stride;
  double t[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) t[c] = a.cur.sig_d[c * st + p];
  a.cur.flag[p] = dp_flag(S, t);
  dp_adapt(S, t);
#pragma unroll
  for (int c = 0; c < 6; ++c) a.cur.sig_d[c * st + p] = t[c];
  a.cur.seq[p] += S.dt_d * a.w.dseq[p];
  a.cur.seqp[p] += S.dt_d * a.w.dseqp[p];
}

// =============================================================== integrators
// advect_LF_half (solver_sph_base.py:96-104).
__global__ __launch_bounds__(256) void k_lf_half(KArgs a, int eos) {
  int s, p;
  if (!tile_thread(a, s, p)) return;
  const SceneDev S = a.scenes[s];
  const int t = a.cur.type[p];
  if (is_real(t)) {
    float4 vt = a.cur.vtmp[p];
    const float4 d = a.w.dv[p];
    vt.w += S.half_dt * a.w.drho[p];
    a.cur.pos[p].w = a.cur.mass[p] / vt.w;
    vt.x += S.half_dt * d.x;
    vt.y += S.half_dt * d.y;
    vt.z += S.half_dt * d.z;
    a.cur.vtmp[p] = vt;
    // the EOS of the next one_step (solver_sph_wc.py:88), fused
    if (eos && S.method == kWC && t == kFluid) {
      const float pv = wc_eos(S, vt.w);
      a.w.p[p] = pv;
      a.w.vp[p] = make_float4(vt.x, vt.y, vt.z, pv / (vt.w * vt.w));
    }
  }
  if (t == kSoil && S.method == kDP) {
    const int st = a.stride;
    const double h = 0.5 * S.dt_d;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const double v = a.w.sigt_d[c * st + p] + h * a.w.dsig_d[c * st + p];
      a.w.sigt_d[c * st + p] = v;
      a.w.sigt_f[c * st + p] = (float)v;
    }
  }
}

// advect_LF (:106-114); advect_SE (:79-87) is the same update.
__global__ __launch_bounds__(256) void k_lf_full(KArgs a) {
  int s, p;
  if (!tile_thread(a, s, p)) return;
  const SceneDev S = a.scenes[s];
  const int t = a.cur.type[p];
  if (is_real(t)) {
    float4 v = a.cur.vel[p];
    const float4 d = a.w.dv[p];
    v.w += S.dt * a.w.drho[p];
    a.cur.pos[p].w = a.cur.mass[p] / v.w;
    v.x += S.dt * d.x;
    v.y += S.dt * d.y;
    v.z += S.dt * d.z;
    a.cur.vel[p] = v;
    if (isnan(v.x) || isnan(v.y) || isnan(v.z) || isnan(v.w)) atomicAdd(&a.w.status[2 * s + 1], 1ull);
  }
  if (t == kSoil && S.method == kDP) {
    const int st = a.stride;
#pragma unroll
    for (int c = 0; c < 6; ++c) a.cur.sig_d[c * st + p] += S.dt_d * a.w.dsig_d[c * st + p];
  }
}

// End of step, fused: advect_LF (unless already run, XSPH case) ->
// advect_pos -> the elementwise part of advect_something (WC chk_density,
// DP return mapping, muI phase 1).  Same per-particle arithmetic as the
// separate passes.
__global__ __launch_bounds__(256) void k_finish(KArgs a, int do_lf, int use_xsph) {
  int s, p;
  if (!tile_thread(a, s, p)) return;
  const SceneDev S = a.scenes[s];
  const int t = a.cur.type[p];
  const int st = a.stride;
  float4 v = a.cur.vel[p];
  float mv = a.cur.pos[p].w;
  if (do_lf) {
    if (is_real(t)) {
      const float4 d = a.w.dv[p];
      v.w += S.dt * a.w.drho[p];
      mv = a.cur.mass[p] / v.w;
      v.x += S.dt * d.x;
      v.y += S.dt * d.y;
      v.z += S.dt * d.z;
      if (isnan(v.x) || isnan(v.y) || isnan(v.z) || isnan(v.w)) atomicAdd(&a.w.status[2 * s + 1], 1ull);
    }
    if (t == kSoil && S.method == kDP) {
#pragma unroll
      for (int c = 0; c < 6; ++c) a.cur.sig_d[c * st + p] += S.dt_d * a.w.dsig_d[c * st + p];
    }
  }
  float4 P = a.cur.pos[p];
  if (is_real(t)) {
    if (use_xsph && S.xsph) {
      const float4 xs = a.w.xs[p];
      P.x += S.dt * (v.x + 0.5f * xs.x);
      P.y += S.dt * (v.y + 0.5f * xs.y);
      P.z += S.dt * (v.z + 0.5f * xs.z);
    } else {
      P.x += S.dt * v.x;
      P.y += S.dt * v.y;
      P.z += S.dt * v.z;
    }
    if (S.dim == 2) P.z = 0.0f;
  }
  if ((S.method == kWC && t == kFluid) || (S.method != kWC && t == kSoil)) {
    v.w = v.w > S.rho0 ? v.w : S.rho0;
    mv = a.cur.mass[p] / v.w;
  }
  P.w = mv;
  a.cur.vel[p] = v;
  a.cur.pos[p] = P;
  if (t != kSoil) return;
  if (S.method == kMUI) {
    a.cur.seq[p] = (double)((float)a.cur.seq[p] + S.dt * (float)a.w.dseq[p]);
#pragma unroll
    for (int c = 0; c < 6; ++c) a.cur.sig_f[c * st + p] = a.w.sigt_f[c * st + p];
  } else if (S.method == kDP) {
    double q[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) q[c] = a.cur.sig_d[c * st + p];
    a.cur.flag[p] = dp_flag(S, q);
    dp_adapt(S, q);
#pragma unroll
    for (int c = 0; c < 6; ++c) a.cur.sig_d[c * st + p] = q[c];
    a.cur.seq[p] += S.dt_d * a.w.dseq[p];
    a.cur.seqp[p] += S.dt_d * a.w.dseqp[p];
  }
}

// XSPH sum (solver_sph_base.py:223-226), from pre-advect positions (D3).
template <int DIM, bool NL>
__global__ __launch_bounds__(256) void k_xsph(KArgs a) {
  int s, p;
  if (!tile_thread(a, s, p)) return;
  const int ti = a.cur.type[p];
  if (!is_real(ti)) return;
  const SceneDev S = a.scenes[s];
  const Centre C = centre_of(a, S, p);
  const int8_t* __restrict__ type = a.cur.type;
  const float4* __restrict__ vel = a.cur.vel;
  const float4 vi = vel[p];
  float sx = 0.f, sy = 0.f, sz = 0.f;
  neighbours_k<DIM, NL>(S, a, p, C, [&](auto kt, int j, float x, float y, float z, float r2,
                                                  float mVj, int) {
    if (type[j] != ti) return;
    const float w = kernel_W_t<decltype(kt)::v>(S, pair_dist(r2));
    const float4 vj = vel[j];
    sx += mVj * (vj.x - vi.x) * w;
    sy += mVj * (vj.y - vi.y) * w;
    sz += mVj * (vj.z - vi.z) * w;
  });
  a.w.xs[p] = make_float4(sx, sy, sz, 0.f);
}

// advect_pos (solver_sph_base.py:228-238): offsets move, cells are
// re-derived by the next hash.
__global__ __launch_bounds__(256) void k_advect_pos(KArgs a, int use_xsph) {
  int s, p;
  if (!tile_thread(a, s, p)) return;
  if (!is_real(a.cur.type[p])) return;
  const SceneDev S = a.scenes[s];
  const float4 v = a.cur.vel[p];
  float4 P = a.cur.pos[p];
  if (use_xsph && S.xsph) {
    const float4 xs = a.w.xs[p];
    P.x += S.dt * (v.x + 0.5f * xs.x);
    P.y += S.dt * (v.y + 0.5f * xs.y);
    P.z += S.dt * (v.z + 0.5f * xs.z);
  } else {
    P.x += S.dt * v.x;
    P.y += S.dt * v.y;
    P.z += S.dt * v.z;
  }
  if (S.dim == 2) P.z = 0.0f;
  a.cur.pos[p] = P;
}

// chk_density for fluid (WC advect_something, solver_sph_wc.py:129-133).
__global__ __launch_bounds__(256) void k_chk_density(KArgs a) {
  int s, p;
  if (!tile_thread(a, s, p)) return;
  if (a.cur.type[p] != kFluid) return;
  const SceneDev S = a.scenes[s];
  float4 v = a.cur.vel[p];
  v.w = v.w > S.rho0 ? v.w : S.rho0;
  a.cur.vel[p] = v;
  a.cur.pos[p].w = a.cur.mass[p] / v.w;
}

// enforce_boundary_2D/3D (solver_sph_base.py:540-601) for boundary == 1.
__global__ __launch_bounds__(256) void k_enforce_boundary(KArgs a) {
  int s, p;
  if (!tile_thread(a, s, p)) return;
  const SceneDev S = a.scenes[s];
  // judge_enforce_bdy (solver_sph_base.py:532-538): flow particles with the
  // collision boundary, dynamic rigid particles always
  const int t = a.cur.type[p];
  const bool dyn = rigid_slot(S, t, a.cur.obj[p]) >= 0;
  if (!((S.boundary == 1 && is_flow(t)) || dyn)) return;
  int cc[3];
  decode_cell(S, a.cur.cell[p], cc[0], cc[1], cc[2]);
  float4 P = a.cur.pos[p];
  float* o[3] = {&P.x, &P.y, &P.z};
  double pos[3];
  for (int ax = 0; ax < 3; ++ax)
    pos[ax] = (ax == 2 && S.dim == 2) ? 0.0 : S.vds[ax] + (double)cc[ax] * S.grid + (double)*o[ax];
  const double dist = S.radius - 1e-8;
  double nrm[3] = {0.0, 0.0, 0.0};
  auto setp = [&](int ax, double x) { *o[ax] = (float)(x - S.vds[ax] - (double)cc[ax] * S.grid); };
  if (pos[0] > S.dend[0] - dist) { nrm[0] += 1.0; setp(0, S.dend[0] - dist); }
  if (pos[0] <= S.dstart[0] + dist) { nrm[0] -= 1.0; setp(0, S.dstart[0] + dist); }
  if (pos[1] <= S.dstart[1] + dist) { nrm[1] -= 1.0; setp(1, S.dstart[1] + dist); }
  if (S.dim == 3) {
    if (pos[2] > S.dend[2] - dist) { nrm[2] += 1.0; setp(2, S.dend[2] - dist); }
    if (pos[2] <= S.dstart[2] + dist) { nrm[2] -= 1.0; setp(2, S.dstart[2] + dist); }
  }
  a.cur.pos[p] = P;
  const double len = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
  if (len > 1e-8) {
    const float u[3] = {(float)(nrm[0] / len), (float)(nrm[1] / len), (float)(nrm[2] / len)};
    float4 v = a.cur.vel[p];
    const float k = 1.3f * (v.x * u[0] + v.y * u[1] + v.z * u[2]);
    v.x -= k * u[0];
    v.y -= k * u[1];
    v.z -= k * u[2];
    a.cur.vel[p] = v;
  }
}

// ======================================================================= RK4
// substep_RK (solver_sph_base.py:134-180).
__global__ __launch_bounds__(256) void k_rk_init(KArgs a) {
  int s, p;
  if (!tile_thread(a, s, p)) return;
  const int t = a.cur.type[p];
  if (is_real(t)) {
    a.w.drho_rk[p] = 0.f;
    a.w.dv_rk[p] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  if (t == kSoil) {
    const int st = a.stride;
    for (int c = 0; c < 6; ++c) a.w.dsig_rk[c * st + p] = 0.0;
  }
}
__global__ __launch_bounds__(256) void k_rk_update(KArgs a, int m) {
  int s, p;
  if (!tile_thread(a, s, p)) return;
  const SceneDev S = a.scenes[s];
  const int t = a.cur.type[p];
  const float mf = (float)m;
  if (is_real(t)) {
    a.w.drho_rk[p] += a.w.drho[p] * mf;
    float4 r = a.w.dv_rk[p];
    const float4 d = a.w.dv[p];
    r.x += d.x * mf;
    r.y += d.y * mf;
    r.z += d.z * mf;
    a.w.dv_rk[p] = r;
  }
  if (t == kSoil && S.method == kDP) {
    const int st = a.stride;
    for (int c = 0; c < 6; ++c) a.w.dsig_rk[c * st + p] += a.w.dsig_d[c * st + p] * m;
  }
}
__global__ __launch_bounds__(256) void k_rk_advect4(KArgs a) {
  int s, p;
  if (!tile_thread(a, s, p)) return;
  const SceneDev S = a.scenes[s];
  const int t = a.cur.type[p];
  if (is_real(t)) {
    const float4 v = a.cur.vel[p];
    const float4 d = a.w.dv[p];
    float4 vt;
    vt.w = S.half_dt * a.w.drho[p] + v.w;
    a.cur.pos[p].w = a.cur.mass[p] / vt.w;
    vt.x = S.half_dt * d.x + v.x;
    vt.y = S.half_dt * d.y + v.y;
    vt.z = S.half_dt * d.z + v.z;
    a.cur.vtmp[p] = vt;
  }
  if (t == kSoil && S.method == kDP) {
    const int st = a.stride;
    for (int c = 0; c < 6; ++c) {
      const double v = 0.5 * S.dt_d * a.w.dsig_d[c * st + p] + a.cur.sig_d[c * st + p];
      a.w.sigt_d[c * st + p] = v;
      a.w.sigt_f[c * st + p] = (float)v;
    }
  }
}
__global__ __launch_bounds__(256) void k_rk_advect(KArgs a) {
  int s, p;
  if (!tile_thread(a, s, p)) return;
  const SceneDev S = a.scenes[s];
  const int t = a.cur.type[p];
  if (is_real(t)) {
    float4 v = a.cur.vel[p];
    const float4 d = a.w.dv_rk[p];
    v.w += S.dt6 * a.w.drho_rk[p];
    a.cur.pos[p].w = a.cur.mass[p] / v.w;
    v.x += S.dt6 * d.x;
    v.y += S.dt6 * d.y;
    v.z += S.dt6 * d.z;
    a.cur.vel[p] = v;
    // the NaN status of k_lf_full, for the RK4 update of the same state
    if (isnan(v.x) || isnan(v.y) || isnan(v.z) || isnan(v.w)) atomicAdd(&a.w.status[2 * s + 1], 1ull);
  }
  if (t == kSoil && S.method == kDP) {
    const int st = a.stride;
    for (int c = 0; c < 6; ++c)
      a.cur.sig_d[c * st + p] += S.dt_d / 6.0 * a.w.dsig_rk[c * st + p];
  }
}

// Neighbour-set digest for tests: count, sum and sum of squares of the
// neighbours' id0 (for_all_neighbors membership, particle_system.py:268).
// only_slot < 0: every particle, output at its engine index; else only the
// particles of that slot, output at (index - base).
template <int DIM>
__global__ __launch_bounds__(256) void k_neighbor_stats(KArgs a, int32_t* __restrict__ cnt,
                                                        long long* __restrict__ s1,
                                                        long long* __restrict__ s2, int only_slot,
                                                        int32_t base) {
  int s, p;
  if (!tile_thread(a, s, p)) return;
  if (only_slot >= 0 && s != only_slot) return;
  const SceneDev S = a.scenes[s];
  const Centre C = centre_of(a, S, p);
  const int32_t* __restrict__ id0 = a.cur.id0;
  int n = 0;
  long long a1 = 0, a2 = 0;
  walk_c<DIM>(S, a, p, C, [&](int j, float x, float y, float z, float r2, float mVj, int) {
    const long long id = id0[j];
    ++n;
    a1 += id;
    a2 += id * id;
  });
  const int o = p - base;
  cnt[o] = n;
  s1[o] = a1;
  s2[o] = a2;
}

// Exhaustive check of pair_dist against the correctly rounded sqrtf over
// the float bit patterns [lo, hi): out[0] counts differences at arguments
// >= 2^-96 (bit pattern 0x0F800000), out[1] results outside [0, 1e-8)
// below it.  Grid-stride over a fixed grid; no memory is read.
__global__ __launch_bounds__(256) void k_check_pair_dist(uint32_t lo, uint32_t hi,
                                                         unsigned long long* __restrict__ out) {
  const uint32_t stride = gridDim.x * blockDim.x;
  unsigned long long bad = 0, small = 0;
  for (uint64_t i = (uint64_t)lo + blockIdx.x * blockDim.x + threadIdx.x; i < hi; i += stride) {
    const float x = __uint_as_float((uint32_t)i);
    const float r = pair_dist(x);
    if (i >= 0x0F800000u) bad += __float_as_uint(r) != __float_as_uint(sqrtf(x));
    else small += !(r >= 0.0f && r < 1e-8f);
  }
  if (bad) atomicAdd(&out[0], bad);
  if (small) atomicAdd(&out[1], small);
}

// =================================================================== globals
// FP64 front / potential / kinetic / mass over flow particles.  Partials per
// 64-particle tile with a fixed shuffle tree, then one block per scene sums
// the scene's tiles in a fixed order: results do not depend on where the
// scene sits in the engine or on what else is resident.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v = fmax(v, __shfl_xor(v, d, 64));
  return v;
}
__global__ __launch_bounds__(256) void k_globals_tile(KArgs a) {
  const int w = wave_id();
  const int2 t = a.tiles[w];
  if (t.x < 0) return;
  const int s = t.x;
  const SceneDev S = a.scenes[s];
  const int p = t.y + (int)(threadIdx.x & 63);
  double fr = -INFINITY, pe = 0.0, ke = 0.0, m = 0.0;
  if (p < S.pbase + S.n && is_flow(a.cur.type[p])) {
    int cc[3];
    decode_cell(S, a.cur.cell[p], cc[0], cc[1], cc[2]);
    const float4 P = a.cur.pos[p];
    const double x = S.vds[0] + (double)cc[0] * S.grid + (double)P.x;
    const double y = S.vds[1] + (double)cc[1] * S.grid + (double)P.y;
    const double z = S.dim == 3 ? S.vds[2] + (double)cc[2] * S.grid + (double)P.z : 0.0;
    const float4 v = a.cur.vel[p];
    const double mi = (double)a.cur.mass[p];
    fr = x;
    pe = mi * -((double)S.grav[0] * x + (double)S.grav[1] * y + (double)S.grav[2] * z);
    ke = 0.5 * mi * ((double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z);
    m = mi;
  }
  fr = wave_max(fr);
  pe = wave_sum(pe);
  ke = wave_sum(ke);
  m = wave_sum(m);
  if ((threadIdx.x & 63) == 0) {
    double* o = a.w.gpart + 4 * (size_t)w;
    o[0] = fr;
    o[1] = pe;
    o[2] = ke;
    o[3] = m;
  }
}
__global__ __launch_bounds__(256) void k_globals_scene(KArgs a, int nslots) {
  const int s = blockIdx.x;
  if (s >= nslots) return;
  const SceneDev S = a.scenes[s];
  if (!S.active) return;
  double fr = -INFINITY, pe = 0.0, ke = 0.0, m = 0.0;
  for (int k = threadIdx.x; k < S.n_tiles; k += blockDim.x) {
    const double* o = a.w.gpart + 4 * (size_t)(S.tile_first + k);
    fr = fmax(fr, o[0]);
    pe += o[1];
    ke += o[2];
    m += o[3];
  }
  __shared__ double red[4][256 / 64];
  fr = wave_max(fr);
  pe = wave_sum(pe);
  ke = wave_sum(ke);
  m = wave_sum(m);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wid] = fr;
    red[1][wid] = pe;
    red[2][wid] = ke;
    red[3][wid] = m;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
      fr = fmax(fr, red[0][k]);
      pe += red[1][k];
      ke += red[2][k];
      m += red[3][k];
    }
    double* o = a.w.gout + 4 * s;
    o[0] = fr;
    o[1] = pe;
    o[2] = ke;
    o[3] = m;
  }
}

// ============================================================== rigid bodies
// solve_rigid_body / solve_constraints (solver_sph_base.py:472-510): shape
// matching of every dynamic rigid object after advect_something, in FP64
// (the reference's default_fp): centre of mass cm = sum m x / sum m,
// A = sum m_V rho (x - cm) (x0 - cm0)^T, R the rotation of A's polar
// decomposition, then x += (cm + R (x0 - cm0) - x).  One pass of per-tile
// partial sums (sum m, m x, w x q^T, w q with w = m_V rho, q = x0 - cm0;
// A = sum w x q^T - cm (sum w q)^T), one block per object in a fixed order,
// then the per-particle move.
__device__ __forceinline__ void abs_position(const SceneDev& S, int32_t cell, const float4& P,
                                             double x[3]) {
  int cc[3];
  decode_cell(S, cell, cc[0], cc[1], cc[2]);
  x[0] = S.vds[0] + (double)cc[0] * S.grid + (double)P.x;
  x[1] = S.vds[1] + (double)cc[1] * S.grid + (double)P.y;
  x[2] = S.dim == 3 ? S.vds[2] + (double)cc[2] * S.grid + (double)P.z : 0.0;
}
constexpr int kRigidSums = 16;
__global__ __launch_bounds__(256) void k_rigid_tile(KArgs a) {
  const int w = wave_id();
  const int2 t = a.tiles[w];
  if (t.x < 0) return;
  const SceneDev S = a.scenes[t.x];
  if (S.nrig == 0) return;  // wave-uniform
  const int lane = (int)(threadIdx.x & 63);
  const int p = t.y + lane;
  int k = -1;
  double v[kRigidSums];
  for (int c = 0; c < kRigidSums; ++c) v[c] = 0.0;
  if (p < S.pbase + S.n) k = rigid_slot(S, a.cur.type[p], a.cur.obj[p]);
  if (k >= 0) {
    double x[3];
    const float4 P = a.cur.pos[p];
    abs_position(S, a.cur.cell[p], P, x);
    const double m = (double)a.cur.mass[p];
    const double wv = (double)P.w * (double)a.cur.vel[p].w;
    const size_t id = (size_t)S.pbase + (size_t)a.cur.id0[p];
    const int st = a.stride;
    double q[3];
    // the initial centre read from the scene table (a runtime index into
    // the local copy S would put S on the stack)
    const double* cm0 = a.scenes[t.x].rig_cm0[k];
    for (int c = 0; c < 3; ++c) q[c] = a.w.x0[(size_t)c * st + id] - cm0[c];
    v[0] = m;
    for (int c = 0; c < 3; ++c) v[1 + c] = m * x[c];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) v[4 + 3 * r + c] = wv * x[r] * q[c];
    for (int c = 0; c < 3; ++c) v[13 + c] = wv * q[c];
  }
  for (int o = 0; o < S.nrig; ++o) {
    double* out = a.w.rigpart + ((size_t)w * kMaxRigid + o) * kRigidSums;
    const bool any = __any(k == o);
    for (int c = 0; c < kRigidSums; ++c) {
      double u = (k == o) ? v[c] : 0.0;
      if (any) u = wave_sum(u);
      if (lane == 0) out[c] = any ? u : 0.0;
    }
  }
}

// Orthogonal factor of the polar decomposition A = R S.  2D (the xy block;
// R_zz = 1): the nearest rotation, (a + d, c - b) normalised.  3D: Newton's
// iteration X <- (X + X^-T) / 2 (Higham 1986) from X = A.  A singular A
// gives I.
__device__ __forceinline__ void polar_rotation(const double A[9], int dim, double R[9]) {
  for (int c = 0; c < 9; ++c) R[c] = (c % 4 == 0) ? 1.0 : 0.0;
  if (dim == 2) {
    const double s = A[0] + A[4], t = A[3] - A[1];
    const double nrm = sqrt(s * s + t * t);
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
    // cofactors: X^-T = cof(X) / det(X)
    double cf[9];
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
      const double e = fabs(xn - X[c]);
      d = e > d ? e : d;
      X[c] = xn;
    }
    if (d <= 1e-15) break;
  }
  for (int c = 0; c < 9; ++c) R[c] = X[c];
}

__global__ __launch_bounds__(256) void k_rigid_solve(KArgs a, int nslots) {
  const int s = blockIdx.x / kMaxRigid, o = blockIdx.x % kMaxRigid;
  if (s >= nslots) return;
  const SceneDev S = a.scenes[s];
  if (!S.active || o >= S.nrig) return;
  double v[kRigidSums];
  for (int c = 0; c < kRigidSums; ++c) v[c] = 0.0;
  for (int k = threadIdx.x; k < S.n_tiles; k += blockDim.x) {
    const double* in = a.w.rigpart + ((size_t)(S.tile_first + k) * kMaxRigid + o) * kRigidSums;
    for (int c = 0; c < kRigidSums; ++c) v[c] += in[c];
  }
  __shared__ double red[kRigidSums][256 / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int c = 0; c < kRigidSums; ++c) {
    v[c] = wave_sum(v[c]);
    if (lane == 0) red[c][wid] = v[c];
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int c = 0; c < kRigidSums; ++c)
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) v[c] += red[c][k];
  double cm[3], A[9], R[9];
  for (int c = 0; c < 3; ++c) cm[c] = v[1 + c] / v[0];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) A[3 * r + c] = v[4 + 3 * r + c] - cm[r] * v[13 + c];
  polar_rotation(A, S.dim, R);
  double* out = a.w.rigout + ((size_t)s * kMaxRigid + o) * 12;
  for (int c = 0; c < 3; ++c) out[c] = cm[c];
  for (int c = 0; c < 9; ++c) out[3 + c] = R[c];
}

__global__ __launch_bounds__(256) void k_rigid_apply(KArgs a) {
  int s, p;
  if (!tile_thread(a, s, p)) return;
  const SceneDev S = a.scenes[s];
  const int k = rigid_slot(S, a.cur.type[p], a.cur.obj[p]);
  if (k < 0) return;
  const double* cr = a.w.rigout + ((size_t)s * kMaxRigid + k) * 12;
  const size_t id = (size_t)S.pbase + (size_t)a.cur.id0[p];
  const int st = a.stride;
  double q[3], x[3];
  const double* cm0 = a.scenes[s].rig_cm0[k];  // as in k_rigid_tile
  for (int c = 0; c < 3; ++c) q[c] = a.w.x0[(size_t)c * st + id] - cm0[c];
  float4 P = a.cur.pos[p];
  abs_position(S, a.cur.cell[p], P, x);
  int cc[3];
  decode_cell(S, a.cur.cell[p], cc[0], cc[1], cc[2]);
  float o[3] = {P.x, P.y, P.z};
#pragma unroll
  for (int r = 0; r < 3; ++r) {  // constant indices: no stack array
    if (r >= S.dim) continue;
    const double goal = cr[r] + (cr[3 + 3 * r] * q[0] + cr[4 + 3 * r] * q[1] + cr[5 + 3 * r] * q[2]);
    const double xn = x[r] + (goal - x[r]) * 1.0;
    o[r] = (float)(xn - S.vds[r] - (double)cc[r] * S.grid);
  }
  P.x = o[0];
  P.y = o[1];
  P.z = o[2];
  a.cur.pos[p] = P;
}

// ============================================================== colour feed
// assign_value_color (solver_sph_base.py:721-789): the colour value of one
// colorTitle in FP64 (the reference's default_fp), for real particles and,
// with show_bdy, dummy particles; 0 for the rest.
__device__ __forceinline__ double norm3(double x, double y, double z) {
  return sqrt(x * x + y * y + z * z);
}
// SplitMix64 (Steele, Lea & Flood, OOPSLA 2014) of (seed, id0) -> [0, 1),
// standing in for ti.random() of title 101.
__device__ __forceinline__ double uniform01(uint32_t seed, int32_t id) {
  uint64_t z = ((uint64_t)seed << 32 | (uint32_t)id) + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}
__global__ __launch_bounds__(256) void k_color_val(KArgs a, int slot, int title, int show_bdy,
                                                   uint32_t seed) {
  int s, p;
  if (!tile_thread(a, s, p) || s != slot) return;
  const SceneDev S = a.scenes[s];
  const int t = a.cur.type[p];
  double v = 0.0;
  if (is_real(t) || (show_bdy && t == kDummy)) {
    const int st = a.stride;
    auto stress = [&](int c) -> double {
      if (a.cur.sig_d) return a.cur.sig_d[c * st + p];
      if (a.cur.sig_f) return (double)a.cur.sig_f[c * st + p];
      return 0.0;  // WC: the fluid's stress field stays zero
    };
    auto position = [&](double x[3]) {
      int cc[3];
      decode_cell(S, a.cur.cell[p], cc[0], cc[1], cc[2]);
      const float4 P = a.cur.pos[p];
      x[0] = S.vds[0] + (double)cc[0] * S.grid + (double)P.x;
      x[1] = S.vds[1] + (double)cc[1] * S.grid + (double)P.y;
      x[2] = S.dim == 3 ? S.vds[2] + (double)cc[2] * S.grid + (double)P.z : 0.0;
    };
    const float4 vel = a.cur.vel[p];
    const float4 dv = a.w.dv[p];
    double x[3];
    switch (title) {
      case 1: v = (double)a.cur.id0[p]; break;
      case 2: v = (double)vel.w; break;
      case 21: v = (double)a.w.drho[p]; break;
      case 3: v = norm3(vel.x, vel.y, vel.z); break;
      case 31: v = (double)vel.x; break;
      case 32: v = -(double)vel.y; break;
      case 33: v = (double)vel.z; break;
      case 34: v = norm3(dv.x, dv.y, dv.z); break;
      case 35: v = (double)dv.x; break;
      case 36: v = (double)dv.y; break;
      case 37: v = (double)dv.z; break;
      case 4: position(x); v = norm3(x[0], x[1], x[2]); break;
      case 41: position(x); v = x[0]; break;
      case 42: position(x); v = x[1]; break;
      case 43: position(x); v = x[2]; break;
      case 44: {
        position(x);
        const size_t k = (size_t)S.pbase + (size_t)a.cur.id0[p];
        v = norm3(x[0] - a.w.x0[k], x[1] - a.w.x0[st + k], x[2] - a.w.x0[2 * (size_t)st + k]);
        break;
      }
      case 51: v = stress(0); break;
      case 52: v = -stress(1); break;
      case 53: v = -stress(2); break;
      case 54: v = stress(3); break;
      case 55: v = stress(4); break;
      case 56: v = stress(5); break;
      case 57: v = -(stress(0) + stress(1) + stress(2)) / 3.0; break;
      case 61: v = a.cur.seq ? a.cur.seq[p] : 0.0; break;
      case 62: v = a.cur.seqp ? a.cur.seqp[p] : 0.0; break;
      case 7: v = (double)a.w.p[p]; break;
      case 8: v = a.cur.flag ? (double)a.cur.flag[p] : 0.0; break;
      case 100: v = (double)a.cur.cell[p]; break;
      case 101: v = uniform01(seed, a.cur.id0[p]); break;
      case 102: v = S.kcorr == 1 ? (double)a.w.Lc[p] : 0.0; break;
      case 103: {
        const float4 vt = a.cur.vtmp[p];
        v = norm3(vt.x, vt.y, vt.z);
        break;
      }
      default: v = 0.0;
    }
  }
  a.w.val[p] = v;
}

// v_maxmin (particle_system.py:387-397): max / min of val over flow
// particles; per-tile partials, then one block per scene, then the
// givenMax / givenMin / fixMax / fixMin rule.  max / min do not depend on the
// reduction order.
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v = fmin(v, __shfl_xor(v, d, 64));
  return v;
}
__global__ __launch_bounds__(256) void k_color_range_tile(KArgs a, int slot) {
  const int w = wave_id();
  const int2 t = a.tiles[w];
  if (t.x != slot) return;  // wave-uniform
  const SceneDev S = a.scenes[t.x];
  const int p = t.y + (int)(threadIdx.x & 63);
  double hi = -INFINITY, lo = INFINITY;
  if (p < S.pbase + S.n && is_flow(a.cur.type[p])) hi = lo = a.w.val[p];
  hi = wave_max(hi);
  lo = wave_min(lo);
  if ((threadIdx.x & 63) == 0) {
    a.w.gpart[4 * (size_t)w] = hi;
    a.w.gpart[4 * (size_t)w + 1] = lo;
  }
}
__global__ __launch_bounds__(256) void k_color_range_scene(KArgs a, int slot, double given_max,
                                                           double given_min, int fix_max,
                                                           int fix_min) {
  const SceneDev S = a.scenes[slot];
  double hi = -INFINITY, lo = INFINITY;
  for (int k = threadIdx.x; k < S.n_tiles; k += blockDim.x) {
    const double* o = a.w.gpart + 4 * (size_t)(S.tile_first + k);
    hi = fmax(hi, o[0]);
    lo = fmin(lo, o[1]);
  }
  __shared__ double red[2][256 / 64];
  hi = wave_max(hi);
  lo = wave_min(lo);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wid] = hi;
    red[1][wid] = lo;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
      hi = fmax(hi, red[0][k]);
      lo = fmin(lo, red[1][k]);
    }
    // vmax if givenmax == -1 or (vmax < givenmax and not fixmax) else givenmax
    a.w.crange[2 * slot] = (given_max == -1.0 || (hi < given_max && !fix_max)) ? hi : given_max;
    a.w.crange[2 * slot + 1] = (given_min == -1.0 || (lo > given_min && !fix_min)) ? lo : given_min;
  }
}

// set_color (particle_system.py:399-407) with the jet map of colormap.py:
// map(x) = clamp((w - |clamp(x) - c|) / w * h), w = wl below c, wr above,
// clamp to [0, 1]; (h, wl, wr, c) = (1.5, .37, .37, .75 / .5 / .25) for
// R / G / B.  FP64 as the reference's Python constants, stored as f32.
__device__ __forceinline__ double clamp01(double x) { return fmax(0.0, fmin(1.0, x)); }
__device__ __forceinline__ double jet_channel(double x, double c) {
  const double h = 1.5, wl = 0.37, wr = 0.37;
  const double w = x < c ? wl : wr;
  return clamp01((w - fabs(clamp01(x) - c)) / w * h);
}
__global__ __launch_bounds__(256) void k_set_color(KArgs a, int slot, int group, int show_bdy) {
  int s, p;
  if (!tile_thread(a, s, p) || s != slot) return;
  const SceneDev S = a.scenes[s];
  const int t = a.cur.type[p];
  const bool member = (group == 0 && is_flow(t)) || (group == 1 && is_real(t)) ||
                      (group == 2 && (is_real(t) || (show_bdy && t == kDummy)));
  float4 c;
  if (member) {
    const double vmax = a.w.crange[2 * slot], vmin = a.w.crange[2 * slot + 1];
    const double vrange1 = 1.0 / (vmax - vmin);
    const double x = (a.w.val[p] - vmin) * vrange1;
    c = make_float4((float)jet_channel(x, 0.75), (float)jet_channel(x, 0.5),
                    (float)jet_channel(x, 0.25), 0.0f);
  } else {
    c = a.w.c0[(size_t)S.pbase + (size_t)a.cur.id0[p]];
  }
  a.w.rgb[p] = c;
}

}  // namespace tsph
