// tsph_engine.hip — host side of libtisphi_hip.so: engine (a set of
// resident scenes of one method and dimension), the step sequence, hipGraph
// capture, transfers and the C ABI declared in include/tisphi.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/tisphi.h"
#include "tsph_kernels.h"

using namespace tsph;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t _e = (expr);                                                             \
    if (_e != hipSuccess)                                                               \
      return fail(TSPH_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));     \
  } while (0)

constexpr int kTileAlign = 64;

template <class T>
int dalloc(T** p, size_t count, size_t& bytes) {
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(TSPH_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  // The engine's stream does not synchronise with the null stream the
  // clear runs on: wait for it here, or a later kernel or copy on the
  // engine's stream could run before it and be overwritten by the zeros.
  e = hipMemset(*p, 0, count * sizeof(T));
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    hipFree(*p);
    *p = nullptr;
    return fail(TSPH_ERR_HIP, std::string("hipMemset: ") + hipGetErrorString(e));
  }
  bytes += count * sizeof(T);
  return TSPH_OK;
}

// First-fit allocator over [0, cap) with aligned starts.
struct RangeAlloc {
  int64_t cap = 0;
  std::map<int64_t, int64_t> used;  // start -> length
  int64_t alloc(int64_t len, int64_t align) {
    int64_t pos = 0;
    for (auto& kv : used) {
      int64_t s = (pos + align - 1) / align * align;
      if (s + len <= kv.first) {
        used[s] = len;
        return s;
      }
      pos = std::max(pos, kv.first + kv.second);
    }
    int64_t s = (pos + align - 1) / align * align;
    if (s + len > cap) return -1;
    used[s] = len;
    return s;
  }
  void release(int64_t s) { used.erase(s); }
};

// Scratch of tsph_neighbor_stats_noalloc: the per-particle count, id sum and
// id square sum of one slot of at most kStatsScratchMax particles, in a
// file-scope device array (every GPU the library runs on holds its own copy
// until the process ends), so that entry point allocates nothing.  Host
// code uses the array only through the device address
// hipGetSymbolAddress returns on each device (kept per device below), never
// through its name.
constexpr int64_t kStatsScratchMax = 262144;
constexpr size_t kStatsCountBytes = (size_t)kStatsScratchMax * 4;
constexpr size_t kStatsScratchBytes = kStatsCountBytes + 2 * (size_t)kStatsScratchMax * 8;
__device__ __align__(256) char scratch[kStatsScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device`, looked up once per device with
// that device current (the caller's current device is restored).
int scratch_on(int device, char** out) {
  if (device < 0 || device >= kMaxDevices) return fail(TSPH_ERR_INVALID_ARG, "device id out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int prev = -1;
    HIP_TRY(hipGetDevice(&prev));
    if (prev != device) HIP_TRY(hipSetDevice(device));
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (prev != device) HIP_TRY(hipSetDevice(prev));
    if (e != hipSuccess || !ptr)
      return fail(TSPH_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<char*>(ptr);
  }
  *out = g_scratch_ptr[device];
  return TSPH_OK;
}

struct Timing {
  bool on = false;
  struct Rec {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Rec> pending;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pool;
  std::map<std::string, std::pair<double, int64_t>> acc;
};

}  // namespace

// Options of the step that the host decides from the resident scenes (a
// captured graph is valid only for the options it was captured with).
struct StepFlags {
  int ti = 2;
  bool xsph = false, collide = false, kcorr = false, rigid = false, rep = false;
  bool operator==(const StepFlags& o) const {
    return ti == o.ti && xsph == o.xsph && collide == o.collide && kcorr == o.kcorr &&
           rigid == o.rigid && rep == o.rep;
  }
};

struct tsph_engine {
  int method = 1, dim = 2, max_scenes = 1, device = 0;
  int32_t cap = 0, ccap = 0, tiles_alloc = 0;
  hipStream_t stream = nullptr;
  size_t bytes = 0;
  int pair_block = 256, elem_block = 256;
  bool use_graph = true;
  bool use_nl = true;    // build per-step neighbour lists
  bool nl_lds = true;    // build them from LDS-staged cell windows
  int nl_stage = kStageDefault;  // candidates per staged window (cell tile)
  bool nl_valid = false; // lists match the current order and positions
  bool cflow_valid = false;  // per-cell flow flags match the current order
  bool fresh_sort = false;   // offsets renormalised, no motion since the sort
  int32_t nl_cap = 0;
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  bool graph_valid = false;
  StepFlags graph_flags;  // options the graphs were captured with
  int parity = 0;  // which state buffer is current
  PState buf[2];
  PWork w;
  SceneDev* d_scenes = nullptr;
  int2* d_tiles = nullptr;
  std::vector<SceneDev> scenes;
  std::vector<tsph_scene_desc> descs;
  std::vector<int64_t> steps;
  std::vector<int2> tiles;
  RangeAlloc palloc, calloc_;
  Timing timing;

  KArgs args() const {
    KArgs a;
    a.scenes = d_scenes;
    a.tiles = d_tiles;
    a.cur = buf[parity];
    a.alt = buf[parity ^ 1];
    a.w = w;
    a.cap = cap;
    a.ccap = ccap;
    a.stride = cap;
    a.nl_cap = nl_cap;
    return a;
  }
  int nblocks_tiles(int block) const { return (tiles_alloc * 64 + block - 1) / block; }
  int active_count() const {
    int k = 0;
    for (auto& s : scenes) k += s.active;
    return k;
  }
};

namespace {

int alloc_state(tsph_engine* E, PState& s) {
  size_t& b = E->bytes;
  const size_t n = E->cap, n6 = 6 * (size_t)E->cap;
  int r;
  if ((r = dalloc(&s.cell, n, b)) || (r = dalloc(&s.pos, n, b)) || (r = dalloc(&s.vel, n, b)) ||
      (r = dalloc(&s.vtmp, n, b)) || (r = dalloc(&s.mass, n, b)) || (r = dalloc(&s.type, n, b)) ||
      (r = dalloc(&s.id0, n, b)) || (r = dalloc(&s.obj, n, b)))
    return r;
  if (E->method == kMUI) {
    if ((r = dalloc(&s.sig_f, n6, b)) || (r = dalloc(&s.seq, n, b))) return r;
  }
  if (E->method == kDP) {
    if ((r = dalloc(&s.sig_d, n6, b)) || (r = dalloc(&s.seq, n, b)) || (r = dalloc(&s.seqp, n, b)) ||
        (r = dalloc(&s.flag, n, b)))
      return r;
  }
  return TSPH_OK;
}

void free_state(PState& s) {
  void* ptrs[] = {s.cell, s.pos, s.vel, s.vtmp, s.mass, s.type, s.id0, s.obj, s.sig_f, s.sig_d,
                  s.seq, s.seqp, s.flag};
  for (void* p : ptrs)
    if (p) hipFree(p);
  s = PState{};
}

int alloc_work(tsph_engine* E) {
  PWork& w = E->w;
  size_t& b = E->bytes;
  const size_t n = E->cap, n6 = 6 * (size_t)E->cap, c = E->ccap;
  const size_t nb = (c + kScanTile - 1) / kScanTile;
  int r;
  if ((r = dalloc(&w.key, n, b)) || (r = dalloc(&w.slot, n, b)) || (r = dalloc(&w.tmp, n, b)) ||
      (r = dalloc(&w.sid, n, b)) || (r = dalloc(&w.p, n, b)) || (r = dalloc(&w.vp, n, b)) || (r = dalloc(&w.Lc, 9 * n, b)) ||
      (r = dalloc(&w.cspm, n, b)) ||
      (r = dalloc(&w.drho, n, b)) || (r = dalloc(&w.dv, n, b)) || (r = dalloc(&w.xs, n, b)) ||
      (r = dalloc(&w.drho_rk, n, b)) || (r = dalloc(&w.dv_rk, n, b)) ||
      (r = dalloc(&w.count, c, b)) || (r = dalloc(&w.start, c + 1, b)) ||
      (r = dalloc(&w.cflow, c, b)) ||
      (r = dalloc(&w.bsum, std::max<size_t>(nb, 1), b)) ||
      (r = dalloc(&w.status, 2 * (size_t)E->max_scenes, b)) ||
      (r = dalloc(&w.gpart, 4 * (size_t)E->tiles_alloc, b)) ||
      (r = dalloc(&w.gout, 4 * (size_t)E->max_scenes, b)) ||
      (r = dalloc(&w.nlc, n, b)) || (r = dalloc(&w.nl, (size_t)E->nl_cap * 64 * (size_t)E->tiles_alloc, b)))
    return r;
  if (E->method != kWC) {
    if ((r = dalloc(&w.sigt_f, n6, b)) || (r = dalloc(&w.dseq, n, b))) return r;
  }
  if (E->method == kDP) {
    if ((r = dalloc(&w.sigt_d, n6, b)) || (r = dalloc(&w.dsig_d, n6, b)) ||
        (r = dalloc(&w.dseqp, n, b)) || (r = dalloc(&w.dsig_rk, n6, b)))
      return r;
  }
  return TSPH_OK;
}

void free_work(PWork& w) {
  void* ptrs[] = {w.key, w.slot, w.tmp, w.sid, w.p, w.vp, w.Lc, w.cspm, w.drho, w.dv, w.sigt_f, w.sigt_d,
                  w.dsig_d, w.dseq, w.dseqp, w.xs, w.drho_rk, w.dv_rk, w.dsig_rk, w.count,
                  w.start, w.bsum, w.status, w.gpart, w.gout, w.nl, w.nlc, w.cflow,
                  w.val, w.rgb, w.x0, w.c0, w.crange, w.rigpart, w.rigout};
  for (void* p : ptrs)
    if (p) hipFree(p);
  w = PWork{};
}

void invalidate_graph(tsph_engine* E) {
  for (auto& g : E->graph)
    if (g) {
      hipGraphExecDestroy(g);
      g = nullptr;
    }
  E->graph_valid = false;
}

// The colour feed's arrays, on first use (an engine that never colours
// does not hold them).
int ensure_colour(tsph_engine* E) {
  PWork& w = E->w;
  if (w.val) return TSPH_OK;
  size_t& b = E->bytes;
  const size_t n = E->cap;
  int r;
  if ((r = dalloc(&w.val, n, b)) || (r = dalloc(&w.rgb, n, b)) || (r = dalloc(&w.x0, 3 * n, b)) ||
      (r = dalloc(&w.c0, n, b)) || (r = dalloc(&w.crange, 2 * (size_t)E->max_scenes, b)))
    return r;
  invalidate_graph(E);  // the kernel arguments captured in the graphs changed
  return TSPH_OK;
}

// Shape-matching arrays, with the first dynamic rigid object (and the
// initial positions they read, shared with the colour feed).
int ensure_rigid(tsph_engine* E) {
  int r = ensure_colour(E);
  if (r || E->w.rigpart) return r;
  size_t& b = E->bytes;
  if ((r = dalloc(&E->w.rigpart, (size_t)E->tiles_alloc * kMaxRigid * kRigidSums, b)) ||
      (r = dalloc(&E->w.rigout, (size_t)E->max_scenes * kMaxRigid * 12, b)))
    return r;
  invalidate_graph(E);
  return TSPH_OK;
}

// ----------------------------------------------------------------- launching
struct Launcher {
  tsph_engine* E;
  int status = TSPH_OK;
  template <class K, class... Args>
  void go(const char* name, K kernel, int grid, int block, Args... args) {
    if (status != TSPH_OK || grid <= 0) return;
    Timing& T = E->timing;
    Timing::Rec rec;
    const bool timed = T.on;
    if (timed) {
      if (T.pool.empty()) {
        hipEvent_t a, b;
        hipEventCreate(&a);
        hipEventCreate(&b);
        T.pool.push_back({a, b});
      }
      rec.name = name;
      rec.a = T.pool.back().first;
      rec.b = T.pool.back().second;
      T.pool.pop_back();
      hipEventRecord(rec.a, E->stream);
    }
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, E->stream, args...);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) status = fail(TSPH_ERR_HIP, std::string(name) + ": " + hipGetErrorString(e));
    if (timed) {
      hipEventRecord(rec.b, E->stream);
      T.pending.push_back(rec);
    }
  }
};

void resolve_timing(tsph_engine* E) {
  Timing& T = E->timing;
  for (auto& r : T.pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      auto& acc = T.acc[r.name];
      acc.first += ms;
      acc.second += 1;
    }
    T.pool.push_back({r.a, r.b});
  }
  T.pending.clear();
}

StepFlags step_flags(tsph_engine* E);

template <int DIM>
struct Seq {
  tsph_engine* E;
  Launcher& L;
  int tb() const { return E->nblocks_tiles(E->elem_block); }
  int pb() const { return E->nblocks_tiles(E->pair_block); }

  // update_grid_id + prefix sum + counting_sort; with fuse_init the move
  // also performs init_real2tmp and the first EOS of the step.
  void sort(bool fuse_init = false) {
    KArgs a = E->args();
    const int eb = E->elem_block;
    L.go("hash", k_hash<DIM>, tb(), eb, a);
    const int n = E->ccap;
    const int nb = (n + kScanTile - 1) / kScanTile;
    L.go("scan", k_scan_reduce, nb, kScanBlock, (const int32_t*)E->w.count, E->w.bsum, n);
    L.go("scan", k_scan_bsum, 1, kScanBlock, E->w.bsum, nb, E->w.start, n);
    L.go("scan", k_scan_down, nb, kScanBlock, E->w.count, (const int32_t*)E->w.bsum, E->w.start,
         E->w.cflow, n);
    L.go("scatter", k_scatter, tb(), eb, a);
    L.go("rank_move", k_rank_move, (E->cap + eb - 1) / eb, eb, a, 1, E->method == kMUI ? 1 : 0,
         E->method == kDP ? 1 : 0, fuse_init ? 1 : 0);
    E->parity ^= 1;
    E->nl_valid = false;
    E->cflow_valid = true;
    E->fresh_sort = true;
  }
  // calc_CSPM_f; with neighbour lists on, the walk also records the lists.
  // skip_fluid: fluid particles get theirs in the step's first fluid pass.
  // skip_all: the WC fused step, whose first fluid and boundary passes form
  // every particle's Shepard sum themselves (a WC scene holds fluid,
  // boundary and rigid particles only; the boundary pass covers the latter
  // two).
  void shepard(bool skip_fluid = false, bool skip_all = false) {
    if (E->use_nl) {
      const int cf = E->cflow_valid ? 1 : 0, cull = E->fresh_sort ? 1 : 0;
      if (E->nl_lds) {
        const KArgs a = E->args();
        if (E->nl_stage == 128)
          L.go("nl_build", k_nl_build_lds<DIM, 128>, E->tiles_alloc, 64, a, cf, cull);
        else if (E->nl_stage == 512)
          L.go("nl_build", k_nl_build_lds<DIM, 512>, E->tiles_alloc, 64, a, cf, cull);
        else
          L.go("nl_build", k_nl_build_lds<DIM, 256>, E->tiles_alloc, 64, a, cf, cull);
      } else
        L.go("nl_build", k_nl_build<DIM>, pb(), E->pair_block, E->args(), cf, cull);
      E->nl_valid = true;
      if (!skip_all)
        L.go("shepard", k_shepard_list<DIM>, pb(), E->pair_block, E->args(), skip_fluid ? 1 : 0);
    } else if (!skip_all) {
      L.go("shepard", k_shepard<DIM>, pb(), E->pair_block, E->args(), skip_fluid ? 1 : 0);
    }
    // calc_kernel_corr (solver_sph_base.py:363-368): CSPM_L after CSPM_f
    if (step_flags(E).kcorr) {
      if (E->nl_valid) L.go("cspm_L", k_cspm_L<DIM, true>, pb(), E->pair_block, E->args());
      else L.go("cspm_L", k_cspm_L<DIM, false>, pb(), E->pair_block, E->args());
    }
  }
  void init_tmp() { L.go("init_tmp", k_init_tmp, tb(), E->elem_block, E->args()); }
  // one_step of the method; with_eos=false when the EOS was fused into the
  // kernel before (the move of the sort, or lf_half); shep: the WC fluid
  // pass also forms the fluid particles' Shepard factors.
  void one_step(bool with_eos = true, bool shep = false) {
    if (E->nl_valid) one_step_t<true>(with_eos, shep);
    else one_step_t<false>(with_eos, shep);
  }
  template <bool NL>
  void one_step_t(bool with_eos, bool shep) {
    KArgs a = E->args();
    const int B = E->pair_block;
    if (E->method == kWC) {
      if (with_eos) L.go("wc_eos", k_wc_eos, tb(), E->elem_block, a);
      if (shep) L.go("wc_bdy", k_wc_bdy<DIM, NL, true>, pb(), B, a);
      else L.go("wc_bdy", k_wc_bdy<DIM, NL, false>, pb(), B, a);
      if (shep) L.go("wc_fluid", k_wc_fluid<DIM, NL, true>, pb(), B, a);
      else L.go("wc_fluid", k_wc_fluid<DIM, NL, false>, pb(), B, a);
    } else if (E->method == kMUI) {
      L.go("mui_soil1", k_mui_soil1<DIM, NL>, pb(), B, a);
      L.go("soil_bdy", k_soil_bdy<DIM, true, NL>, pb(), B, a);
      L.go("mui_soil3", k_mui_soil3<DIM, NL>, pb(), B, a);
    } else {
      L.go("dp_adapt", k_dp_adapt, tb(), E->elem_block, a);
      L.go("soil_bdy", k_soil_bdy<DIM, false, NL>, pb(), B, a);
      L.go("dp_soil", k_dp_soil<DIM, NL>, pb(), B, a);
    }
  }
  void lf_half(bool eos = false) { L.go("lf_half", k_lf_half, tb(), E->elem_block, E->args(), eos ? 1 : 0); }
  void lf_full() { L.go("lf_full", k_lf_full, tb(), E->elem_block, E->args()); }
  void substep(int ti, bool shep_first = false) {
    if (ti == 1) {
      one_step(true, shep_first);
      lf_full();
    } else if (ti == 2) {
      one_step(true, shep_first);
      lf_half();
      one_step();
      lf_full();
    } else {
      KArgs a = E->args();
      L.go("rk", k_rk_init, tb(), E->elem_block, a);
      const int m[4] = {1, 2, 2, 1};
      for (int k = 0; k < 4; ++k) {
        one_step(true, shep_first && k == 0);
        L.go("rk", k_rk_update, tb(), E->elem_block, a, m[k]);
        if (k < 3) L.go("rk", k_rk_advect4, tb(), E->elem_block, a);
      }
      L.go("rk", k_rk_advect, tb(), E->elem_block, a);
    }
  }
  void advect_pos(bool xsph) {
    E->fresh_sort = false;
    KArgs a = E->args();
    if (xsph) {
      if (E->nl_valid) L.go("xsph", k_xsph<DIM, true>, pb(), E->pair_block, a);
      else L.go("xsph", k_xsph<DIM, false>, pb(), E->pair_block, a);
    }
    L.go("advect_pos", k_advect_pos, tb(), E->elem_block, a, xsph ? 1 : 0);
  }
  void advect_something() {
    KArgs a = E->args();
    if (E->method == kWC) {
      L.go("chk_density", k_chk_density, tb(), E->elem_block, a);
    } else if (E->method == kMUI) {
      L.go("mui_adv", k_mui_adv1, tb(), E->elem_block, a);
      L.go("mui_regu", k_mui_regu<DIM>, pb(), E->pair_block, a);
      L.go("mui_adv", k_mui_adv2, tb(), E->elem_block, a);
    } else {
      L.go("dp_adv", k_dp_adv, tb(), E->elem_block, a);
    }
  }
  void enforce_boundary() {
    E->fresh_sort = false;
    L.go("enforce_boundary", k_enforce_boundary, tb(), E->elem_block, E->args());
  }
  // solve_rigid_body (solver_sph_base.py:472-475): shape matching of the
  // dynamic rigid objects.
  void solve_rigid() {
    E->fresh_sort = false;
    KArgs a = E->args();
    L.go("rigid", k_rigid_tile, E->nblocks_tiles(256), 256, a);
    L.go("rigid", k_rigid_solve, E->max_scenes * kMaxRigid, 256, a, E->max_scenes);
    L.go("rigid", k_rigid_apply, tb(), E->elem_block, a);
  }
  // SPHBase.step (solver_sph_base.py:41-51) as the pass sequence, one
  // launch per reference kernel.
  void step_passes(int ti, bool xsph, bool collide, bool rigid) {
    sort();
    shepard();
    init_tmp();
    substep(ti);
    advect_pos(xsph);
    advect_something();
    if (rigid) solve_rigid();
    if (collide || rigid) enforce_boundary();
  }
  // The same step with elementwise passes fused into their neighbours
  // (bitwise identical results, fewer launches); all resident scenes share
  // the engine's time-integration scheme.
  void step(int ti, bool xsph, bool collide, bool rigid) {
    sort(true);
    // WC: every particle's Shepard sum is formed in the first boundary and
    // fluid passes (the boundary pass runs first and reads it right away)
    const bool shep = E->method == kWC;
    shepard(shep, shep);
    if (ti == 2) {
      one_step(false, shep);
      lf_half(true);
      one_step(false);
    } else if (ti == 1) {
      one_step(false, shep);
    } else {
      substep(4, shep);
    }
    const bool lf_in_finish = ti != 4 && !xsph;
    KArgs a = E->args();
    if (xsph) {
      if (ti != 4) lf_full();
      if (E->nl_valid) L.go("xsph", k_xsph<DIM, true>, pb(), E->pair_block, a);
      else L.go("xsph", k_xsph<DIM, false>, pb(), E->pair_block, a);
    }
    L.go("finish", k_finish, tb(), E->elem_block, a, lf_in_finish ? 1 : 0, xsph ? 1 : 0);
    E->fresh_sort = false;
    if (E->method == kMUI) {
      L.go("mui_regu", k_mui_regu<DIM>, pb(), E->pair_block, a);
      L.go("mui_adv", k_mui_adv2, tb(), E->elem_block, a);
    }
    if (rigid) solve_rigid();
    if (collide || rigid) enforce_boundary();
  }
  void globals() {
    KArgs a = E->args();
    L.go("globals", k_globals_tile, E->nblocks_tiles(256), 256, a);
    L.go("globals", k_globals_scene, E->max_scenes, 256, a, E->max_scenes);
  }
};

StepFlags step_flags(tsph_engine* E) {
  StepFlags f;
  bool first = true;
  for (size_t s = 0; s < E->scenes.size(); ++s) {
    if (!E->scenes[s].active) continue;
    if (first) f.ti = E->scenes[s].ti;
    first = false;
    f.xsph = f.xsph || E->scenes[s].xsph;
    f.collide = f.collide || E->scenes[s].boundary == 1;
    f.kcorr = f.kcorr || E->scenes[s].kcorr != 0;
    f.rigid = f.rigid || E->scenes[s].nrig > 0;
    f.rep = f.rep || E->scenes[s].rep != 0;
  }
  return f;
}

int run_step_eager(tsph_engine* E) {
  Launcher L{E};
  StepFlags f = step_flags(E);
  if (E->dim == 2) Seq<2>{E, L}.step(f.ti, f.xsph, f.collide, f.rigid);
  else Seq<3>{E, L}.step(f.ti, f.xsph, f.collide, f.rigid);
  return L.status;
}

int build_graphs(tsph_engine* E) {
  invalidate_graph(E);
  const int p0 = E->parity;
  for (int k = 0; k < 2; ++k) {
    const int par = (p0 + k) & 1;
    E->parity = par;
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(E->stream, hipStreamCaptureModeThreadLocal));
    int st = run_step_eager(E);
    hipError_t e = hipStreamEndCapture(E->stream, &g);
    if (st != TSPH_OK) {
      if (g) hipGraphDestroy(g);
      E->parity = p0;
      return st;
    }
    if (e != hipSuccess) {
      E->parity = p0;
      return fail(TSPH_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    }
    hipGraphExec_t ex = nullptr;
    e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) {
      E->parity = p0;
      return fail(TSPH_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    E->graph[par] = ex;
  }
  E->parity = p0;
  E->graph_valid = true;
  E->graph_flags = step_flags(E);
  return TSPH_OK;
}

int upload_tables(tsph_engine* E) {
  // Tiles: scenes in slot order, 64 particles each.
  std::fill(E->tiles.begin(), E->tiles.end(), make_int2(-1, 0));
  int w = 0;
  for (size_t s = 0; s < E->scenes.size(); ++s) {
    SceneDev& S = E->scenes[s];
    S.tile_first = w;
    S.n_tiles = 0;
    if (!S.active) continue;
    for (int p = 0; p < S.n; p += 64) {
      E->tiles[w++] = make_int2((int)s, S.pbase + p);
      S.n_tiles++;
    }
  }
  HIP_TRY(hipMemcpyAsync(E->d_tiles, E->tiles.data(), E->tiles.size() * sizeof(int2),
                         hipMemcpyHostToDevice, E->stream));
  HIP_TRY(hipMemcpyAsync(E->d_scenes, E->scenes.data(), E->scenes.size() * sizeof(SceneDev),
                         hipMemcpyHostToDevice, E->stream));
  HIP_TRY(hipStreamSynchronize(E->stream));
  return TSPH_OK;
}

int validate_desc(const tsph_engine* E, const tsph_scene_desc* d) {
  if (!d) return fail(TSPH_ERR_INVALID_ARG, "null scene description");
  if (d->method != E->method) return fail(TSPH_ERR_INVALID_ARG, "scene method differs from the engine's");
  // one fused step sequence per engine: its residents share the integrator
  for (const SceneDev& o : E->scenes)
    if (o.active && o.ti != d->time_integration)
      return fail(TSPH_ERR_INVALID_ARG, "scene time integration differs from the resident scenes'");
  if (d->dim != E->dim) return fail(TSPH_ERR_INVALID_ARG, "scene dimension differs from the engine's");
  if (d->boundary < 0 || d->boundary > 4) return fail(TSPH_ERR_INVALID_ARG, "boundary must be 0..4");
  if (d->kernel != 0 && d->kernel != 1) return fail(TSPH_ERR_INVALID_ARG, "kernel must be 0 or 1");
  if (d->kernel_corr < 0 || d->kernel_corr > 2)
    return fail(TSPH_ERR_INVALID_ARG, "kernelCorrection must be 0, 1 or 2");
  if (d->time_integration != 1 && d->time_integration != 2 && d->time_integration != 4)
    return fail(TSPH_ERR_UNSUPPORTED, "timeIntegration must be 1, 2 or 4");
  if (d->n_particles < 0) return fail(TSPH_ERR_INVALID_ARG, "negative particle count");
  for (int a = 0; a < 3; ++a)
    if (d->grid_num[a] < 1) return fail(TSPH_ERR_INVALID_ARG, "grid_num must be >= 1");
  if (d->dim == 2 && d->grid_num[2] != 1) return fail(TSPH_ERR_INVALID_ARG, "2D scenes need grid_num[2] == 1");
  if (!(d->grid_size > 0) || !(d->support_radius > 0) || !(d->smoothing_len > 0))
    return fail(TSPH_ERR_INVALID_ARG, "non-positive length scale");
  if (d->support_radius > d->grid_size * (1 + 1e-12))
    return fail(TSPH_ERR_INVALID_ARG, "support radius exceeds the grid cell");
  return TSPH_OK;
}

SceneDev make_scene(const tsph_engine* E, const tsph_scene_desc* d) {
  SceneDev S;
  std::memset(&S, 0, sizeof(S));
  S.active = 1;
  S.dim = d->dim;
  S.method = d->method;
  S.boundary = d->boundary;
  S.kernel = d->kernel;
  S.ti = d->time_integration;
  S.xsph = d->xsph ? 1 : 0;
  const double ie = std::nearbyint(d->exponent);
  S.iexp = (ie == d->exponent && std::fabs(ie) <= 64) ? (int)ie : INT32_MIN;
  for (int a = 0; a < 3; ++a) {
    S.gn[a] = d->grid_num[a];
    S.vds[a] = d->vdomain_start[a];
    S.dstart[a] = d->domain_start[a];
    S.dend[a] = d->domain_end[a];
    S.grav[a] = (float)d->gravity[a];
  }
  S.n = (int32_t)d->n_particles;
  S.ncell = d->grid_num[0] * d->grid_num[1] * d->grid_num[2];
  S.grid = d->grid_size;
  S.radius = d->particle_radius;
  S.dt_d = d->dt;
  S.g = (float)d->grid_size;
  const float rs = (float)d->support_radius;
  S.rs2 = rs * rs;
  S.rs2c = S.rs2 * 1.0001f;
  // repulsive layer (boundary 3 / 4): WC and muI add its force; the DP
  // solver does not (solver_sph_dp.py:262-264 leaves it out)
  S.kcorr = d->kernel_corr;
  S.rep = (d->boundary == 3 || d->boundary == 4) && d->method != kDP;
  S.rep_k = (float)(0.01 * d->vsound * d->vsound);
  S.rep_d = (float)d->particle_diameter;
  S.rep_h = (float)(0.75 * d->smoothing_len);
  const double h = d->smoothing_len, h1 = 1.0 / h;
  S.h1 = (float)h1;
  const double kw = d->dim == 2 ? 7.0 / (4.0 * M_PI) : 21.0 / (2.0 * M_PI);
  S.kW = (float)(kw * std::pow(h1, d->dim));
  const double kc = d->dim == 2 ? 15.0 / 7.0 / M_PI : 3.0 / 2.0 / M_PI;
  S.kC = (float)(kc * std::pow(h1, d->dim));
  S.kG = (float)(-5.0 * kw * std::pow(h1, d->dim) * h1 * h1);
  S.kGc = (float)(kc * std::pow(h1, d->dim) * h1 * h1);
  S.eps = (float)1e-8;
  S.dt = (float)d->dt;
  S.half_dt = (float)(0.5 * d->dt);
  S.dt6 = (float)(d->dt / 6.0);
  S.rho0 = (float)d->density0;
  S.visc_c1 = (float)(2.0 * (d->dim + 2) * d->viscosity);
  S.visc_c2 = (float)(0.01 * h * h);
  S.stiffness = (float)d->stiffness;
  S.exponent = (float)d->exponent;
  S.vsound2 = (float)(d->vsound * d->vsound);
  S.coh = (float)d->cohesion;
  S.mu = (float)std::tan(d->friction);
  S.young = (float)d->young;
  S.h2 = (float)(h * h);
  const double t = std::tan(d->friction);
  S.alpha = t / std::sqrt(9 + 12 * t * t);
  S.kc = 3 * d->cohesion / std::sqrt(9 + 12 * t * t);
  S.G = d->young / (2 * (1 + d->poisson));
  S.K = d->young / (3 * (1 - 2 * d->poisson));
  S.sin_psi = std::sin(d->dilatancy);
  S.eps_f = 1e-4;
  (void)E;
  return S;
}

bool slot_ok(const tsph_engine* E, int slot) {
  return E && slot >= 0 && slot < (int)E->scenes.size() && E->scenes[slot].active;
}

// ------------------------------------------------------------- transfers
struct FieldInfo {
  int comps;
  bool is_int;
};
FieldInfo field_info(int f) {
  switch (f) {
    case TSPH_F_POS: case TSPH_F_VEL: case TSPH_F_D_VEL: case TSPH_F_V_TMP: case TSPH_F_OFFSET:
    case TSPH_F_COLOR: case TSPH_F_X0: case TSPH_F_COLOR0:
      return {3, false};
    case TSPH_F_STRESS: case TSPH_F_STRESS_TMP: case TSPH_F_D_STRESS:
      return {6, false};
    case TSPH_F_CSPM_L:
      return {9, false};
    case TSPH_F_TYPE: case TSPH_F_ID0: case TSPH_F_OBJ_ID: case TSPH_F_FLAG_RETMAP: case TSPH_F_CELL:
      return {1, true};
    default:
      return {1, false};
  }
}

template <class T>
int d2h(tsph_engine* E, const T* dptr, int64_t off, int64_t n, std::vector<T>& out) {
  out.resize(n);
  if (n == 0) return TSPH_OK;
  if (!dptr) return fail(TSPH_ERR_INVALID_ARG, "field not held by this engine's method");
  HIP_TRY(hipMemcpyAsync(out.data(), dptr + off, n * sizeof(T), hipMemcpyDeviceToHost, E->stream));
  HIP_TRY(hipStreamSynchronize(E->stream));
  return TSPH_OK;
}
template <class T>
int h2d(tsph_engine* E, T* dptr, int64_t off, const std::vector<T>& in) {
  if (in.empty()) return TSPH_OK;
  if (!dptr) return fail(TSPH_ERR_INVALID_ARG, "field not held by this engine's method");
  HIP_TRY(hipMemcpyAsync(dptr + off, in.data(), in.size() * sizeof(T), hipMemcpyHostToDevice, E->stream));
  HIP_TRY(hipStreamSynchronize(E->stream));
  return TSPH_OK;
}

int transfer(tsph_engine* E, int slot, int field, void* host, size_t bytes, bool up) {
  if (!slot_ok(E, slot)) return fail(TSPH_ERR_INVALID_ARG, "bad slot");
  if (up) E->nl_valid = E->cflow_valid = E->fresh_sort = false;
  if (field < 0 || field >= TSPH_F_COUNT) return fail(TSPH_ERR_INVALID_ARG, "bad field id");
  const SceneDev& S = E->scenes[slot];
  const int64_t n = S.n, off = S.pbase;
  HIP_TRY(hipStreamSynchronize(E->stream));
  if (field == TSPH_F_CELL_START) {
    if (up) return fail(TSPH_ERR_INVALID_ARG, "cell_start is download-only");
    const int64_t C = S.ncell;
    if (bytes != (size_t)(C + 1) * 4) return fail(TSPH_ERR_INVALID_ARG, "size mismatch for cell_start");
    std::vector<int32_t> st;
    int r = d2h(E, (const int32_t*)E->w.start, S.cbase, C + 1, st);
    if (r) return r;
    int32_t* o = (int32_t*)host;
    for (int64_t c = 0; c <= C; ++c) o[c] = st[c] - st[0];
    return TSPH_OK;
  }
  const FieldInfo fi = field_info(field);
  const size_t want = (size_t)n * fi.comps * (fi.is_int ? 4 : 8);
  if (bytes != want) return fail(TSPH_ERR_INVALID_ARG, "size mismatch: expected " + std::to_string(want) + " bytes");
  double* hd = (double*)host;
  int32_t* hi = (int32_t*)host;
  PState& cur = E->buf[E->parity];
  PWork& w = E->w;
  const int64_t st = E->cap;
  int r = TSPH_OK;
  auto f4 = [&](float4* arr, int lane0, int nl, bool pos_abs) -> int {
    std::vector<float4> v;
    if ((r = d2h(E, (const float4*)arr, off, n, v))) return r;
    std::vector<int32_t> cell;
    if (pos_abs && (r = d2h(E, (const int32_t*)cur.cell, off, n, cell))) return r;
    for (int64_t i = 0; i < n; ++i) {
      float* e = &v[i].x;
      int c[3] = {0, 0, 0};
      if (pos_abs && !up) {
        const int32_t k = cell[i];
        const int yz = S.gn[1] * S.gn[2];
        c[0] = k / yz;
        c[1] = (k - c[0] * yz) / S.gn[2];
        c[2] = k - c[0] * yz - c[1] * S.gn[2];
      }
      for (int l = 0; l < nl; ++l) {
        const int ax = lane0 + l;
        if (up) {
          double x = hd[nl * i + l];
          if (pos_abs) {
            if (ax == 2 && S.dim == 2) {
              e[ax] = 0.0f;
              c[2] = 0;
              continue;
            }
            const double rel = x - S.vds[ax];
            const double cf = rel / S.grid;
            // the cell grid covers the virtual domain; outside it the
            // reference's pos_to_index would address cells that do not exist
            if (!(cf >= 0.0 && cf < (double)S.gn[ax]))
              return fail(TSPH_ERR_INVALID_ARG, "particle " + std::to_string(i) + " lies outside the cell grid");
            const int ci = (int)cf;  // pos_to_index, particle_system.py:216-218
            c[ax] = ci;
            e[ax] = (float)(rel - (double)ci * S.grid);
          } else {
            e[ax] = (float)x;
          }
        } else {
          if (pos_abs)
            hd[nl * i + l] = (ax == 2 && S.dim == 2) ? 0.0 : S.vds[ax] + (double)c[ax] * S.grid + (double)e[ax];
          else
            hd[nl * i + l] = (double)e[ax];
        }
      }
      if (up && pos_abs) cell[i] = (c[0] * S.gn[1] + c[1]) * S.gn[2] + c[2];
    }
    if (up) {
      if ((r = h2d(E, arr, off, v))) return r;
      if (pos_abs && (r = h2d(E, cur.cell, off, cell))) return r;
    }
    return TSPH_OK;
  };
  auto f1 = [&](float* arr) -> int {
    std::vector<float> v;
    if (up) {
      v.resize(n);
      for (int64_t i = 0; i < n; ++i) v[i] = (float)hd[i];
      return h2d(E, arr, off, v);
    }
    if ((r = d2h(E, (const float*)arr, off, n, v))) return r;
    for (int64_t i = 0; i < n; ++i) hd[i] = v[i];
    return TSPH_OK;
  };
  auto d1 = [&](double* arr) -> int {
    std::vector<double> v;
    if (up) {
      v.assign(hd, hd + n);
      return h2d(E, arr, off, v);
    }
    if ((r = d2h(E, (const double*)arr, off, n, v))) return r;
    std::memcpy(hd, v.data(), n * 8);
    return TSPH_OK;
  };
  auto i1 = [&](int32_t* arr) -> int {
    std::vector<int32_t> v;
    if (up) {
      v.assign(hi, hi + n);
      return h2d(E, arr, off, v);
    }
    if ((r = d2h(E, (const int32_t*)arr, off, n, v))) return r;
    std::memcpy(hi, v.data(), n * 4);
    return TSPH_OK;
  };
  // 6-component stress: float (muI) or double (DP), component-major.
  auto s6 = [&](float* af, double* ad, bool mirror_f) -> int {
    for (int c = 0; c < 6; ++c) {
      if (ad) {
        std::vector<double> v;
        if (up) {
          v.resize(n);
          for (int64_t i = 0; i < n; ++i) v[i] = hd[6 * i + c];
          if ((r = h2d(E, ad, off + c * st, v))) return r;
          if (mirror_f && af) {
            std::vector<float> vf(n);
            for (int64_t i = 0; i < n; ++i) vf[i] = (float)v[i];
            if ((r = h2d(E, af, off + c * st, vf))) return r;
          }
        } else {
          if ((r = d2h(E, (const double*)ad, off + c * st, n, v))) return r;
          for (int64_t i = 0; i < n; ++i) hd[6 * i + c] = v[i];
        }
      } else if (af) {
        std::vector<float> v;
        if (up) {
          v.resize(n);
          for (int64_t i = 0; i < n; ++i) v[i] = (float)hd[6 * i + c];
          if ((r = h2d(E, af, off + c * st, v))) return r;
        } else {
          if ((r = d2h(E, (const float*)af, off + c * st, n, v))) return r;
          for (int64_t i = 0; i < n; ++i) hd[6 * i + c] = v[i];
        }
      } else {
        if (!up)
          for (int64_t i = 0; i < n; ++i) hd[6 * i + c] = 0.0;
      }
    }
    return TSPH_OK;
  };
  const bool dp = E->method == kDP, mui = E->method == kMUI;
  switch (field) {
    case TSPH_F_POS: return f4(cur.pos, 0, 3, true);
    case TSPH_F_OFFSET: return f4(cur.pos, 0, 3, false);
    case TSPH_F_M_V: return f4(cur.pos, 3, 1, false);
    case TSPH_F_VEL: return f4(cur.vel, 0, 3, false);
    case TSPH_F_DENSITY: return f4(cur.vel, 3, 1, false);
    case TSPH_F_V_TMP: return f4(cur.vtmp, 0, 3, false);
    case TSPH_F_DENSITY_TMP: return f4(cur.vtmp, 3, 1, false);
    case TSPH_F_D_VEL: return f4(w.dv, 0, 3, false);
    case TSPH_F_MASS: return f1(cur.mass);
    case TSPH_F_PRESSURE: return f1(w.p);
    case TSPH_F_CSPM_F: return f1(w.cspm);
    case TSPH_F_CSPM_L: {
      if (up) return fail(TSPH_ERR_INVALID_ARG, "cspm_L is download-only");
      for (int c = 0; c < 9; ++c) {
        std::vector<float> v;
        if ((r = d2h(E, (const float*)(w.Lc + (size_t)c * st), off, n, v))) return r;
        for (int64_t i = 0; i < n; ++i) hd[9 * i + c] = v[i];
      }
      return TSPH_OK;
    }
    case TSPH_F_D_DENSITY: return f1(w.drho);
    case TSPH_F_TYPE: {
      std::vector<int8_t> v;
      if (up) {
        // A scene holds the material of its method, rigid blocks and
        // boundary particles only (WC: fluid; muI / DP: soil); the pair
        // kernels rely on it.
        const int flow = E->method == kWC ? kFluid : kSoil;
        v.resize(n);
        for (int64_t i = 0; i < n; ++i) {
          const int t = hi[i];
          if (t != flow && t != kDummy && t != kRep && t != kRigid)
            return fail(TSPH_ERR_UNSUPPORTED, "particle type " + std::to_string(t) +
                                                  " in a scene of method " + std::to_string(E->method) +
                                                  " (only its own material and boundary particles)");
          v[i] = (int8_t)t;
        }
        return h2d(E, cur.type, off, v);
      }
      if ((r = d2h(E, (const int8_t*)cur.type, off, n, v))) return r;
      for (int64_t i = 0; i < n; ++i) hi[i] = v[i];
      return TSPH_OK;
    }
    case TSPH_F_ID0:
      // kernels index the scene's id0-ordered arrays (initial positions and
      // colours) with it: a value outside [0, n) would read past the scene
      if (up)
        for (int64_t i = 0; i < n; ++i)
          if (hi[i] < 0 || hi[i] >= n)
            return fail(TSPH_ERR_INVALID_ARG, "id0 " + std::to_string(hi[i]) + " out of range [0, " +
                                                  std::to_string(n) + ")");
      return i1(cur.id0);
    case TSPH_F_OBJ_ID: return i1(cur.obj);
    case TSPH_F_CELL:
      if (up)
        for (int64_t i = 0; i < n; ++i)
          if (hi[i] < 0 || hi[i] >= S.ncell)
            return fail(TSPH_ERR_INVALID_ARG, "cell id " + std::to_string(hi[i]) + " out of range");
      return i1(cur.cell);
    case TSPH_F_FLAG_RETMAP:
      if (!dp) {
        if (up) return TSPH_OK;
        std::memset(hi, 0, n * 4);
        return TSPH_OK;
      }
      return i1(cur.flag);
    case TSPH_F_STRESS: return s6(mui ? cur.sig_f : nullptr, dp ? cur.sig_d : nullptr, false);
    case TSPH_F_STRESS_TMP: return s6(w.sigt_f, dp ? w.sigt_d : nullptr, true);
    case TSPH_F_D_STRESS: return s6(nullptr, dp ? w.dsig_d : nullptr, false);
    case TSPH_F_STRAIN_EQU:
      if (!cur.seq) { if (!up) std::memset(hd, 0, n * 8); return TSPH_OK; }
      if (mui && up) {
        std::vector<double> v(n);
        for (int64_t i = 0; i < n; ++i) v[i] = (double)(float)hd[i];
        return h2d(E, cur.seq, off, v);
      }
      return d1(cur.seq);
    case TSPH_F_STRAIN_EQU_P:
      if (!cur.seqp) { if (!up) std::memset(hd, 0, n * 8); return TSPH_OK; }
      return d1(cur.seqp);
    case TSPH_F_D_STRAIN_EQU:
      if (!w.dseq) { if (!up) std::memset(hd, 0, n * 8); return TSPH_OK; }
      return d1(w.dseq);
    case TSPH_F_D_STRAIN_EQU_P:
      if (!w.dseqp) { if (!up) std::memset(hd, 0, n * 8); return TSPH_OK; }
      return d1(w.dseqp);
    case TSPH_F_VAL:
    case TSPH_F_COLOR:
    case TSPH_F_X0:
    case TSPH_F_COLOR0: {
      if (up && (field == TSPH_F_VAL || field == TSPH_F_COLOR))
        return fail(TSPH_ERR_INVALID_ARG, "val and color are download-only");
      if ((r = ensure_colour(E))) return r;
      if (field == TSPH_F_VAL) return d1(w.val);
      if (field == TSPH_F_X0) {
        for (int c = 0; c < 3; ++c) {
          std::vector<double> v;
          if (up) {
            v.resize(n);
            for (int64_t i = 0; i < n; ++i) v[i] = hd[3 * i + c];
            if ((r = h2d(E, w.x0 + (size_t)c * st, off, v))) return r;
          } else {
            if ((r = d2h(E, (const double*)(w.x0 + (size_t)c * st), off, n, v))) return r;
            for (int64_t i = 0; i < n; ++i) hd[3 * i + c] = v[i];
          }
        }
        return TSPH_OK;
      }
      return f4(field == TSPH_F_COLOR ? w.rgb : w.c0, 0, 3, false);
    }
  }
  return fail(TSPH_ERR_INVALID_ARG, "unhandled field");
}

}  // namespace

// =============================================================== the C ABI
extern "C" {

int tsph_abi_version(void) { return TSPH_ABI_VERSION; }

const char* tsph_last_error(void) { return g_err.c_str(); }

int tsph_device_count(int* n) {
  if (!n) return fail(TSPH_ERR_INVALID_ARG, "null pointer");
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  *n = c;
  return TSPH_OK;
}

int tsph_engine_create(int method, int dim, int64_t particle_capacity, int64_t cell_capacity,
                       int max_scenes, void** out_engine) {
  if (!out_engine) return fail(TSPH_ERR_INVALID_ARG, "null out pointer");
  *out_engine = nullptr;
  if (method < 1 || method > 3) return fail(TSPH_ERR_INVALID_ARG, "method must be 1, 2 or 3");
  if (dim != 2 && dim != 3) return fail(TSPH_ERR_INVALID_ARG, "dim must be 2 or 3");
  if (particle_capacity < 1 || particle_capacity > (1ll << 26))
    return fail(TSPH_ERR_INVALID_ARG, "particle capacity out of range");
  if (cell_capacity < 1 || cell_capacity > (int64_t)kScanTile * kScanTile)
    return fail(TSPH_ERR_INVALID_ARG, "cell capacity out of range");
  if (max_scenes < 1 || max_scenes > 65536) return fail(TSPH_ERR_INVALID_ARG, "max_scenes out of range");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(TSPH_ERR_HIP, "no HIP device available");
  // Particle ranges start on 64-particle tiles; each scene may waste <1 tile.
  const int64_t pc = (particle_capacity + kTileAlign - 1) / kTileAlign * kTileAlign + (int64_t)kTileAlign * max_scenes;
  if (pc >= (1ll << 26))
    return fail(TSPH_ERR_INVALID_ARG, "particle capacity (with per-scene alignment) must stay below 2^26");
  tsph_engine* E = new tsph_engine();
  E->method = method;
  E->dim = dim;
  E->max_scenes = max_scenes;
  hipGetDevice(&E->device);
  {
    // set-up of tsph_neighbor_stats_noalloc's scratch on this device
    char* scr = nullptr;
    if (int rs = scratch_on(E->device, &scr)) {
      delete E;
      return rs;
    }
  }
  E->cap = (int32_t)pc;
  E->ccap = (int32_t)cell_capacity;
  // 3D lattice: 92 neighbours (+30 exactly at r_s); 2D: 24 (+4).  Denser
  // particles fall back to the stencil walk.
  E->nl_cap = dim == 3 ? 160 : 48;
  E->tiles_alloc = (int32_t)((pc / 64 + 15) / 16 * 16);
  E->palloc.cap = pc;
  E->calloc_.cap = cell_capacity;
  E->scenes.assign(max_scenes, SceneDev{});
  for (auto& s : E->scenes) s.active = 0;
  E->descs.assign(max_scenes, tsph_scene_desc{});
  E->steps.assign(max_scenes, 0);
  E->tiles.assign(E->tiles_alloc, make_int2(-1, 0));
  int r = TSPH_OK;
  if (hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking) != hipSuccess) r = fail(TSPH_ERR_HIP, "hipStreamCreate failed");
  if (!r) r = alloc_state(E, E->buf[0]);
  if (!r) r = alloc_state(E, E->buf[1]);
  if (!r) r = alloc_work(E);
  if (!r) r = dalloc(&E->d_scenes, max_scenes, E->bytes);
  if (!r) r = dalloc(&E->d_tiles, E->tiles_alloc, E->bytes);
  if (!r) r = upload_tables(E);
  if (r) {
    tsph_engine_destroy(E);
    return r;
  }
  *out_engine = E;
  return TSPH_OK;
}

int tsph_engine_destroy(void* engine) {
  tsph_engine* E = (tsph_engine*)engine;
  if (!E) return fail(TSPH_ERR_INVALID_ARG, "null engine");
  if (E->stream) hipStreamSynchronize(E->stream);
  invalidate_graph(E);
  free_state(E->buf[0]);
  free_state(E->buf[1]);
  free_work(E->w);
  if (E->d_scenes) hipFree(E->d_scenes);
  if (E->d_tiles) hipFree(E->d_tiles);
  for (auto& r : E->timing.pending) E->timing.pool.push_back({r.a, r.b});
  for (auto& pr : E->timing.pool) {
    hipEventDestroy(pr.first);
    hipEventDestroy(pr.second);
  }
  if (E->stream) hipStreamDestroy(E->stream);
  delete E;
  return TSPH_OK;
}

int tsph_engine_bytes(void* engine, int64_t* out_bytes) {
  tsph_engine* E = (tsph_engine*)engine;
  if (!E || !out_bytes) return fail(TSPH_ERR_INVALID_ARG, "null argument");
  *out_bytes = (int64_t)E->bytes;
  return TSPH_OK;
}

int tsph_add_scene(void* engine, const tsph_scene_desc* desc, int* out_slot) {
  tsph_engine* E = (tsph_engine*)engine;
  if (!E || !out_slot) return fail(TSPH_ERR_INVALID_ARG, "null argument");
  int r = validate_desc(E, desc);
  if (r) return r;
  int slot = -1;
  for (int s = 0; s < E->max_scenes; ++s)
    if (!E->scenes[s].active) {
      slot = s;
      break;
    }
  if (slot < 0) return fail(TSPH_ERR_CAPACITY, "no free scene slot");
  const int64_t ncell = (int64_t)desc->grid_num[0] * desc->grid_num[1] * desc->grid_num[2];
  const int64_t pb = E->palloc.alloc(std::max<int64_t>(desc->n_particles, 1), kTileAlign);
  if (pb < 0) return fail(TSPH_ERR_CAPACITY, "not enough particle capacity");
  const int64_t cb = E->calloc_.alloc(ncell, 1);
  if (cb < 0) {
    E->palloc.release(pb);
    return fail(TSPH_ERR_CAPACITY, "not enough cell capacity");
  }
  SceneDev S = make_scene(E, desc);
  S.pbase = (int32_t)pb;
  S.cbase = (int32_t)cb;
  // Static particle -> slot map and zeroed state for the range; on any
  // failure the reserved ranges go back to the allocators.
  r = [&]() -> int {
    HIP_TRY(hipStreamSynchronize(E->stream));
    std::vector<int32_t> sid(desc->n_particles, slot);
    int rr;
    if ((rr = h2d(E, E->w.sid, pb, sid))) return rr;
    for (int b = 0; b < 2; ++b) {
      PState& P = E->buf[b];
      const size_t n = desc->n_particles;
      if (n == 0) break;
      HIP_TRY(hipMemsetAsync(P.cell + pb, 0, n * 4, E->stream));
      HIP_TRY(hipMemsetAsync(P.pos + pb, 0, n * sizeof(float4), E->stream));
      HIP_TRY(hipMemsetAsync(P.vel + pb, 0, n * sizeof(float4), E->stream));
      HIP_TRY(hipMemsetAsync(P.vtmp + pb, 0, n * sizeof(float4), E->stream));
      HIP_TRY(hipMemsetAsync(P.mass + pb, 0, n * 4, E->stream));
      HIP_TRY(hipMemsetAsync(P.type + pb, 0, n, E->stream));
      for (int c = 0; c < 6; ++c) {
        if (P.sig_f) HIP_TRY(hipMemsetAsync(P.sig_f + c * (size_t)E->cap + pb, 0, n * 4, E->stream));
        if (P.sig_d) HIP_TRY(hipMemsetAsync(P.sig_d + c * (size_t)E->cap + pb, 0, n * 8, E->stream));
      }
      if (P.seq) HIP_TRY(hipMemsetAsync(P.seq + pb, 0, n * 8, E->stream));
      if (P.seqp) HIP_TRY(hipMemsetAsync(P.seqp + pb, 0, n * 8, E->stream));
      if (P.flag) HIP_TRY(hipMemsetAsync(P.flag + pb, 0, n * 4, E->stream));
    }
    HIP_TRY(hipMemsetAsync(E->w.status + 2 * slot, 0, 16, E->stream));
    return TSPH_OK;
  }();
  if (r) {
    E->palloc.release(pb);
    E->calloc_.release(cb);
    return r;
  }
  E->scenes[slot] = S;
  E->descs[slot] = *desc;
  E->nl_valid = E->cflow_valid = E->fresh_sort = false;
  E->steps[slot] = 0;
  if ((r = upload_tables(E))) {
    const std::string msg = g_err;
    E->scenes[slot].active = 0;
    E->palloc.release(pb);
    E->calloc_.release(cb);
    return fail(r, msg);
  }
  *out_slot = slot;
  return TSPH_OK;
}

int tsph_remove_scene(void* engine, int slot) {
  tsph_engine* E = (tsph_engine*)engine;
  if (!slot_ok(E, slot)) return fail(TSPH_ERR_INVALID_ARG, "bad slot");
  HIP_TRY(hipStreamSynchronize(E->stream));
  E->palloc.release(E->scenes[slot].pbase);
  E->calloc_.release(E->scenes[slot].cbase);
  E->scenes[slot].active = 0;
  E->nl_valid = E->cflow_valid = E->fresh_sort = false;
  return upload_tables(E);
}

int tsph_upload(void* engine, int slot, int field, const void* host, size_t bytes) {
  tsph_engine* E = (tsph_engine*)engine;
  if (!E || !host) return fail(TSPH_ERR_INVALID_ARG, "null argument");
  return transfer(E, slot, field, (void*)host, bytes, true);
}

int tsph_download(void* engine, int slot, int field, void* host, size_t bytes) {
  tsph_engine* E = (tsph_engine*)engine;
  if (!E || !host) return fail(TSPH_ERR_INVALID_ARG, "null argument");
  return transfer(E, slot, field, host, bytes, false);
}

int tsph_step(void* engine, int nsteps) {
  tsph_engine* E = (tsph_engine*)engine;
  if (!E || nsteps < 0) return fail(TSPH_ERR_INVALID_ARG, "bad argument");
  if (E->active_count() == 0 || nsteps == 0) return TSPH_OK;
  if (E->use_graph && !E->timing.on) {
    // recapture when the residents ask for other step options than the
    // graphs hold (a scene with XSPH, collisions or kernel correction added)
    if (!E->graph_valid || !(E->graph_flags == step_flags(E))) {
      int r = build_graphs(E);
      if (r) return r;
    }
    for (int k = 0; k < nsteps; ++k) {
      HIP_TRY(hipGraphLaunch(E->graph[E->parity], E->stream));
      E->parity ^= 1;
    }
  } else {
    for (int k = 0; k < nsteps; ++k) {
      int r = run_step_eager(E);
      if (r) return r;
    }
  }
  for (size_t s = 0; s < E->scenes.size(); ++s)
    if (E->scenes[s].active) E->steps[s] += nsteps;
  return TSPH_OK;
}

int tsph_sync(void* engine) {
  tsph_engine* E = (tsph_engine*)engine;
  if (!E) return fail(TSPH_ERR_INVALID_ARG, "null engine");
  HIP_TRY(hipStreamSynchronize(E->stream));
  resolve_timing(E);
  std::vector<unsigned long long> st(2 * E->max_scenes);
  HIP_TRY(hipMemcpy(st.data(), E->w.status, st.size() * 8, hipMemcpyDeviceToHost));
  for (int s = 0; s < E->max_scenes; ++s) {
    if (!E->scenes[s].active) continue;
    if (st[2 * s]) return fail(TSPH_ERR_PARTICLE_ESCAPED, "scene " + std::to_string(s) + ": " + std::to_string(st[2 * s]) + " particle(s) left the grid");
    if (st[2 * s + 1]) return fail(TSPH_ERR_NAN, "scene " + std::to_string(s) + ": NaN in density/velocity");
  }
  return TSPH_OK;
}

int tsph_run_pass(void* engine, int pass_id) {
  tsph_engine* E = (tsph_engine*)engine;
  if (!E) return fail(TSPH_ERR_INVALID_ARG, "null engine");
  if (E->active_count() == 0) return TSPH_OK;
  Launcher L{E};
  StepFlags f = step_flags(E);
  auto run = [&](auto seq) {
    switch (pass_id) {
      case TSPH_PASS_SORT: seq.sort(); break;
      case TSPH_PASS_SHEPARD: seq.shepard(); break;
      case TSPH_PASS_INIT_TMP: seq.init_tmp(); break;
      case TSPH_PASS_ONE_STEP: seq.one_step(); break;
      case TSPH_PASS_LF_HALF: seq.lf_half(); break;
      case TSPH_PASS_LF_FULL: case TSPH_PASS_SE: seq.lf_full(); break;
      case TSPH_PASS_ADVECT_POS: seq.advect_pos(f.xsph); break;
      case TSPH_PASS_ADVECT_SOMETHING: seq.advect_something(); break;
      case TSPH_PASS_ENFORCE_BOUNDARY: seq.enforce_boundary(); break;
      case TSPH_PASS_STEP: seq.step_passes(f.ti, f.xsph, f.collide, f.rigid); break;
      case TSPH_PASS_SOLVE_RIGID: seq.solve_rigid(); break;
      default: L.status = fail(TSPH_ERR_INVALID_ARG, "bad pass id");
    }
  };
  if (E->dim == 2) run(Seq<2>{E, L});
  else run(Seq<3>{E, L});
  if (L.status) return L.status;
  if (pass_id == TSPH_PASS_STEP)
    for (size_t s = 0; s < E->scenes.size(); ++s)
      if (E->scenes[s].active) E->steps[s] += 1;
  HIP_TRY(hipStreamSynchronize(E->stream));
  return TSPH_OK;
}

int tsph_scene_steps(void* engine, int slot, int64_t* out_steps) {
  tsph_engine* E = (tsph_engine*)engine;
  if (!slot_ok(E, slot) || !out_steps) return fail(TSPH_ERR_INVALID_ARG, "bad argument");
  *out_steps = E->steps[slot];
  return TSPH_OK;
}

int tsph_set_graph(void* engine, int enable) {
  tsph_engine* E = (tsph_engine*)engine;
  if (!E) return fail(TSPH_ERR_INVALID_ARG, "null engine");
  E->use_graph = enable != 0;
  return TSPH_OK;
}

int tsph_set_neighbor_list(void* engine, int enable) {
  tsph_engine* E = (tsph_engine*)engine;
  if (!E) return fail(TSPH_ERR_INVALID_ARG, "null engine");
  HIP_TRY(hipStreamSynchronize(E->stream));
  if (enable < 0 || enable > 2) return fail(TSPH_ERR_INVALID_ARG, "neighbour-list mode must be 0, 1 or 2");
  E->use_nl = enable != 0;
  E->nl_lds = enable == 2;
  E->nl_valid = false;
  invalidate_graph(E);
  return TSPH_OK;
}

int tsph_set_cell_tile(void* engine, int stage) {
  tsph_engine* E = (tsph_engine*)engine;
  if (!E) return fail(TSPH_ERR_INVALID_ARG, "null engine");
  if (stage != 128 && stage != 256 && stage != 512)
    return fail(TSPH_ERR_INVALID_ARG, "cell tile must be 128, 256 or 512 candidates");
  HIP_TRY(hipStreamSynchronize(E->stream));
  E->nl_stage = stage;
  invalidate_graph(E);
  return TSPH_OK;
}

int tsph_set_block(void* engine, int pair_block, int elem_block) {
  tsph_engine* E = (tsph_engine*)engine;
  if (!E) return fail(TSPH_ERR_INVALID_ARG, "null engine");
  auto ok = [](int b) { return b == 64 || b == 128 || b == 256; };
  if (!ok(pair_block) || !ok(elem_block)) return fail(TSPH_ERR_INVALID_ARG, "block sizes must be 64, 128 or 256");
  HIP_TRY(hipStreamSynchronize(E->stream));
  E->pair_block = pair_block;
  E->elem_block = elem_block;
  invalidate_graph(E);
  return TSPH_OK;
}

int tsph_globals(void* engine, int slot, double out[4]) {
  tsph_engine* E = (tsph_engine*)engine;
  if (!slot_ok(E, slot) || !out) return fail(TSPH_ERR_INVALID_ARG, "bad argument");
  Launcher L{E};
  if (E->dim == 2) Seq<2>{E, L}.globals();
  else Seq<3>{E, L}.globals();
  if (L.status) return L.status;
  HIP_TRY(hipMemcpyAsync(out, E->w.gout + 4 * slot, 32, hipMemcpyDeviceToHost, E->stream));
  HIP_TRY(hipStreamSynchronize(E->stream));
  return TSPH_OK;
}

int tsph_set_rigid_object(void* engine, int slot, int object_id, const double rest_cm[3]) {
  tsph_engine* E = (tsph_engine*)engine;
  if (!slot_ok(E, slot) || !rest_cm) return fail(TSPH_ERR_INVALID_ARG, "bad argument");
  SceneDev& S = E->scenes[slot];
  int k = 0;
  while (k < S.nrig && S.rig_obj[k] != object_id) ++k;  // a known object: new rest centre
  if (k == kMaxRigid)
    return fail(TSPH_ERR_CAPACITY, "at most " + std::to_string(kMaxRigid) + " dynamic rigid objects per scene");
  int r = ensure_rigid(E);
  if (r) return r;
  S.rig_obj[k] = object_id;
  for (int c = 0; c < 3; ++c) S.rig_cm0[k][c] = rest_cm[c];
  if (k == S.nrig) S.nrig++;
  HIP_TRY(hipStreamSynchronize(E->stream));
  HIP_TRY(hipMemcpy(E->d_scenes, E->scenes.data(), E->scenes.size() * sizeof(SceneDev),
                    hipMemcpyHostToDevice));
  invalidate_graph(E);
  return TSPH_OK;
}

int tsph_assign_value_color(void* engine, int slot, int color_title, int show_bdy, uint32_t seed) {
  tsph_engine* E = (tsph_engine*)engine;
  if (!slot_ok(E, slot)) return fail(TSPH_ERR_INVALID_ARG, "bad slot");
  int r = ensure_colour(E);
  if (r) return r;
  Launcher L{E};
  L.go("color", k_color_val, E->nblocks_tiles(256), 256, E->args(), slot, color_title,
       show_bdy ? 1 : 0, seed);
  return L.status;
}

int tsph_v_maxmin(void* engine, int slot, double given_max, double given_min, int fix_max,
                  int fix_min, double* out_vmax, double* out_vmin) {
  tsph_engine* E = (tsph_engine*)engine;
  if (!slot_ok(E, slot) || !out_vmax || !out_vmin) return fail(TSPH_ERR_INVALID_ARG, "bad argument");
  int r = ensure_colour(E);
  if (r) return r;
  Launcher L{E};
  const KArgs a = E->args();
  L.go("color", k_color_range_tile, E->nblocks_tiles(256), 256, a, slot);
  L.go("color", k_color_range_scene, 1, 256, a, slot, given_max, given_min, fix_max ? 1 : 0,
       fix_min ? 1 : 0);
  if (L.status) return L.status;
  double v[2];
  HIP_TRY(hipMemcpyAsync(v, E->w.crange + 2 * slot, 16, hipMemcpyDeviceToHost, E->stream));
  HIP_TRY(hipStreamSynchronize(E->stream));
  *out_vmax = v[0];
  *out_vmin = v[1];
  return TSPH_OK;
}

int tsph_set_color(void* engine, int slot, int color_group, int show_bdy) {
  tsph_engine* E = (tsph_engine*)engine;
  if (!slot_ok(E, slot)) return fail(TSPH_ERR_INVALID_ARG, "bad slot");
  int r = ensure_colour(E);
  if (r) return r;
  Launcher L{E};
  L.go("color", k_set_color, E->nblocks_tiles(256), 256, E->args(), slot, color_group,
       show_bdy ? 1 : 0);
  return L.status;
}

int tsph_status(void* engine, int slot, int64_t out[2]) {
  tsph_engine* E = (tsph_engine*)engine;
  if (!slot_ok(E, slot) || !out) return fail(TSPH_ERR_INVALID_ARG, "bad argument");
  unsigned long long st[2];
  HIP_TRY(hipStreamSynchronize(E->stream));
  HIP_TRY(hipMemcpy(st, E->w.status + 2 * slot, 16, hipMemcpyDeviceToHost));
  out[0] = (int64_t)st[0];
  out[1] = (int64_t)st[1];
  return TSPH_OK;
}

int tsph_kernel_time(void* engine, const char* name, double* out_ms, int64_t* out_launches) {
  tsph_engine* E = (tsph_engine*)engine;
  if (!E || !name || !out_ms || !out_launches) return fail(TSPH_ERR_INVALID_ARG, "null argument");
  HIP_TRY(hipStreamSynchronize(E->stream));
  resolve_timing(E);
  auto it = E->timing.acc.find(name);
  if (it == E->timing.acc.end()) {
    *out_ms = 0.0;
    *out_launches = 0;
    return TSPH_OK;
  }
  *out_launches = it->second.second;
  *out_ms = it->second.second ? it->second.first / it->second.second : 0.0;
  return TSPH_OK;
}

int tsph_timing(void* engine, int enable) {
  tsph_engine* E = (tsph_engine*)engine;
  if (!E) return fail(TSPH_ERR_INVALID_ARG, "null engine");
  HIP_TRY(hipStreamSynchronize(E->stream));
  resolve_timing(E);
  E->timing.acc.clear();
  E->timing.on = enable != 0;
  return TSPH_OK;
}

int tsph_neighbor_stats(void* engine, int slot, int32_t* count, int64_t* id_sum,
                        int64_t* id_sqsum) {
  tsph_engine* E = (tsph_engine*)engine;
  if (!slot_ok(E, slot) || !count || !id_sum || !id_sqsum) return fail(TSPH_ERR_INVALID_ARG, "bad argument");
  const SceneDev& S = E->scenes[slot];
  int32_t* dc = nullptr;
  long long *d1 = nullptr, *d2 = nullptr;
  size_t bytes = 0;
  int r = dalloc(&dc, E->cap, bytes);
  if (!r) r = dalloc(&d1, E->cap, bytes);
  if (!r) r = dalloc(&d2, E->cap, bytes);
  if (!r) {
    Launcher L{E};
    if (E->dim == 2)
      L.go("neighbor_stats", k_neighbor_stats<2>, E->nblocks_tiles(256), 256, E->args(), dc, d1, d2, -1, 0);
    else
      L.go("neighbor_stats", k_neighbor_stats<3>, E->nblocks_tiles(256), 256, E->args(), dc, d1, d2, -1, 0);
    r = L.status;
  }
  if (!r) {
    hipError_t e = hipStreamSynchronize(E->stream);
    if (e == hipSuccess) e = hipMemcpy(count, dc + S.pbase, S.n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(id_sum, d1 + S.pbase, S.n * 8, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(id_sqsum, d2 + S.pbase, S.n * 8, hipMemcpyDeviceToHost);
    if (e != hipSuccess) r = fail(TSPH_ERR_HIP, std::string("neighbor_stats: ") + hipGetErrorString(e));
  }
  if (dc) hipFree(dc);
  if (d1) hipFree(d1);
  if (d2) hipFree(d2);
  return r;
}

int tsph_check_sqrt(uint64_t out[2]) {
  if (!out) return fail(TSPH_ERR_INVALID_ARG, "null argument");
  unsigned long long* d = nullptr;
  HIP_TRY(hipMalloc((void**)&d, 2 * sizeof(unsigned long long)));
  hipError_t e = hipMemset(d, 0, 2 * sizeof(unsigned long long));
  if (e == hipSuccess) {
    // every finite non-negative float: [+0, +inf)
    hipLaunchKernelGGL(k_check_pair_dist, dim3(8192), dim3(256), 0, nullptr, 0u, 0x7F800000u, d);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  unsigned long long h[2] = {0, 0};
  if (e == hipSuccess) e = hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost);
  hipFree(d);
  if (e != hipSuccess) return fail(TSPH_ERR_HIP, std::string("check_sqrt: ") + hipGetErrorString(e));
  out[0] = h[0];
  out[1] = h[1];
  return TSPH_OK;
}

int tsph_neighbor_stats_noalloc(void* engine, int slot, int32_t* count, int64_t* id_sum,
                                int64_t* id_sqsum) {
  tsph_engine* E = (tsph_engine*)engine;
  if (!slot_ok(E, slot) || !count || !id_sum || !id_sqsum) return fail(TSPH_ERR_INVALID_ARG, "bad argument");
  const SceneDev& S = E->scenes[slot];
  if (S.n > kStatsScratchMax)
    return fail(TSPH_ERR_INVALID_ARG, "neighbor_stats_noalloc: " + std::to_string(S.n) +
                                          " particles in the slot, at most " +
                                          std::to_string(kStatsScratchMax));
  char* base = nullptr;
  int r = scratch_on(E->device, &base);
  if (r) return r;
  int32_t* dc = reinterpret_cast<int32_t*>(base);
  long long* d1 = reinterpret_cast<long long*>(base + kStatsCountBytes);
  long long* d2 = d1 + kStatsScratchMax;
  // Launched directly (no timing events are created).  Every particle of
  // the slot writes its three entries, so nothing needs clearing; the array
  // keeps the last call's contents, and calls that overlap on one GPU race.
  if (E->dim == 2)
    hipLaunchKernelGGL(k_neighbor_stats<2>, dim3(E->nblocks_tiles(256)), dim3(256), 0, E->stream, E->args(), dc,
                       d1, d2, slot, S.pbase);
  else
    hipLaunchKernelGGL(k_neighbor_stats<3>, dim3(E->nblocks_tiles(256)), dim3(256), 0, E->stream, E->args(), dc,
                       d1, d2, slot, S.pbase);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(E->stream);
  if (e == hipSuccess && S.n > 0) e = hipMemcpy(count, dc, S.n * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess && S.n > 0) e = hipMemcpy(id_sum, d1, S.n * 8, hipMemcpyDeviceToHost);
  if (e == hipSuccess && S.n > 0) e = hipMemcpy(id_sqsum, d2, S.n * 8, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return fail(TSPH_ERR_HIP, std::string("neighbor_stats_noalloc: ") + hipGetErrorString(e));
  return TSPH_OK;
}

int tsph_neighbor_lists(void* engine, int slot, int32_t* count, int32_t* ids, int8_t* codes,
                        int64_t ids_cap, int64_t* out_total) {
  tsph_engine* E = (tsph_engine*)engine;
  if (!slot_ok(E, slot) || !count || !out_total || ids_cap < 0 || (ids_cap > 0 && (!ids || !codes)))
    return fail(TSPH_ERR_INVALID_ARG, "bad argument");
  HIP_TRY(hipStreamSynchronize(E->stream));
  if (!E->use_nl || !E->nl_valid)
    return fail(TSPH_ERR_INVALID_ARG, "no neighbour lists match the current particle order "
                                      "(list mode 0, or no list build since the last sort / upload)");
  const SceneDev& S = E->scenes[slot];
  const int64_t n = S.n, base = S.pbase;
  if (base % 64 != 0) return fail(TSPH_ERR_INVALID_ARG, "scene range not tile aligned");
  const int64_t ntiles = (n + 63) / 64, per_tile = (int64_t)E->nl_cap * 64;
  std::vector<int32_t> nlc, id0;
  std::vector<uint32_t> nl;
  int r;
  if ((r = d2h(E, (const int32_t*)E->w.nlc, base, n, nlc)) ||
      (r = d2h(E, (const int32_t*)E->buf[E->parity].id0, base, n, id0)) ||
      (r = d2h(E, (const uint32_t*)E->w.nl, (base / 64) * per_tile, ntiles * per_tile, nl)))
    return r;
  int64_t total = 0;
  for (int64_t i = 0; i < n; ++i) {
    const int32_t c = nlc[i];
    if (c < 0 || c > E->nl_cap) {
      count[i] = -1;  // walks the stencil in the pair passes
      continue;
    }
    count[i] = c;
    const int64_t lb = (i >> 6) * per_tile + ((i & 63) << 4);
    for (int32_t k = 0; k < c; ++k) {
      const uint32_t e = nl[lb + (int64_t)nl_at(k)];
      const int64_t j = (int64_t)(e >> kNlShift) - base;
      if (j < 0 || j >= n) return fail(TSPH_ERR_HIP, "neighbour list entry outside the scene");
      if (total < ids_cap) {
        ids[total] = id0[j];
        codes[total] = (int8_t)(e & 63u);
      }
      ++total;
    }
  }
  *out_total = total;
  if (total > ids_cap) return fail(TSPH_ERR_INVALID_ARG, "ids_cap too small: " + std::to_string(total) + " entries");
  return TSPH_OK;
}

int tsph_create(const tsph_scene_desc* desc, void** out_engine) {
  if (!desc || !out_engine) return fail(TSPH_ERR_INVALID_ARG, "null argument");
  const int64_t ncell = (int64_t)desc->grid_num[0] * desc->grid_num[1] * desc->grid_num[2];
  int r = tsph_engine_create(desc->method, desc->dim, std::max<int64_t>(desc->n_particles, 1), ncell, 1, out_engine);
  if (r) return r;
  int slot = -1;
  r = tsph_add_scene(*out_engine, desc, &slot);
  if (r) {
    std::string msg = g_err;
    tsph_engine_destroy(*out_engine);
    *out_engine = nullptr;
    return fail(r, msg);
  }
  return TSPH_OK;
}

int tsph_destroy(void* engine) { return tsph_engine_destroy(engine); }

}  // extern "C"
