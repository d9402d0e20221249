/*
 * tisphi.h — C ABI of libtisphi_hip.so, the MI355X (gfx950) SPH particle engine.
 *
 * The reference (tiSPHi) has no FFI: its "operator API" is the Taichi
 * @ti.data_oriented classes ParticleSystem / SPHBase / WCSPHSolver /
 * MUISPHSolver / DPSPHSolver.  Each entry point below replaces one piece of
 * that surface; the Python facade (tisphi_amd/eng/) binds them with ctypes
 * and re-exposes the reference's class interface.
 *
 * Conventions
 *   - Every function returns an int status: TSPH_OK (0) or a negative code.
 *     tsph_last_error() returns a thread-local message for the last failure.
 *     The library never aborts the process.
 *   - An engine owns all device memory.  Host buffers are borrowed for the
 *     duration of the call only.  Handles are opaque.
 *   - One HIP stream per engine; an engine is not re-entrant, distinct
 *     engines may be driven from distinct threads.
 *   - An engine holds up to max_scenes independent scenes ("slots") of one
 *     simulation method and dimension.  The single-scene functions
 *     (tsph_create ...) build a one-slot engine and address slot 0.
 *   - Host-side field data is float64 / int32 in the engine's current
 *     particle order (the order produced by the last counting sort; right
 *     after tsph_add_scene it is the upload order).  TSPH_F_ID0 carries the
 *     generation index of each particle.
 */
#ifndef TISPHI_H
#define TISPHI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TSPH_ABI_VERSION 1

/* ---- status codes ------------------------------------------------------ */
enum {
  TSPH_OK = 0,
  TSPH_ERR_INVALID_ARG = -1,
  TSPH_ERR_OOM = -2,
  TSPH_ERR_HIP = -3,
  TSPH_ERR_PARTICLE_ESCAPED = -4, /* a particle left the virtual grid      */
  TSPH_ERR_NAN = -5,              /* NaN detected in density/velocity     */
  TSPH_ERR_UNSUPPORTED = -6,      /* option of the scene format not built */
  TSPH_ERR_CAPACITY = -7,         /* no free slot / particle / cell room  */
};

/* ---- particle material types (particle_system.py:23-24) ---------------- */
enum {
  TSPH_MAT_FLUID = 1,
  TSPH_MAT_SOIL = 2,
  TSPH_MAT_RIGID = 11,
  TSPH_MAT_DUMMY = -1,
  TSPH_MAT_REPULSIVE = -2,
};

/* ---- simulation methods (simulation.py:14-22) -------------------------- */
enum { TSPH_METHOD_WC = 1, TSPH_METHOD_MUI = 2, TSPH_METHOD_DP = 3 };

/* ---- scene description --------------------------------------------------
 * Plain data derived from the scene JSON by the host
 * (ParticleSystem.__init__, particle_system.py:14-58; SPHBase.__init__,
 * solver_sph_base.py:14-21; solver constructors solver_sph_wc.py:12-21,
 * solver_sph_muI.py:12-25, solver_sph_dp.py:12-29).                        */
typedef struct tsph_scene_desc {
  int32_t dim;              /* 2 or 3                                        */
  int32_t method;           /* TSPH_METHOD_*                                 */
  int32_t boundary;         /* 0 none, 1 collision, 2 dummy, 3 repulsive,
                               4 dummy + repulsive                           */
  int32_t kernel;           /* 0 cubic spline, 1 Wendland C2                 */
  int32_t kernel_corr;      /* 0 none, 1 CSPM, 2 MLS switch (zero gradient)  */
  int32_t time_integration; /* 1 symplectic Euler, 2 leap-frog, 4 RK4        */
  int32_t xsph;             /* XSPH position update on/off                   */
  int32_t grid_num[3];      /* cells per axis; grid_num[2] == 1 in 2D        */
  int64_t n_particles;      /* particles of this scene                      */
  double vdomain_start[3];  /* virtual domain start (domain - grid)          */
  double domain_start[3];   /* physical domain (collision boundary)         */
  double domain_end[3];
  double grid_size;         /* ceil(kappa*kh) * diameter                     */
  double particle_radius;
  double particle_diameter;
  double smoothing_len;     /* kh * diameter                                 */
  double support_radius;    /* kappa * smoothing_len                         */
  double gravity[3];
  double dt;                /* from calc_dt_CFL, solver_sph_base.py:209-212  */
  /* material of mat_fluid[0] / mat_soil[0] */
  double density0;
  double viscosity;         /* WC                                            */
  double stiffness;         /* WC Tait B                                     */
  double exponent;          /* WC Tait gamma                                 */
  double cohesion;          /* soil, Pa                                      */
  double friction;          /* soil, radians                                 */
  double young;             /* soil E                                        */
  double poisson;           /* soil nu                                       */
  double dilatancy;         /* soil, radians                                 */
  double vsound;            /* solver sound speed                            */
} tsph_scene_desc;

/* ---- fields for tsph_upload / tsph_download -----------------------------
 * count = n_particles of the slot; "x3" = 3 components per particle
 * (row-major [n][3]); "x6" = stress components xx,yy,zz,xy,yz,zx.           */
enum {
  TSPH_F_POS = 0,           /* f64 x3, absolute position                    */
  TSPH_F_VEL = 1,           /* f64 x3                                        */
  TSPH_F_DENSITY = 2,       /* f64                                           */
  TSPH_F_MASS = 3,          /* f64                                           */
  TSPH_F_M_V = 4,           /* f64                                           */
  TSPH_F_PRESSURE = 5,      /* f64                                           */
  TSPH_F_TYPE = 6,          /* i32 mat_type                                  */
  TSPH_F_ID0 = 7,           /* i32                                           */
  TSPH_F_OBJ_ID = 8,        /* i32                                           */
  TSPH_F_STRESS = 9,        /* f64 x6                                        */
  TSPH_F_STRESS_TMP = 10,   /* f64 x6                                        */
  TSPH_F_D_STRESS = 11,     /* f64 x6                                        */
  TSPH_F_STRAIN_EQU = 12,   /* f64                                           */
  TSPH_F_STRAIN_EQU_P = 13, /* f64                                           */
  TSPH_F_D_STRAIN_EQU = 14, /* f64                                           */
  TSPH_F_D_STRAIN_EQU_P = 15, /* f64                                         */
  TSPH_F_FLAG_RETMAP = 16,  /* i32                                           */
  TSPH_F_CSPM_F = 17,       /* f64 Shepard factor                            */
  TSPH_F_D_DENSITY = 18,    /* f64                                           */
  TSPH_F_D_VEL = 19,        /* f64 x3                                        */
  TSPH_F_DENSITY_TMP = 20,  /* f64                                           */
  TSPH_F_V_TMP = 21,        /* f64 x3                                        */
  TSPH_F_CELL = 22,         /* i32 scene-local flattened cell id             */
  TSPH_F_OFFSET = 23,       /* f64 x3 cell-local offset (stored as f32)      */
  TSPH_F_CELL_START = 24,   /* i32 x (cells+1), download only: scene-local
                               particle range of each cell                   */
  TSPH_F_CSPM_L = 25,       /* f64 x9, download only: inverse CSPM_L
                               (row-major 3x3; kernelCorrection 1; zero
                               with 2, whose corrected gradient is 0)       */
  /* colour feed (tsph_assign_value_color / tsph_set_color) */
  TSPH_F_VAL = 26,          /* f64, download only: colour value `val`        */
  TSPH_F_COLOR = 27,        /* f64 x3, download only: colour (stored as f32) */
  TSPH_F_X0 = 28,           /* f64 x3 initial position, indexed by id0 (the
                               scene's original particle order)              */
  TSPH_F_COLOR0 = 29,       /* f64 x3 initial colour (stored as f32),
                               indexed by id0                                */
  TSPH_F_COUNT = 30
};

/* ---- single passes (test hooks; tsph_step runs the fused sequence) ----- */
enum {
  TSPH_PASS_SORT = 0,             /* hash + scan + stable counting sort     */
  TSPH_PASS_SHEPARD = 1,          /* calc_CSPM_f                            */
  TSPH_PASS_INIT_TMP = 2,         /* init_real2tmp                          */
  TSPH_PASS_ONE_STEP = 3,         /* the method's one_step                  */
  TSPH_PASS_LF_HALF = 4,          /* advect_LF_half                         */
  TSPH_PASS_LF_FULL = 5,          /* advect_LF                              */
  TSPH_PASS_ADVECT_POS = 6,       /* advect_pos (+XSPH)                     */
  TSPH_PASS_ADVECT_SOMETHING = 7, /* solver-specific post-update            */
  TSPH_PASS_STEP = 8,             /* one full step(), one launch per pass   */
  TSPH_PASS_SE = 9,               /* advect_SE                              */
  TSPH_PASS_ENFORCE_BOUNDARY = 10, /* enforce_boundary (boundary == 1, and
                                     dynamic rigid particles)               */
  TSPH_PASS_SOLVE_RIGID = 11      /* solve_rigid_body: shape matching       */
};

/* ---- version / errors --------------------------------------------------- */
int tsph_abi_version(void);
const char* tsph_last_error(void);
/* Number of HIP devices visible (0 without a GPU). */
int tsph_device_count(int* n);

/* ---- engine lifecycle --------------------------------------------------- */
/* Engine for scenes of one method/dim.  Capacities bound the sum over
 * scenes of particles and of grid cells.  Replaces the Taichi field
 * allocation of ParticleSystem.__init__ (eng/particle_system.py:12-133) and
 * the ti.init device-memory reservation (run_simulation.py:23), sized once for
 * every scene the engine will hold. */
int tsph_engine_create(int method, int dim, int64_t particle_capacity,
                       int64_t cell_capacity, int max_scenes,
                       void** out_engine);
int tsph_engine_destroy(void* engine);
/* Device bytes held by the engine (for HBM-budget admission). */
int tsph_engine_bytes(void* engine, int64_t* out_bytes);

/* ---- scenes ----------------------------------------------------------- */
/* Reserve a slot and its particle/cell ranges; fields start at zero.
 * Replaces Simulation.__init__ / build_solver (eng/simulation.py:8-22): the
 * solver type and its constants come from desc.  The scene's method must be
 * the engine's and its time integration that of the resident scenes (one
 * fused step sequence per engine): TSPH_ERR_INVALID_ARG otherwise. */
int tsph_add_scene(void* engine, const tsph_scene_desc* desc, int* out_slot);
/* Free a slot (no reference counterpart: the reference runs one scene per
 * process). */
int tsph_remove_scene(void* engine, int slot);
/* Replace / read one particle field of a slot.  Upload replaces
 * ParticleSystem.add_particle / _add_particles (eng/particle_system.py:275-319)
 * and set_id0 (:208-215); download replaces the copy_to_numpy* kernels
 * (:413-457) and the .to_numpy() reads of dump() (:459-545). */
int tsph_upload(void* engine, int slot, int field, const void* host,
                size_t bytes);
int tsph_download(void* engine, int slot, int field, void* host,
                  size_t bytes);

/* ---- stepping ----------------------------------------------------------- */
/* Advance every active scene by nsteps reference steps (SPHBase.step,
 * eng/solver_sph_base.py:41-51: initialize_particle_system, calc_CSPM_f,
 * substep, advect_pos, advect_something, enforce_boundary).  Asynchronous on
 * the engine stream; a scene leaves the stepping set with tsph_remove_scene. */
int tsph_step(void* engine, int nsteps);
/* Wait for the engine stream; reports escaped / NaN particles. */
int tsph_sync(void* engine);
/* Run one pass on every active scene (single-pass parity hooks): the
 * @ti.kernel methods listed with TSPH_PASS_* (eng/particle_system.py:228-258,
 * eng/solver_sph_base.py:63-245, 393-398, 525-597, and the solvers'
 * one_step / advect_something). */
int tsph_run_pass(void* engine, int pass_id);
/* Steps taken by a slot since it was added (the step counter of
 * eng/ui_sim.py's main loop). */
int tsph_scene_steps(void* engine, int slot, int64_t* out_steps);
/* Use a captured hipGraph for tsph_step (1, default) or plain launches (0). */
int tsph_set_graph(void* engine, int enable);
/* Threads per block of the pair kernels and of the per-particle kernels
 * (64, 128 or 256). */
int tsph_set_block(void* engine, int pair_block, int elem_block);
/* Neighbour search mode: 2 (default) records per-step neighbour lists in
 * one stencil walk that reads LDS-staged cell windows and replays them in
 * the later pair passes of the step; 1 does the same with the walk reading
 * global memory; 0 walks the cell stencil in every pass.  All visit the same pairs in the same
 * order (for_all_neighbors, eng/particle_system.py:259-269). */
int tsph_set_neighbor_list(void* engine, int enable);
/* Cell tile of the mode-2 list build: how many candidate particles one
 * wave stages in LDS per row offset (128, 256 (default) or 512; 2, 4 or
 * 8 KiB per wave).  A window that does not fit is walked from global
 * memory instead, so every size records the same lists.  No reference
 * counterpart (a launch shape, like tsph_set_block). */
int tsph_set_cell_tile(void* engine, int stage);

/* ---- reductions --------------------------------------------------------- */
/* FP64 global quantities of one scene (no reference counterpart; the
 * reference reads these from dumped fields on the host, e.g. v_maxmin,
 * eng/particle_system.py:387-397):
 * out[0] surge front = max x over flow particles,
 * out[1] potential energy = sum m * (-g . x),
 * out[2] kinetic energy   = sum 1/2 m |v|^2,
 * out[3] mass             = sum m   (flow particles). */
int tsph_globals(void* engine, int slot, double out[4]);

/* ---- rigid bodies ---------------------------------------------------------
 * Rigid blocks (TSPH_MAT_RIGID particles; ParticleSystem init_block,
 * particle_system.py:150-174) are walls to the flow passes (dummy
 * extrapolation: solver_sph_wc.py:90-103, solver_sph_muI.py:96-110,
 * solver_sph_dp.py:220-231).  A static block keeps d_vel = 0 (WC :125-126,
 * muI :130-132, DP :272-274).  A dynamic block (isDynamic) is registered
 * here with its object id and its rest centre of mass (calc_rigid_rest_m,
 * solver_sph_base.py:467-470, 517-518, the mass-weighted mean of the
 * initial positions): its particles take g minus the flow particles' pair
 * forces (WC :68-71, 105-106; muI :45-46, 111-112; DP :164-165, 233-234),
 * move with the integrator, are put back into shape by solve_rigid_body
 * (solver_sph_base.py:472-510; initial positions from TSPH_F_X0) and are
 * always kept inside the domain (judge_enforce_bdy, :532-538).  At most 4
 * dynamic objects per scene. */
int tsph_set_rigid_object(void* engine, int slot, int object_id, const double rest_cm[3]);
/* (registering a known object id again replaces its rest centre) */

/* ---- colour feed ---------------------------------------------------------
 * The reference's per-frame visual pipeline (eng/ui_sim.py:284-286) without
 * the window, on the device, in the same three calls.  Values are FP64 as
 * the reference's fields (default_fp=f64, run_simulation.py:23) formed from
 * this engine's state; colours are f32 as its colour field.
 *
 * solver.assign_value_color (eng/solver_sph_base.py:721-789): `val` of the
 * colorTitle for every real particle, and for dummy particles when
 * show_bdy; the others hold 0 (the value the reference never overwrites).
 * Title 44 (distance from the initial position) reads TSPH_F_X0; title 101
 * (ti.random) is a counter-based uniform in [0, 1) of (id0, seed), which no
 * check can pin to the reference's generator; unknown titles give 0. */
int tsph_assign_value_color(void* engine, int slot, int color_title, int show_bdy,
                            uint32_t seed);
/* ps.v_maxmin (eng/particle_system.py:387-397): max / min of `val` over flow
 * particles (-inf / +inf with none), then the givenMax / givenMin / fixMax /
 * fixMin rule; the pair is kept for tsph_set_color and returned. */
int tsph_v_maxmin(void* engine, int slot, double given_max, double given_min, int fix_max,
                  int fix_min, double* out_vmax, double* out_vmin);
/* ps.set_color (eng/particle_system.py:399-407): the jet colour
 * (eng/colormap.py:8-48) of (val - vmin) / (vmax - vmin) for the particles
 * of colorGroup (0 flow, 1 real, 2 real and, with show_bdy, dummy); every
 * other particle gets its TSPH_F_COLOR0 colour (the reference leaves its
 * initial colour in place). */
int tsph_set_color(void* engine, int slot, int color_group, int show_bdy);

/* ---- status / timing ---------------------------------------------------- */
/* Device status word: out[0] escaped particles, out[1] NaN particles
 * (counted since the scene was added). */
int tsph_status(void* engine, int slot, int64_t out[2]);
/* Average device time (ms) of the named kernel class over the launches
 * since the last reset, measured with HIP events on the engine stream.
 * name: "pair_fluid" etc.; returns count in *out_launches. */
int tsph_kernel_time(void* engine, const char* name, double* out_ms,
                     int64_t* out_launches);
int tsph_timing(void* engine, int enable);

/* ---- test hooks ----------------------------------------------------------- */
/* Per particle of the slot (current order): number of neighbours
 * (j != i, |x_ij| < r_s; for_all_neighbors, particle_system.py:259-269) and
 * the sum and sum of squares of their TSPH_F_ID0 values. */
int tsph_neighbor_stats(void* engine, int slot, int32_t* count, int64_t* id_sum,
                        int64_t* id_sqsum);
/* Self-check of the pair distance the pair kernels use (v_sqrt_f32 plus a
 * one-ulp residual correction, in place of the correctly rounded sqrtf and
 * its small-argument rescaling): over every finite non-negative float,
 * out[0] = results that differ from sqrtf at arguments >= 2^-96, out[1] =
 * results outside [0, 1e-8) below 2^-96 (where the smoothing kernels are 0
 * for either).  Both 0 means every kernel value equals the one formed
 * from the correctly rounded root, as the reference's |x_ij| (.norm()).
 * Runs on the current device. */
int tsph_check_sqrt(uint64_t out[2]);
/* tsph_neighbor_stats without any allocation (no counterpart in the
 * reference; for callers that cannot allocate device memory: code under
 * HIP graph capture, tight serving loops).  Same outputs for one slot of
 * at most 262,144 particles (TSPH_ERR_INVALID_ARG beyond); its scratch is a
 * fixed 5 MiB device array of the library that every GPU holds once, shared
 * by all calls on that GPU: calls that overlap on one GPU race. */
int tsph_neighbor_stats_noalloc(void* engine, int slot, int32_t* count, int64_t* id_sum,
                                int64_t* id_sqsum);
/* The neighbour lists the pair passes replay (recorded by the last list
 * build: TSPH_PASS_SHEPARD or a step, list mode 1 or 2), per particle of the
 * slot in the current order: the pairs for_all_neighbors
 * (particle_system.py:259-269) visits, in its order, of which each pass
 * reads -- every neighbour of a flow particle, the flow neighbours of a
 * boundary particle, the flow and rigid neighbours of a rigid one.
 * count[i] = entries of particle i, or -1 when it has more neighbours than a
 * list holds (its passes walk the cell stencil); ids / codes = each entry's
 * TSPH_F_ID0 and stencil cell ((dx+1) << 4 | (dy+1) << 2 | (dz+1), d = the
 * neighbour's cell minus i's), concatenated in particle and list order
 * (room for ids_cap entries); *out_total = entries in all lists.
 * TSPH_ERR_INVALID_ARG when no lists match the current order (list mode 0,
 * or no build since the last sort or upload) or ids_cap is too small. */
int tsph_neighbor_lists(void* engine, int slot, int32_t* count, int32_t* ids, int8_t* codes,
                        int64_t ids_cap, int64_t* out_total);

/* ---- single-scene convenience (slot 0 of a one-slot engine) ------------
 * The reference's one-scene ParticleSystem + solver pair
 * (eng/simulation.py:8-22) as one handle. */
int tsph_create(const tsph_scene_desc* desc, void** out_engine);
int tsph_destroy(void* engine);

#ifdef __cplusplus
}
#endif

#endif /* TISPHI_H */
