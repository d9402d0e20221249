/*
 * oracle.h — C interface of the CPU oracle (test infrastructure only).
 *
 * The oracle is a C++17/OpenMP restatement of the reference tiSPHi solver
 * (eng/particle_system.py, eng/solver_sph_*.py) used to check the HIP
 * engine and to time a CPU baseline.  It is never linked into, loaded by,
 * or called from the product library.
 *
 * Field ids, pass ids and the parameter block have the same numeric values
 * and layout as the engine's (include/tisphi.h) so that one Python helper
 * can feed both the same state.
 */
#ifndef TSPH_ORACLE_H
#define TSPH_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
  int32_t dim, method, boundary, kernel, kernel_corr, time_integration, xsph;
  int32_t grid_num[3];
  int64_t n_particles;
  double vdomain_start[3];
  double domain_start[3];
  double domain_end[3];
  double grid_size, particle_radius, particle_diameter, smoothing_len,
      support_radius;
  double gravity[3];
  double dt;
  double density0, viscosity, stiffness, exponent;
  double cohesion, friction, young, poisson, dilatancy, vsound;
} orc_params;

/* precision: 32 (pair arithmetic and state in float, DP constitutive update
 * and reductions in double — the engine's mixed mode) or 64 (all double). */
void* orc_create(const orc_params* p, int precision);
void orc_destroy(void* h);
int orc_set(void* h, int field, const void* data);
int orc_get(void* h, int field, void* data);
int orc_pass(void* h, int pass_id);
int orc_step(void* h, int nsteps);
int orc_globals(void* h, double out[4]);
/* Neighbour list of particle i in the current order (strict r < r_s test
 * on the stored representation); returns the count, writes up to cap. */
int64_t orc_neighbors(void* h, int64_t i, int64_t* out, int64_t cap);
/* Every particle's neighbour list at once: counts[i]; id0 and stencil cell
 * code ((dx+1) << 4 | (dy+1) << 2 | (dz+1)) of each neighbour, in particle
 * and walk order.  Returns the total entries; writes ids / codes only when
 * cap holds them. */
int64_t orc_neighbor_csr(void* h, int64_t* counts, int32_t* ids, int8_t* codes, int64_t cap);
int64_t orc_escaped(void* h);
/* Dynamic rigid object (shape matching) with its rest centre of mass. */
int orc_set_rigid_object(void* h, int obj, const double cm[3]);
/* Orthogonal polar factor R of a row-major 3x3 A (dim 2: the xy block). */
int orc_polar(const double* A, int dim, double* R);
void orc_set_threads(int n);
/* Function-level hooks for closed-form tests. */
int orc_kernel_eval(const orc_params* p, int precision, const double* r, int64_t n, double* W,
                    double* gradfac);
int orc_dp_adapt(const orc_params* p, double* st6, int64_t n, int32_t* flags);
int orc_dp_rate(const orc_params* p, const double* L9, const double* st6, int64_t n, double* ds6,
                double* dseq, double* dseqp);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
