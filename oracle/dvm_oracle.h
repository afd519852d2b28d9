/*
 * dvm_oracle.h -- CPU oracle of the DVM time step (TEST INFRASTRUCTURE ONLY).
 *
 * A plain-C, serial, FP64 restatement of the reference Fortran hot path, kept
 * in the reference's evaluation order so that it reproduces its rounding.
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may load
 * it, and only as the checker; the product path never calls it.
 *
 * Its entry points mirror include/dvm_hip.h with an `orc_` prefix so that a test
 * drives the oracle and the HIP library with the same config and inputs.
 */
#ifndef DVM_ORACLE_H
#define DVM_ORACLE_H

#include <stdint.h>
#include "../include/dvm_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

int  orc_ctx_create(const dvm_config* cfg, orc_ctx** out);
void orc_ctx_destroy(orc_ctx* ctx);
const char* orc_last_error(const orc_ctx* ctx);
int  orc_upload_f(orc_ctx* ctx, int32_t species, const double* f);
int  orc_download_f(orc_ctx* ctx, int32_t species, double* f);
int  orc_moments_ext(orc_ctx* ctx, double* ext);
int  orc_moments(orc_ctx* ctx, double* props_out);
int  orc_collide(orc_ctx* ctx, int32_t ntime, int64_t* samples_out);
int  orc_convect(orc_ctx* ctx, int32_t species, int32_t ntime);
int  orc_step(orc_ctx* ctx, int32_t ntime);
int  orc_remap(orc_ctx* ctx, int64_t n, const double* uvw, double f_off,
               int32_t* stencil, double* mass);
int  orc_collide_trace(orc_ctx* ctx, int32_t ntime, int32_t cell, int64_t max_records,
                       int32_t* irec, double* drec, int64_t* n_out);
/* dvm_l1: per-cell sum |f - M| / n against the discrete Maxwellian at the
   cell's moments, scaled to sum M = n (-1 for an empty or cold cell). */
int  orc_l1(orc_ctx* ctx, double* l1);
/* Parity mode of dvm_set_uniform_stream: uniforms from an explicit list. */
int  orc_set_uniform_stream(orc_ctx* ctx, const double* u, int64_t n);
int  orc_uniform_stream_position(orc_ctx* ctx, int64_t* consumed);
int  orc_set_wall_heat_flux_freq(orc_ctx* ctx, int32_t every);
int  orc_wall_heat_flux(orc_ctx* ctx, double* q, int32_t* ntime_out);
/* Collision phase restricted to cells [c0, c1) (CPU-baseline sampling). */
int  orc_collide_cells(orc_ctx* ctx, int32_t ntime, int32_t c0, int32_t c1);

/* Test hooks. orc_set_f_keep_cache replaces f (and invalidates the moments)
   without resetting the equilibrium cache, so that a test can move the
   cell state between collisions the way convection does; orc_grid_info
   returns the derived integers of the velocity grid in the layout of the
   reference's grid fields (num_points, i_zero, i_neg, i_pos, j_zero, j_neg,
   j_pos, k_zero, k_neg, k_pos, num_points_x) and beta3_min. */
int  orc_set_f_keep_cache(orc_ctx* ctx, const double* f);
int  orc_grid_info(orc_ctx* ctx, int32_t* ints /* [11] */, double* beta3_min);

/* Building blocks exposed for unit tests. */
void orc_uniform_axis(int32_t n, double vmin, double vmax, double* out);
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
int32_t orc_referenced_search(const double* cumul, int32_t n, const int32_t* left,
                              const int32_t* right, double u);
void orc_build_buckets(const double* cumul, int32_t n, int32_t* left, int32_t* right);
double orc_maxwellian(double x, double y, double z, double coeff, double mass, double u,
                      double v, double w, double temp);
double orc_cross_section(int32_t coeff_kind, double m_red, double omega, double d, double visc,
                         double tv);
/* CDF arrays of one cell after the last collide (or computed on demand):
   eq_cumul[n+1], neq_cumul[n+1], eq_left/right[1000], neq_left/right[1000]. */
int  orc_cell_cdf(orc_ctx* ctx, int32_t cell, double* eq_cumul, double* neq_cumul,
                  int32_t* eq_left, int32_t* eq_right, int32_t* neq_left, int32_t* neq_right,
                  int32_t* sample_counts /* [3] */);

#ifdef __cplusplus
}
#endif
#endif
