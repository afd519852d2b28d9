/*
 * dvm_hip.h -- C-ABI of the MI355X (gfx950) DVM time step.
 *
 * The library replaces the hot path of the reference's serial Fortran time step:
 *   collide  (variance-reduced Monte Carlo collision integral + conservative remap)
 *   convect  (diffuse / specular / neighbour ghosts, boundary-cell and interior upwind sweep)
 *   moments  (per-cell density, velocity, temperature)
 * behind plain C entry points (no C++ or torch types), callable from Fortran
 * through ISO_C_BINDING and from Python through ctypes.
 *
 * Conventions
 *   - Every entry point returns 0 on success and a negative DVM_E_* code on
 *     failure; dvm_last_error() then holds a message (the reference's own
 *     convention is `write(*,*) ...; stop`, which the Fortran wrappers restore).
 *   - Distributions are exchanged as one contiguous FP64 slab
 *     f[cell][node], cell = ix + nx*iy (ix fastest, as phi(nx,ny) in Fortran),
 *     node = i + nvx*(j + nvy*k) (i fastest, as value(i,j,k) in Fortran).
 *   - Velocity-grid indices are 0-based on both sides of the boundary.
 *   - One host thread per context; each context owns one HIP stream.
 */
#ifndef DVM_HIP_H
#define DVM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------- */
#define DVM_OK            0
#define DVM_E_ARG        -1   /* invalid argument / configuration            */
#define DVM_E_HIP        -2   /* HIP runtime error                           */
#define DVM_E_NOMEM      -3   /* device allocation failed                    */
#define DVM_E_STATE      -4   /* call out of order (e.g. collide before f)   */
#define DVM_E_REMAP      -5   /* remap found no exterior point (ext flag 0)  */

/* ---- enumerations (values equal the reference's deck integers) --------- */
/* collision solver, collisions/collision_integral.f90:35-38 */
#define DVM_COLL_NONE     0
#define DVM_COLL_FMC      2   /* full Monte Carlo, collisions/full_mc.f90:110-641 */
#define DVM_COLL_VR       3
/* side boundary types, conditions/boundary_conditions.f90:56-59 */
#define DVM_BC_WALL       1
#define DVM_BC_INOUT      2
#define DVM_BC_ZEROGRAD   3
#define DVM_BC_COUETTE    4
/* cross-section coefficient, solver/scaling.f90:34-36 */
#define DVM_CS_CLARKE     1
#define DVM_CS_BIRD       2
#define DVM_CS_VKI        3
/* physical-space dimension of the run */
#define DVM_DIM_0D        0   /* space-homogeneous: one cell, no convection   */
#define DVM_DIM_1D        1   /* 1-D in x, boundary semantics of the zone-grid 1-D solver */
#define DVM_DIM_2D        2   /* 2-D in x,y, boundary semantics of the trunk solver       */
/* collision arithmetic */
#define DVM_PREC_FP64     0   /* everything FP64 (parity path)                */
#define DVM_PREC_FP32     1   /* FP32 collision arithmetic, FP64 remap weights,
                                 moments and conservation correction          */
#define DVM_PREC_FP32_REMAP32 2 /* as DVM_PREC_FP32, with the remap (nearest
                                 nodes and weights) in FP32 too on equally
                                 spaced axes (FP64 remap on other axes)       */
/* sample-count rule */
#define DVM_SAMPLES_REFERENCE 0 /* nint(dt*T^(2/3)*nA*nB/(2 Kn rms^2 beta3)) per integral */
#define DVM_SAMPLES_EXPLICIT  1 /* samples_per_cell split over the 3 integrals
                                   in the reference's nA*nB proportions       */
/* 1-D sweep order (dvm_config.conv_order) */
#define DVM_CONV_UPWIND1  1
#define DVM_CONV_MINMOD2  2

/* ---- per-cell property record returned by dvm_moments ------------------ */
#define DVM_P_DENS        0   /* number density                 */
#define DVM_P_U           1   /* x velocity                     */
#define DVM_P_V           2   /* y velocity                     */
#define DVM_P_W           3   /* z velocity                     */
#define DVM_P_TRTEMP      4   /* translational temperature      */
#define DVM_P_TEMP        5   /* species (overall) temperature  */
#define DVM_P_PRESSURE    6   /* 0.5 * n * T                    */
#define DVM_P_AVGSPEED    7   /* sum f |v| / n                  */
#define DVM_NPROP         8

/* ---- per-cell output moments returned by dvm_moments_ext --------------- */
#define DVM_X_QX          0   /* heat flux x                    */
#define DVM_X_QY          1
#define DVM_X_QZ          2
#define DVM_X_TXX         3   /* viscous stress (pressure on the diagonal) */
#define DVM_X_TYY         4
#define DVM_X_TZZ         5
#define DVM_X_TXY         6
#define DVM_X_TXZ         7
#define DVM_X_TYZ         8
#define DVM_X_ENTROPY     9   /* -sum |f| log |f|               */
#define DVM_NEXT         10

/* Wall record indices: wall[side][...] with side 0 left, 1 right, 2 bottom, 3 top */
#define DVM_W_DENS 0
#define DVM_W_U    1
#define DVM_W_V    2
#define DVM_W_W    3
#define DVM_W_TEMP 4

typedef struct dvm_config {
  /* physical space (physical_grid.f90:42-55) */
  int32_t dim;              /* DVM_DIM_*                                        */
  int32_t nx, ny;           /* nx_space, ny_space (ny = 1 for 0-D and 1-D)      */
  double  dx, dy;
  /* velocity grid axes, copied at create (velocity_grid.f90:271-350) */
  int32_t nvx, nvy, nvz;
  const double* vx;         /* [nvx] ascending                                  */
  const double* vy;         /* [nvy]                                            */
  const double* vz;         /* [nvz]                                            */
  /* time stepping and scaling */
  double  dt;               /* get_deltat                                       */
  double  kn;               /* get_knudsen_number                               */
  /* species (single monatomic species; species_data.f90:46-78) */
  double  mass, diam, omega, alpha, viscosity, t_viscosity;
  int32_t csection_coeff;   /* DVM_CS_*                                         */
  /* collision integral */
  int32_t coll_solver;      /* DVM_COLL_NONE, DVM_COLL_FMC or DVM_COLL_VR       */
  double  cdf_tolerance;    /* set_cdf_tolerance                                */
  int32_t num_repl;         /* set_num_repl_locations (NRepl)                   */
  double  coln_rms;         /* ColnRMS                                          */
  int32_t sample_mode;      /* DVM_SAMPLES_*                                    */
  int64_t samples_per_cell; /* for DVM_SAMPLES_EXPLICIT                         */
  uint64_t seed;            /* Philox4x32-10 key                                */
  int32_t precision;        /* DVM_PREC_*                                       */
  /* convection and boundaries */
  int32_t enable_convection;
  int32_t bc_left, bc_right;/* DVM_BC_* (bottom/top are walls in 2-D, as trunk) */
  double  acc_coeff;        /* accommodation: 1 diffuse, 0 specular             */
  double  wall[4][5];       /* [side][DVM_W_*]                                  */
  /* 1-D sweep order: 1 the reference's donor-cell upwind (0 is taken as 1);
     2 limited second-order upwind (minmod slopes), which the reference does
     not have */
  int32_t conv_order;
} dvm_config;

typedef struct dvm_ctx dvm_ctx;

/* Creates a context on the current HIP device and allocates the resident
   distribution, CDF and moment buffers.
   Replaces: initialize_solver (solver/dvm_solver.f90:92-229) allocation of
   phi/delta_phi/properties, create_velocity_grid (grid/velocity_grid.f90:271),
   create_collision_integral / create_variance_reduction
   (collisions/collision_integral.f90:98, collisions/variance_reduction.f90:71). */
int dvm_ctx_create(const dvm_config* cfg, dvm_ctx** out);

/* Frees every device buffer of the context.
   Replaces: destroy_solver (solver/dvm_solver.f90:371-414). */
void dvm_ctx_destroy(dvm_ctx* ctx);

/* Message of the last failing call on this context (or of the last failing
   dvm_ctx_create when ctx is NULL). */
const char* dvm_last_error(const dvm_ctx* ctx);

/* Host -> device copy of f[ncell][nnode]; also resets the equilibrium cache
   as create_variance_reduction does (variance_reduction.f90:119-125).
   Replaces: the host-side phi(:,:,:)%value fill of initialize_dist_func
   (conditions/initial_conditions.f90:212). */
int dvm_upload_f(dvm_ctx* ctx, int32_t species, const double* f);

/* Device -> host copy of f[ncell][nnode]. */
int dvm_download_f(dvm_ctx* ctx, int32_t species, double* f);

/* Per-cell moments of the resident f, kept on the device for the next
   collide call and copied to props_out[ncell][DVM_NPROP] when non-NULL.
   Replaces: calculate_physical_properties (dist_funcs/physical_props.f90:285)
   called per cell at solver/dvm_solver.f90:277,336. */
int dvm_moments(dvm_ctx* ctx, double* props_out);

/* Output moments of the resident f per cell, ext[ncell][DVM_NEXT]: heat flux
   (compute_heat_flux, dist_funcs/physical_props.f90:1914-1986), viscous
   stress (compute_viscous_stress, :1988-2065) and entropy (compute_entropy,
   :1871-1912), from f and its moments (which this call refreshes). Used by
   the writers of input_output/visualization.f90:1591-1716 (get_data_to_output
   :2296-2633); not on the time-step path. */
int dvm_moments_ext(dvm_ctx* ctx, double* ext);

/* Wall heat fluxes of a 2-D grid, replacing the four wall files written
   inside set_boundary_nodes (conditions/boundary_conditions.f90:872-923,
   1069-1115, 1262-1307, 1450-1495) at the interval of
   set_wall_heat_flux_dump_freq (:108-116). During the convection of every
   step whose ntime is a multiple of `every` (0: never), each cell along a
   diffuse / specular wall records
     q = m/2 * ( sum over the incoming half of v_n f |c - u_cell|^2
               + sum over the outgoing half of v_n g |c - u_wall|^2 ),
   f the post-collision distribution, g the wall ghost, u_cell the bulk
   velocity of the moments the step started with, v_n the wall-normal
   velocity component (x for left / right, y for bottom / top). */
int dvm_set_wall_heat_flux_freq(dvm_ctx* ctx, int32_t every);
/* The last recorded fluxes q[side][pos] (side 0 left, 1 right, 2 bottom,
   3 top; pos the cell index along the wall, q holds 4 * max(nx, ny) values,
   sides that are not walls hold 0) and the step they belong to (-1: none). */
int dvm_wall_heat_flux(dvm_ctx* ctx, double* q, int32_t* ntime_out);

/* One collision phase over every cell at time step ntime (>= 1): CDF build with
   the cached equilibrium, the three sampled VR integrals, phi := delta_phi.
   samples_out[ncell] (may be NULL) receives tot_colls per cell.
   Replaces: compute_collision_integral (collisions/collision_integral.f90:163)
   called per cell at solver/dvm_solver.f90:305-315. */
int dvm_collide(dvm_ctx* ctx, int32_t ntime, int64_t* samples_out);

/* One convection phase: boundary ghosts, boundary-cell and interior upwind
   update, phi := phi_conv.
   Replaces: compute_convection (convection/convection.f90:69) called at
   solver/dvm_solver.f90:318-329 (2-D) and the 1-D first_order_scheme +
   apply_boundary_conditions of the zone-grid solver
   (code_versions/grid_ref_code/src/convection/convection.f90:111-256). */
int dvm_convect(dvm_ctx* ctx, int32_t species, int32_t ntime);

/* Fused step: collide -> convect -> moments, the body of the time loop
   (solver/dvm_solver.f90:299-357) without its screen and file output. */
int dvm_step(dvm_ctx* ctx, int32_t ntime);

/* Remaps n off-grid velocities uvw[n][3] with weight f_off onto the context's
   velocity grid on the device: stencil[n][7] = (o_i,o_j,o_k,a,b,c,flag),
   mass[n][7] = (f_o,f_ix,f_iy,f_iz,f_ex,f_ey,f_ez).
   Replaces: perform_remapping (collisions/remapping.f90:55-90). */
int dvm_remap(dvm_ctx* ctx, int64_t n, const double* uvw, double f_off,
              int32_t* stencil, double* mass);

/* Parity trace of one cell's collision phase at time step ntime, computed on
   the device from the resident f without modifying it. For each sample s of
   integral t (t = 1..3, in order) one record of 32 int32:
     [0] integral, [1] sample index, [2] node A, [3] node B, [4] skipped (g<1e-13),
     [5..11] stencil of the first post-collision velocity (repl location 1),
     [12..18] stencil of the second, [19..31] zero (reserved),
   and 16 doubles: [0] kin_depl A, [1] kin_depl B, [2..8] mass of remap 1,
   [9..15] mass of remap 2. max_records bounds the output; n_out receives the
   number of samples of the cell. */
int dvm_collide_trace(dvm_ctx* ctx, int32_t ntime, int32_t cell, int64_t max_records,
                      int32_t* irec, double* drec, int64_t* n_out);

/* Distance of each cell's f from equilibrium, l1[ncell] = sum |f - M| / n,
   with M the discrete Maxwellian (times beta^3) at the cell's own density,
   velocity and translational temperature, scaled so that sum M = n (-1 for an
   empty or cold cell); the same measure as dvm_batch_l1. No counterpart in
   the reference, which has no convergence measure: it is the per-step output
   of space-homogeneous relaxation runs and their stopping rule. */
int dvm_l1(dvm_ctx* ctx, double* l1);

/* dvm_l1 for callers that cannot allocate device memory (code under HIP graph
   capture, tight serving and relaxation loops, harnesses that forbid
   hipMalloc): the same results, staged in a file-scope device scratch array
   of 2^20 doubles that the library holds on each GPU it runs on until the
   process ends, instead of a buffer allocated per call. Refuses contexts with
   more than 2^20 cells (DVM_E_ARG). Calls that overlap on one GPU share the
   array and race. No counterpart in the reference. */
int dvm_l1_noalloc(dvm_ctx* ctx, double* l1);

/* Parity mode: the collision phase takes its uniforms from the explicit list
   u[0..n) instead of the counter-based generator, so that one list can be fed
   to this library and to a CPU implementation. Each collision reads the list
   from a cursor on: cells in order, in each cell its samples in order
   (integral 1, 2, 3), 2 + 2*NRepl values per sample: uA, uB (the two partner
   picks), then (phi, theta) of each replenish location, the order in which the
   reference draws them (variance_reduction.f90:686-816,
   replenishing_collisions.f90:134-354). The cursor starts at 0 and advances by
   what each collision read; a collision that needs more values than are left
   fails with DVM_E_ARG before it changes anything. dvm_collide_trace reads the
   traced cell's samples from the cursor without moving it. Values must lie in
   [0, 1); n = 0 returns to the counter-based generator. In this mode each
   collision synchronises the host with the device once.
   Replaces: get_rand (utils/random_number_generation.f90:140-162), the
   reference's sequential stream, for parity runs. */
int dvm_set_uniform_stream(dvm_ctx* ctx, const double* u, int64_t n);
/* Values of the list read so far (the cursor). */
int dvm_uniform_stream_position(dvm_ctx* ctx, int64_t* consumed);

/* Device time per phase, measured with HIP events on the context's stream
   around every phase that dvm_step / dvm_collide / dvm_convect / dvm_moments
   enqueue. Phases: 0 CDF build, 1 collision prefill (only when a cell's slice
   exceeds LDS), 2 collision sampling (the collide kernel alone), 3 FP32
   conservation correction, 4 wall ghosts, 5 upwind sweep, 6 moments.
   dvm_timing_reset waits for pending work and zeroes the totals;
   dvm_timing_read waits for pending work and returns the totals in ms since
   the last reset and the number of calls they cover.
   (No counterpart in the reference, which times phases with cpu_time at
   solver/dvm_solver.f90:294-366.) */
#define DVM_NPHASE 7
int dvm_timing_reset(dvm_ctx* ctx);
int dvm_timing_read(dvm_ctx* ctx, double* phase_ms /* [DVM_NPHASE] */, int64_t* calls);

/* ------------------------------------------------------------------------
   Batched solves. Many independent space-homogeneous (dim 0), 1-D and 2-D solves,
   each with its own velocity grid, cell count, dt, Kn, sample count, seed and
   step counter, resident in shared device pools and advanced together by one
   replayed HIP graph per step. Each instance behaves as the loop body of
   solve() (solver/dvm_solver.f90:299-357) for its own deck; the batch itself
   has no counterpart in the reference, which runs one solve per process.
   The Philox samples of an instance depend only on its seed, its own cell
   indices and its own step count, so an instance's trajectory is the one a
   dvm_ctx with the same config produces, whatever else is resident. */
typedef struct dvm_batch dvm_batch;

/* Creates the pools: node_capacity velocity nodes (summed over the cells of
   all resident instances), cell_capacity cells, max_slots instances. The
   species / collision fields of `species` (mass, diam, omega, alpha,
   csection_coeff, viscosity, t_viscosity, coll_solver = DVM_COLL_VR,
   cdf_tolerance, num_repl, coln_rms, sample_mode, precision) are shared by
   every instance admitted later. */
int dvm_batch_create(const dvm_config* species, int32_t max_slots, int64_t node_capacity,
                     int32_t cell_capacity, dvm_batch** out);
void dvm_batch_destroy(dvm_batch* b);
/* Message of the last failing call on b (or of the last failing
   dvm_batch_create / dvm_batch_footprint when b is NULL). */
const char* dvm_batch_last_error(const dvm_batch* b);

/* Pool nodes and HBM bytes an instance would occupy (admission by footprint). */
int dvm_batch_footprint(const dvm_config* inst, int32_t precision, int64_t* nodes, int64_t* bytes);

/* Admits one instance (dim 0, 1 or 2; its species fields must equal the batch's;
   a 2-D instance brings its own walls, swept by per-slot launches inside the
   step's graph)
   with initial f0[nx][nnode], computes its moments and returns its slot.
   DVM_E_NOMEM with a message when no slot or pool range is free: the caller
   waits for a retirement and tries again.
   Replaces, per instance: initialize_solver + initialize_dist_func
   (solver/dvm_solver.f90:92-229). */
int dvm_batch_admit(dvm_batch* b, const dvm_config* inst, const double* f0, int32_t* slot);
/* Releases a slot's pool ranges (waits for queued steps first). */
int dvm_batch_retire(dvm_batch* b, int32_t slot);

/* Enqueues nsteps steps of every resident instance (no host wait). */
int dvm_batch_step(dvm_batch* b, int32_t nsteps);
/* Waits for queued steps; reports a remap failure flagged by any of them. */
int dvm_batch_synchronize(dvm_batch* b);

/* Current f[nx][nnode] / moments [nx][DVM_NPROP] of one instance. */
int dvm_batch_download_f(dvm_batch* b, int32_t slot, double* f);
int dvm_batch_moments(dvm_batch* b, int32_t slot, double* props);
/* l1[max_slots]: for each resident space-homogeneous instance
   sum |f - M| / n, M the discrete Maxwellian (times beta^3) at the
   instance's own n, u, T scaled to sum M = n; -1 for other slots. A
   convergence measure of this library; the reference has none. */
int dvm_batch_l1(dvm_batch* b, double* l1);
/* Steps enqueued since admission, pool nodes, cells, and whether the
   instance's cells accumulate in LDS (1) or with global atomics (0). */
int dvm_batch_slot_info(dvm_batch* b, int32_t slot, int32_t* steps, int64_t* nodes, int32_t* ncell,
                        int32_t* in_lds);
int dvm_batch_capacity(dvm_batch* b, int64_t* free_nodes, int64_t* largest_free, int32_t* free_cells,
                       int32_t* free_slots);

/* Collision tiles: lanes per workgroup (256, 512, 768, 1024), samples per lane
   picked in lockstep (1, 2) and samples per work item (a cell's samples are
   split into items of that many, each accumulating in its own LDS window and
   adding it to the cell atomically; 0, the default, sizes items from the
   resident load: about four per CU, at least 2^15 samples).
   dvm_batch_get_tiles returns lanes, pick_k, the item size in use and the
   persistent grid size. */
int dvm_batch_set_tiles(dvm_batch* b, int32_t lanes, int32_t pick_k, int32_t item_samples);
int dvm_batch_get_tiles(dvm_batch* b, int32_t* tiles /* [4] */);
/* Autotuner: times `steps` replays of the resident set for each candidate
   tile, restoring the resident state before each candidate and at the end,
   keeps the fastest (tiles_out[3], item size 0 = sized from the load) and
   its ms per step. */
int dvm_batch_tune(dvm_batch* b, int32_t steps, int32_t* tiles_out, double* ms_out);

#ifdef __cplusplus
}
#endif

#endif /* DVM_HIP_H */
