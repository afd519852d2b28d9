/*
 * dvm_oracle.c -- CPU oracle of the DVM time step (TEST INFRASTRUCTURE ONLY).
 *
 * Serial FP64 C restatement of the reference's hot path, in its evaluation
 * order (left-to-right Fortran expression order, no FP contraction: build with
 * -ffp-contract=off). Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline may load it, as the checker; the product never calls it.
 *
 * Reference files followed (paths relative to software/trunk/src unless noted):
 *   utils/constants.f90:30            pi = 3.1459265358979323 (typo kept)
 *   utils/math_utils.f90:38-60        isgn / dsgn
 *   utils/sort_and_search.f90:40-86   referenced_search
 *   utils/sort_and_search.f90:130-183 binary_search_with_first_guess
 *   grid/velocity_grid.f90:271-712    grid, beta, zero points, global2local_map
 *   dist_funcs/dist_func.f90:317-344  compute_maxwellian
 *   dist_funcs/physical_props.f90:285-485,879-1087,1459-1507 moments
 *   solver/scaling.f90:163-201        calculate_cross_section
 *   collisions/collision_utilities.f90:106-352 pair counts, partners, post-collision velocities
 *   collisions/variance_reduction.f90:234-1285 CDF build and the three VR integrals
 *   collisions/replenishing_collisions.f90:134-434 elastic replenishment
 *   collisions/remapping.f90:55-438   conservative 7-point remap
 *   convection/convection.f90:69-276  2-D interior upwind
 *   conditions/boundary_conditions.f90:537-1656,1852-1929 2-D ghosts and boundary-cell update
 *   code_versions/grid_ref_code/src/convection/convection.f90:111-256 1-D upwind
 *   code_versions/grid_ref_code/src/conditions/boundary_conditions.f90:319-444,783-973 1-D BCs
 *
 * The reference draws uniforms from one sequential GSL stream; this oracle,
 * like the HIP library, draws them from counter-based Philox4x32-10 keyed by
 * the seed with counter (sample, integral | slot<<2, cell, ntime): block slot 0
 * gives (u_A, u_B, u_phi, u_theta) of replenish location 1, block slot r gives
 * (u_phi, u_theta) of location r >= 2. orc_set_uniform_stream replaces them by
 * an explicit list read in the reference's draw order (cells in order, samples
 * in order, u_A, u_B, then u_phi, u_theta per location), the parity mode that
 * feeds one list to both codes.
 *
 * orc_l1 (the distance of f from its equilibrium Maxwellian) and the
 * limited second-order 1-D sweep have no counterpart in the reference; they
 * are pinned by invariants and closed forms in tests/.
 */
#include "dvm_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define PI_REF 3.1459265358979323     /* constants.f90:30, the reference's value */
#define DOUBLE_TOL 1.0e-13            /* constants.f90:52 */
#define NBUCKET 1000

/* ------------------------------------------------------------------ grid */
typedef struct {
  int n[3];
  int num_points;
  double* ax[3];     /* axis coordinates */
  double* beta[3];   /* quadrature half-gaps */
  double vmin[3], vmax[3], factor[3];
  int zero[3], neg[3], pos[3];
  double beta3_min;
} ogrid;

struct orc_ctx {
  dvm_config cfg;
  ogrid g;
  int ncell;
  double* f;          /* [ncell][N] phi        */
  double* fd;         /* [ncell][N] delta_phi  */
  double* props;      /* [ncell][DVM_NPROP]    */
  int props_valid;
  int f_valid;
  /* per-cell equilibrium cache (cdf_eq of variance_reduction.f90:36) */
  double* eq_df;      /* [ncell][N+1] */
  double* eq_par;     /* [ncell][5] dens,u,v,w,temp */
  /* scratch for one cell (cdf_neq is a global scratch in the reference) */
  double* eqc;        /* [N+1] */
  double* neqc;       /* [N+1] */
  int32_t eql[NBUCKET], eqr[NBUCKET], neql[NBUCKET], neqr[NBUCKET];
  double sigma, vhs_exponent;
  /* wall heat fluxes (set_wall_heat_flux_dump_freq, boundary_conditions.f90:108) */
  int heat_every, heat_on, heat_ntime;
  double* heat;       /* [4][max(nx,ny)] */
  /* explicit uniform list (dvm_set_uniform_stream): su_n values, the next
     collision reads from su_cursor on; su_base is the current cell's offset */
  double* su;
  int64_t su_n, su_cursor, su_base;
  char err[256];
};

static int set_err(orc_ctx* c, int code, const char* fmt, ...) {
  if (c) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(c->err, sizeof c->err, fmt, ap);
    va_end(ap);
  }
  return code;
}

const char* orc_last_error(const orc_ctx* ctx) { return ctx ? ctx->err : "null context"; }

/* Fortran nint: round half away from zero. */
static long long fnint(double x) { return llround(x); }
/* isgn (math_utils.f90:38): nint(sign(1.0d0, x)); sign of -0.0 is -1. */
static int isgn(double x) { return signbit(x) ? -1 : 1; }
/* dsgn (math_utils.f90:50) */
static double dsgn(double x) { return signbit(x) ? -1.0 : 1.0; }
/* Fortran integer sign(1, k): +1 for k >= 0. */
static int isign1(int k) { return k >= 0 ? 1 : -1; }

/* velocity_grid.f90:636-648 generate_uniform_grid */
void orc_uniform_axis(int32_t n, double vmin, double vmax, double* out) {
  for (int i = 0; i < n; ++i) out[i] = vmin + (double)i * (vmax - vmin) / (double)(n - 1);
}

/* velocity_grid.f90:653-689 find_zero_point */
static void find_zero_point(const double* a, int n, int* zero, int* pos, int* neg) {
  int mn = 0, mx = n - 1;
  if (a[mn] > 0.0) { *zero = -1; *pos = mn; *neg = -1; return; }
  for (int i = mn + 1; i <= mx; ++i) {
    if (a[i] > 0.0) {
      *zero = i - 1; *pos = i; *neg = i - 1;
      if (a[i - 1] > -DOUBLE_TOL) *neg = i - 2;
      return;
    }
  }
  *zero = mx + 1; *pos = mx + 1; *neg = mx;
}

static int grid_setup(ogrid* g, const dvm_config* cfg) {
  const double* src[3] = {cfg->vx, cfg->vy, cfg->vz};
  int n[3] = {cfg->nvx, cfg->nvy, cfg->nvz};
  for (int d = 0; d < 3; ++d) {
    if (n[d] < 2 || !src[d]) return -1;
    g->n[d] = n[d];
    g->ax[d] = (double*)malloc(sizeof(double) * n[d]);
    g->beta[d] = (double*)malloc(sizeof(double) * n[d]);
    memcpy(g->ax[d], src[d], sizeof(double) * n[d]);
    for (int i = 1; i < n[d]; ++i)
      if (!(g->ax[d][i] > g->ax[d][i - 1])) return -1;
    g->vmin[d] = g->ax[d][0];
    g->vmax[d] = g->ax[d][n[d] - 1];
    /* velocity_grid.f90:296-312 */
    g->factor[d] = (double)(n[d] - 1) / (g->vmax[d] - g->vmin[d]);
    /* compute_beta, velocity_grid.f90:486-540 */
    double* a = g->ax[d];
    double* b = g->beta[d];
    b[0] = a[1] - a[0];
    b[n[d] - 1] = a[n[d] - 1] - a[n[d] - 2];
    for (int i = 1; i < n[d] - 1; ++i) b[i] = (a[i + 1] - a[i]) / 2.0 + (a[i] - a[i - 1]) / 2.0;
    find_zero_point(a, n[d], &g->zero[d], &g->pos[d], &g->neg[d]);
  }
  g->num_points = n[0] * n[1] * n[2];
  /* find_min_max_beta3, velocity_grid.f90:542-597 */
  double b3min = g->beta[0][0] * g->beta[1][0] * g->beta[2][0];
  for (int k = 0; k < n[2]; ++k)
    for (int j = 0; j < n[1]; ++j)
      for (int i = 0; i < n[0]; ++i) {
        double b3 = g->beta[0][i] * g->beta[1][j] * g->beta[2][k];
        if (b3 < b3min) b3min = b3;
      }
  g->beta3_min = b3min;
  return 0;
}

/* --------------------------------------------------------------- search */
/* sort_and_search.f90:130-183; the loop is bounded (the reference would spin
   forever on an unbracketed value; no valid input reaches that bound). */
static int bsearch_first_guess(int index, double val, const double* a, int mn, int mx) {
  if (a[mn] > val) return mn;
  if (a[mx] < val) return mx;
  if (index < mn) index = mn;   /* guess outside the axis: the result does not depend on it */
  if (index > mx) index = mx;
  int left = mn, right = mx;
  if (fabs(val - a[index]) < DOUBLE_TOL) return index;
  if (index == right) index = index - 1;
  for (int it = 0; it < 128; ++it) {
    if (val > a[index] && val <= a[index + 1]) {
      if (fabs(val - a[index]) > fabs(val - a[index + 1])) index = index + 1;
      return index;
    }
    if (val > a[index]) left = index; else right = index;
    index = left + (right - left) / 2;
  }
  return index;
}

/* sort_and_search.f90:40-86 */
int32_t orc_referenced_search(const double* cumul, int32_t n, const int32_t* left_tab,
                              const int32_t* right_tab, double u) {
  if (u < DOUBLE_TOL) return 0;
  int i = (int)ceil(u * 1000.0);
  if (i < 1) i = 1;
  if (i > NBUCKET) i = NBUCKET;
  int left = left_tab[i - 1], right = right_tab[i - 1];
  int index = left + (right - left) / 2;
  if (left == right) return index;
  double adj = u * cumul[n];
  for (int it = 0; it < 128; ++it) {
    if (adj > cumul[index] && adj <= cumul[index + 1]) return index;
    if (adj > cumul[index]) left = index; else right = index;
    index = left + (right - left) / 2;
  }
  return index;
}

/* variance_reduction.f90:468-493 bucket tables (l never reads past cumul[n]). */
void orc_build_buckets(const double* cumul, int32_t n, int32_t* left, int32_t* right) {
  int l = 1;
  double mx = 0.001 * cumul[n];
  for (int ref = 1; ref <= NBUCKET; ++ref) {
    left[ref - 1] = l - 1;
    while (l <= n && (double)ref * mx > cumul[l]) {
      l = l + 1;
    }
    right[ref - 1] = l;
  }
  right[NBUCKET - 1] = n;
}

/* ---------------------------------------------------------------- philox */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Four uniforms in [0,1) with 32 random bits each (x * 2^-32, the resolution
   of the reference's gsl_rng_uniform on mt19937) from one Philox4x32-10 block. */
static void uniforms4(uint64_t seed, uint32_t sample, uint32_t tag, uint32_t cell, uint32_t ntime,
                      double u[4]) {
  uint32_t ctr[4] = {sample, tag, cell, ntime};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t o[4];
  orc_philox4x32_10(ctr, key, o);
  for (int q = 0; q < 4; ++q) u[q] = (double)o[q] * 0x1.0p-32;
}

/* ----------------------------------------------------------------- remap */
/* remapping.f90:396-438 find_boundary_stencil */
static void find_boundary_stencil(int* sa, int* sb, int* sc, double abc, int oi, int oj, int ok,
                                  const double* X, const double* Y, const double* Z,
                                  int imin, int imax, int jmin, int jmax, int kmin, int kmax) {
  while (oi + *sa > imin && oi + *sa < imax) {
    if (fabs(abc) > 0.5 * fabs(X[oi + *sa] - X[oi])) *sa = *sa + isign1(*sa); else break;
  }
  while (oj + *sb > jmin && oj + *sb < jmax && oj - *sb > jmin && oj - *sb < jmax) {
    if (fabs(abc) > 0.5 * fabs(Y[oj + *sb] - Y[oj])) *sb = *sb + isign1(*sb); else break;
  }
  while (ok + *sc > kmin && ok + *sc < kmax && ok - *sc > kmin && ok - *sc < kmax) {
    if (fabs(abc) > 0.5 * fabs(Z[ok + *sc] - Z[ok])) *sc = *sc + isign1(*sc); else break;
  }
}

/* remapping.f90:92-394 compute_remapping_quantities (non-uniform branch, which
   trunk always takes: uniform_grid = .false. at :121-122). Returns the exterior
   flag (0 is the reference's fatal case). */
static int remap1(const ogrid* g, double u, double v, double w, double f_off, int st[7], double m[7]) {
  const double *X = g->ax[0], *Y = g->ax[1], *Z = g->ax[2];
  int imax = g->n[0] - 1, jmax = g->n[1] - 1, kmax = g->n[2] - 1;
  int imin = 0, jmin = 0, kmin = 0;
  int corner = 0;
  long long gi;
  int oi, oj, ok;
  gi = fnint(g->factor[0] * (u - g->vmin[0]));
  oi = bsearch_first_guess((int)(gi < -1 ? -1 : (gi > imax + 1 ? imax + 1 : gi)), u, X, imin, imax);
  if (oi >= imax) { oi = imax; corner++; } else if (oi <= imin) { oi = imin; corner++; }
  gi = fnint(g->factor[1] * (v - g->vmin[1]));
  oj = bsearch_first_guess((int)(gi < -1 ? -1 : (gi > jmax + 1 ? jmax + 1 : gi)), v, Y, jmin, jmax);
  if (oj >= jmax) { oj = jmax; corner++; } else if (oj <= jmin) { oj = jmin; corner++; }
  gi = fnint(g->factor[2] * (w - g->vmin[2]));
  ok = bsearch_first_guess((int)(gi < -1 ? -1 : (gi > kmax + 1 ? kmax + 1 : gi)), w, Z, kmin, kmax);
  if (ok >= kmax) { ok = kmax; corner++; } else if (ok <= kmin) { ok = kmin; corner++; }
  if (corner == 3) {
    int su = isign1(oi - (imin + 1)), sv = isign1(oj - (jmin + 1)), sw = isign1(ok - (kmin + 1));
    oi -= su; oj -= sv; ok -= sw;
  }
  double ox = X[oi], oy = Y[oj], oz = Z[ok];
  double a = u - ox, b = v - oy, c = w - oz;
  int sa = isgn(a), sb = isgn(b), sc = isgn(c);
  if (oi + sa > imax || oi + sa < imin) {
    sa = -sa;
    find_boundary_stencil(&sa, &sb, &sc, a, oi, oj, ok, X, Y, Z, imin, imax, jmin, jmax, kmin, kmax);
  }
  if (oj + sb > jmax || oj + sb < jmin) {
    sb = -sb;
    find_boundary_stencil(&sb, &sa, &sc, b, oj, oi, ok, Y, X, Z, jmin, jmax, imin, imax, kmin, kmax);
  }
  if (ok + sc > kmax || ok + sc < kmin) {
    sc = -sc;
    find_boundary_stencil(&sc, &sb, &sa, c, ok, oj, oi, Z, Y, X, kmin, kmax, jmin, jmax, imin, imax);
  }
  double ix = X[oi + sa] - ox, iy = Y[oj + sb] - oy, iz = Z[ok + sc] - oz;
  int fx = 0, fy = 0, fz = 0;
  double ex = 0.0, ey = 0.0, ez = 0.0;
  if (oi - sa <= imax && oi - sa >= imin) { fx = 1; ex = X[oi - sa] - ox; }
  if (oj - sb <= jmax && oj - sb >= jmin) { fy = 2; ey = Y[oj - sb] - oy; }
  if (ok - sc <= kmax && ok - sc >= kmin) { fz = 4; ez = Z[ok - sc] - oz; }
  int flag = fx + fy + fz;
  double num = f_off * (a * a + b * b + c * c - a * ix - b * iy - c * iz);
  double fe, fex = 0.0, fey = 0.0, fez = 0.0;
  switch (flag) {
    case 7: fe = num / (ex * (ex - ix) + ey * (ey - iy) + ez * (ez - iz)); fex = fe; fey = fe; fez = fe; break;
    case 1: fe = num / (ex * (ex - ix)); fex = fe; break;
    case 2: fe = num / (ey * (ey - iy)); fey = fe; break;
    case 4: fe = num / (ez * (ez - iz)); fez = fe; break;
    case 3: fe = num / (ex * (ex - ix) + ey * (ey - iy)); fex = fe; fey = fe; break;
    case 5: fe = num / (ex * (ex - ix) + ez * (ez - iz)); fex = fe; fez = fe; break;
    case 6: fe = num / (ey * (ey - iy) + ez * (ez - iz)); fey = fe; fez = fe; break;
    default:
      st[0] = oi; st[1] = oj; st[2] = ok; st[3] = sa; st[4] = sb; st[5] = sc; st[6] = 0;
      for (int q = 0; q < 7; ++q) m[q] = 0.0;
      return 0;
  }
  double fix = (f_off * a - fex * ex) / ix;
  double fiy = (f_off * b - fey * ey) / iy;
  double fiz = (f_off * c - fez * ez) / iz;
  double fo = f_off - fix - fiy - fiz - fex - fey - fez;
  st[0] = oi; st[1] = oj; st[2] = ok; st[3] = sa; st[4] = sb; st[5] = sc; st[6] = flag;
  m[0] = fo; m[1] = fix; m[2] = fiy; m[3] = fiz; m[4] = fex; m[5] = fey; m[6] = fez;
  return flag;
}

/* replenishing_collisions.f90:356-434 apply_phi_replenishment and
   boundary_conditions.f90:1852-1929 apply_bc_remapping: scale * mass(q) added on
   the stencil. `bc_form` selects the latter's product order (dens * repl). */
static void scatter_stencil(double* phi, const ogrid* g, const int st[7], const double m[7],
                            double s) {
  int nx = g->n[0], nxy = g->n[0] * g->n[1];
  int i = st[0], j = st[1], k = st[2], a = st[3], b = st[4], c = st[5], fl = st[6];
#define AT(ii, jj, kk) phi[(ii) + nx * (jj) + nxy * (kk)]
  AT(i, j, k) += s * m[0];
  AT(i + a, j, k) += s * m[1];
  AT(i, j + b, k) += s * m[2];
  AT(i, j, k + c) += s * m[3];
  if (fl & 1) AT(i - a, j, k) += s * m[4];
  if (fl & 2) AT(i, j - b, k) += s * m[5];
  if (fl & 4) AT(i, j, k - c) += s * m[6];
#undef AT
}

int orc_remap(orc_ctx* ctx, int64_t n, const double* uvw, double f_off, int32_t* stencil,
              double* mass) {
  for (int64_t q = 0; q < n; ++q) {
    int st[7];
    double m[7];
    remap1(&ctx->g, uvw[3 * q], uvw[3 * q + 1], uvw[3 * q + 2], f_off, st, m);
    for (int t = 0; t < 7; ++t) { stencil[7 * q + t] = st[t]; mass[7 * q + t] = m[t]; }
  }
  return DVM_OK;
}

/* --------------------------------------------------------------- context */
static double gamma_lanczos(double x);

int orc_ctx_create(const dvm_config* cfg, orc_ctx** out) {
  *out = NULL;
  orc_ctx* c = (orc_ctx*)calloc(1, sizeof(orc_ctx));
  if (!c) return DVM_E_NOMEM;
  c->cfg = *cfg;
  if (cfg->nx < 1 || cfg->ny < 1 || cfg->num_repl < 1) { free(c); return DVM_E_ARG; }
  if (grid_setup(&c->g, cfg) != 0) { free(c); return DVM_E_ARG; }
  c->cfg.vx = c->g.ax[0]; c->cfg.vy = c->g.ax[1]; c->cfg.vz = c->g.ax[2];
  c->ncell = cfg->nx * cfg->ny;
  size_t N = (size_t)c->g.num_points;
  c->f = (double*)calloc((size_t)c->ncell * N, sizeof(double));
  c->fd = (double*)calloc((size_t)c->ncell * N, sizeof(double));
  c->props = (double*)calloc((size_t)c->ncell * DVM_NPROP, sizeof(double));
  c->eq_df = (double*)calloc((size_t)c->ncell * (N + 1), sizeof(double));
  c->eq_par = (double*)calloc((size_t)c->ncell * 5, sizeof(double));
  c->eqc = (double*)calloc(N + 1, sizeof(double));
  c->neqc = (double*)calloc(N + 1, sizeof(double));
  if (!c->f || !c->fd || !c->props || !c->eq_df || !c->eq_par || !c->eqc || !c->neqc) {
    orc_ctx_destroy(c);
    return DVM_E_NOMEM;
  }
  /* calculate_cross_section (scaling.f90:163-201), single species: d1 = d2,
     m_red = m*m/(m+m), omega_AB = (w+w)/2 */
  double d = cfg->diam, m = cfg->mass;
  double m_red = m * m / (m + m);
  double omega_ab = 0.5 * (cfg->omega + cfg->omega);
  double sigma_vhs = 0.25 * (d + d) * (d + d);
  if (cfg->csection_coeff == DVM_CS_BIRD) {
    c->sigma = sigma_vhs * pow(m_red, 0.5 - omega_ab) / gamma_lanczos(2.5 - omega_ab);
  } else if (cfg->csection_coeff == DVM_CS_VKI) {
    c->sigma = (3.32335097045 * pow(m_red, 1.0 - omega_ab) * pow(cfg->t_viscosity, omega_ab)) /
               (gamma_lanczos(4.5 - omega_ab) * cfg->viscosity);
  } else {
    c->sigma = sigma_vhs;
  }
  c->vhs_exponent = 2.0 - 2.0 * omega_ab;
  *out = c;
  return DVM_OK;
}

void orc_ctx_destroy(orc_ctx* c) {
  if (!c) return;
  for (int d = 0; d < 3; ++d) { free(c->g.ax[d]); free(c->g.beta[d]); }
  free(c->f); free(c->fd); free(c->props); free(c->eq_df); free(c->eq_par);
  free(c->eqc); free(c->neqc); free(c->heat); free(c->su);
  free(c);
}

/* math_utils.f90:83-135 gamma_calculator (Lanczos, coefficients as listed there) */
static double gamma_lanczos(double x) {
  static const double a[9] = {0.99999999999980993, 676.5203681218851, -1259.1392167224028,
                              771.32342877765313, -176.61502916214059, 12.507343278686905,
                              -0.13857109526572012, 9.9843695780195716e-6, 1.5056327351493116e-7};
  if (x < 0.5) return PI_REF / (sin(PI_REF * x) * gamma_lanczos(1.0 - x));
  x = x - 1.0;
  double y = a[0];
  for (int i = 1; i <= 8; ++i) y = y + a[i] / (x + (double)i);
  double z = x + 7.0 + 0.5;
  return sqrt(2.0 * PI_REF) * pow(z, x + 0.5) * exp(-z) * y;
}

int orc_upload_f(orc_ctx* c, int32_t species, const double* f) {
  if (species != 0) return set_err(c, DVM_E_ARG, "only species 0 is supported");
  size_t n = (size_t)c->ncell * c->g.num_points;
  memcpy(c->f, f, n * sizeof(double));
  memcpy(c->fd, f, n * sizeof(double));
  memset(c->eq_par, 0, sizeof(double) * 5 * (size_t)c->ncell);
  memset(c->eq_df, 0, sizeof(double) * (size_t)c->ncell * (c->g.num_points + 1));
  c->f_valid = 1;
  c->props_valid = 0;
  return DVM_OK;
}

int orc_set_f_keep_cache(orc_ctx* c, const double* f) {
  size_t n = (size_t)c->ncell * c->g.num_points;
  memcpy(c->f, f, n * sizeof(double));
  memcpy(c->fd, f, n * sizeof(double));
  c->f_valid = 1;
  c->props_valid = 0;
  return DVM_OK;
}

int orc_grid_info(orc_ctx* c, int32_t* ints, double* beta3_min) {
  const ogrid* g = &c->g;
  ints[0] = g->num_points;
  for (int d = 0; d < 3; ++d) {
    ints[1 + 3 * d] = g->zero[d];
    ints[2 + 3 * d] = g->neg[d];
    ints[3 + 3 * d] = g->pos[d];
  }
  ints[10] = g->n[0];
  *beta3_min = g->beta3_min;
  return DVM_OK;
}

int orc_download_f(orc_ctx* c, int32_t species, double* f) {
  if (species != 0) return set_err(c, DVM_E_ARG, "only species 0 is supported");
  memcpy(f, c->f, (size_t)c->ncell * c->g.num_points * sizeof(double));
  return DVM_OK;
}

/* --------------------------------------------------------------- moments */
/* physical_props.f90:285-485 for one monatomic species. */
static void moments1(const ogrid* g, const double* f, double mass, double* p) {
  int nx = g->n[0], ny = g->n[1], nz = g->n[2];
  const double *X = g->ax[0], *Y = g->ax[1], *Z = g->ax[2];
  double dens = 0.0;
  for (int l = 0; l < g->num_points; ++l) dens = dens + f[l];
  for (int q = 0; q < DVM_NPROP; ++q) p[q] = 0.0;
  p[DVM_P_DENS] = dens;
  if (dens == 0.0) return;
  double u = 0.0, v = 0.0, w = 0.0, sp = 0.0;
  int l = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i, ++l) {
        u = u + f[l] * X[i];
        v = v + f[l] * Y[j];
        w = w + f[l] * Z[k];
        sp = sp + f[l] * sqrt(X[i] * X[i] + Y[j] * Y[j] + Z[k] * Z[k]);
      }
  u = u / dens; v = v / dens; w = w / dens; sp = sp / dens;
  double t = 0.0;
  l = 0;
  for (int k = 0; k < nz; ++k) {
    double czsq = (Z[k] - w); czsq = czsq * czsq;
    for (int j = 0; j < ny; ++j) {
      double cysq = (Y[j] - v); cysq = cysq * cysq;
      for (int i = 0; i < nx; ++i, ++l) {
        double cxsq = (X[i] - u); cxsq = cxsq * cxsq;
        t = t + f[l] * (cxsq + cysq + czsq);
      }
    }
  }
  t = 2.0 * mass * t / (3.0 * dens);
  p[DVM_P_U] = u; p[DVM_P_V] = v; p[DVM_P_W] = w;
  p[DVM_P_TRTEMP] = t;
  /* compute_overall_temperature (physical_props.f90:1428-1457) with zero rot/vib dof */
  p[DVM_P_TEMP] = (3.0 * t + 0.0 * 0.0 + 0.0 * 0.0) / (3.0 + 0.0 + 0.0);
  p[DVM_P_PRESSURE] = 0.5 * dens * t;
  p[DVM_P_AVGSPEED] = sp;
}

/* Output moments per cell in the reference's loop order: compute_heat_flux
   (physical_props.f90:1914-1986), compute_viscous_stress (:1988-2065),
   compute_entropy (:1871-1912, C = 1). ext = {qx qy qz txx tyy tzz txy txz
   tyz entropy}; uses the moments of the current f. */
int orc_moments_ext(orc_ctx* c, double* ext) {
  int rc = orc_moments(c, NULL);
  if (rc) return rc;
  const ogrid* g = &c->g;
  const int N = g->num_points;
  const double mass = c->cfg.mass;
  for (int cell = 0; cell < c->ncell; ++cell) {
    const double* f = c->f + (size_t)cell * N;
    const double* p = c->props + (size_t)cell * DVM_NPROP;
    const double dens = p[DVM_P_DENS], u = p[DVM_P_U], v = p[DVM_P_V], w = p[DVM_P_W];
    double q[3] = {0, 0, 0}, t[6] = {0, 0, 0, 0, 0, 0}, ent = 0.0;
    int l = 0;
    for (int k = 0; k < g->n[2]; ++k) {
      const double z = g->ax[2][k];
      const double czsq = (z - w) * (z - w);
      for (int j = 0; j < g->n[1]; ++j) {
        const double y = g->ax[1][j];
        const double cysq = (y - v) * (y - v);
        for (int i = 0; i < g->n[0]; ++i, ++l) {
          const double x = g->ax[0][i];
          const double cxsq = (x - u) * (x - u);
          const double csq = cxsq + cysq + czsq;
          const double fl = f[l];
          q[0] = q[0] + (x - u) * csq * fl;
          q[1] = q[1] + (y - v) * csq * fl;
          q[2] = q[2] + (z - w) * csq * fl;
          t[0] = t[0] - (x - u) * (x - u) * fl;
          t[1] = t[1] - (y - v) * (y - v) * fl;
          t[2] = t[2] - (z - w) * (z - w) * fl;
          t[3] = t[3] - (x - u) * (y - v) * fl;
          t[4] = t[4] - (x - u) * (z - w) * fl;
          t[5] = t[5] - (y - v) * (z - w) * fl;
          if (fl != 0.0) ent = ent - fabs(fl) * log(1.0 * fabs(fl));
        }
      }
    }
    double* o = ext + (size_t)cell * DVM_NEXT;
    for (int a = 0; a < 3; ++a) o[a] = (dens == 0.0) ? 0.0 : q[a] * (0.5 * mass);
    for (int a = 0; a < 3; ++a) o[3 + a] = p[DVM_P_PRESSURE] + mass * t[a];
    for (int a = 3; a < 6; ++a) o[3 + a] = mass * t[a];
    o[9] = ent;
  }
  return DVM_OK;
}

int orc_moments(orc_ctx* c, double* props_out) {
  if (!c->f_valid) return set_err(c, DVM_E_STATE, "moments before upload");
  size_t N = (size_t)c->g.num_points;
  for (int cell = 0; cell < c->ncell; ++cell)
    moments1(&c->g, c->f + cell * N, c->cfg.mass, c->props + (size_t)cell * DVM_NPROP);
  c->props_valid = 1;
  if (props_out) memcpy(props_out, c->props, sizeof(double) * DVM_NPROP * (size_t)c->ncell);
  return DVM_OK;
}

/* ------------------------------------------------------------------- CDF */
typedef struct {
  double dens, neq_dens, eq_dens, temp;
  int num[3];
  double factor[3];
  int fmc; /* full Monte Carlo: integral-3 samples from the |f| CDF, signs dsgn(f) */
} cell_coll;

/* variance_reduction.f90:234-496 compute_cdf_vr for the like-species pair
   (same = AA). eq_df of the cell is updated when the cache misses (commit=1). */
static void cdf_vr(orc_ctx* c, int cell, int commit, double* eq_df_out) {
  const ogrid* g = &c->g;
  int N = g->num_points;
  const double* p = c->props + (size_t)cell * DVM_NPROP;
  const double* f = c->f + (size_t)cell * N;
  double* eq_df = c->eq_df + (size_t)cell * (N + 1);
  double* par = c->eq_par + (size_t)cell * 5;
  double mass = c->cfg.mass;
  double dens = p[DVM_P_DENS], u = p[DVM_P_U], v = p[DVM_P_V], w = p[DVM_P_W];
  double temp = p[DVM_P_TEMP];
  if (temp < DOUBLE_TOL) temp = p[DVM_P_TEMP];
  double coeff = 0.0;
  if (dens > 0.0)
    coeff = dens * sqrt(mass * mass * mass / (PI_REF * PI_REF * PI_REF * temp * temp * temp));
  double tol = c->cfg.cdf_tolerance;
  int new_eq = (fabs(dens - par[0]) > tol * dens || fabs(u - par[1]) > tol * u ||
                fabs(v - par[2]) > tol * v || fabs(temp - par[4]) > tol * temp);
  double* e = eq_df;
  if (!commit && eq_df_out) {
    memcpy(eq_df_out, eq_df, sizeof(double) * (N + 1));
    e = eq_df_out;
  }
  e[0] = 0.0;
  if (new_eq) {
    int l = 1;
    for (int k = 0; k < g->n[2]; ++k)
      for (int j = 0; j < g->n[1]; ++j)
        for (int i = 0; i < g->n[0]; ++i, ++l) {
          double x = g->ax[0][i], y = g->ax[1][j], z = g->ax[2][k];
          double beta3 = g->beta[0][i] * g->beta[1][j] * g->beta[2][k];
          double phi_eq;
          if (temp > 0.0) {
            /* dist_func.f90:317-344 compute_maxwellian */
            double cx = (x - u); cx = cx * cx;
            double cy = (y - v); cy = cy * cy;
            double cz = (z - w); cz = cz * cz;
            double csq = cx + cy + cz;
            phi_eq = coeff * exp(-csq * mass / temp);
          } else {
            phi_eq = 0.0;
          }
          phi_eq = phi_eq * beta3;
          e[l] = e[l - 1] + phi_eq;
        }
    if (commit) { par[0] = dens; par[1] = u; par[2] = v; par[4] = temp; }
  }
  double* eqc = c->eqc;
  double* neqc = c->neqc;
  eqc[0] = 0.0;
  neqc[0] = 0.0;
  for (int l = 1; l <= N; ++l) {
    double phi_eq = e[l] - e[l - 1];
    double phi_neq = f[l - 1] - phi_eq;
    neqc[l] = neqc[l - 1] + fabs(phi_neq);
    eqc[l] = eqc[l - 1] + phi_eq;
  }
  orc_build_buckets(eqc, N, c->eql, c->eqr);
  orc_build_buckets(neqc, N, c->neql, c->neqr);
}

/* collision_utilities.f90:106-140 compute_num_coll_pairs and the explicit-count
   rule; variance_reduction.f90:598-604,679,831,976 factors. */
static void pair_counts(orc_ctx* c, int cell, cell_coll* cc) {
  const dvm_config* cfg = &c->cfg;
  const double* p = c->props + (size_t)cell * DVM_NPROP;
  int N = c->g.num_points;
  cc->dens = p[DVM_P_DENS];
  cc->temp = p[DVM_P_TEMP];
  cc->neq_dens = c->neqc[N];
  cc->eq_dens = c->eqc[N];
  double factor_coeff = cfg->dt * 0.5 / cfg->kn;
  double dA[3], dB[3], wA[3], wB[3];
  dA[0] = 1.0;                                  dB[0] = 0.5 * cc->neq_dens / cc->dens;
  dA[1] = 0.5 * cc->neq_dens / cc->dens;        dB[1] = 1.0;
  dA[2] = 0.5 * cc->neq_dens / cc->dens;        dB[2] = 0.5 * cc->neq_dens / cc->dens;
  wA[0] = cc->eq_dens;  wB[0] = cc->neq_dens;
  wA[1] = cc->neq_dens; wB[1] = cc->eq_dens;
  wA[2] = cc->neq_dens; wB[2] = cc->neq_dens;
  double tot = 0.0;
  if (cfg->sample_mode == DVM_SAMPLES_EXPLICIT)
    tot = dA[0] * dB[0] + dA[1] * dB[1] + dA[2] * dB[2];
  for (int t = 0; t < 3; ++t) {
    long long num;
    if (cfg->sample_mode == DVM_SAMPLES_EXPLICIT) {
      num = (tot > 0.0) ? fnint((double)cfg->samples_per_cell * (dA[t] * dB[t]) / tot) : 0;
    } else {
      double rms_sq = cfg->coln_rms * cfg->coln_rms;
      double b3a = c->g.beta3_min * pow(cfg->mass, 3.0 / 2.0);
      double b3b = c->g.beta3_min * pow(cfg->mass, 3.0 / 2.0);
      double beta3 = 0.5 * (b3a + b3b);
      double ct = 0.5 * (cc->temp + cc->temp);
      ct = pow(ct, 2.0 / 3.0);
      num = fnint((0.5 * cfg->dt * ct * dA[t] * dB[t]) / (cfg->kn * rms_sq * beta3));
    }
    if (num < 0) num = 0;
    /* the library's bound on one integral's samples (a third of 2^30) */
    if (num > (1LL << 30) / 3) num = (1LL << 30) / 3;
    cc->num[t] = (int)num;
    cc->factor[t] = (num > 0) ? factor_coeff * c->sigma * wA[t] * wB[t] / (double)num : 0.0;
  }
}

/* full_mc.f90:110-180 compute_cdf_fmc: sequential prefix of |f| in node order
   and its 1000-bucket table (the same table rule as compute_cdf_vr). Held in
   the deviation slots (neqc, neql, neqr); the equilibrium slots are zeroed. */
static void cdf_fmc(orc_ctx* c, int cell) {
  int N = c->g.num_points;
  const double* f = c->f + (size_t)cell * N;
  double* eqc = c->eqc;
  double* neqc = c->neqc;
  eqc[0] = 0.0;
  neqc[0] = 0.0;
  for (int l = 1; l <= N; ++l) {
    neqc[l] = neqc[l - 1] + fabs(f[l - 1]);
    eqc[l] = 0.0;
  }
  orc_build_buckets(eqc, N, c->eql, c->eqr);
  orc_build_buckets(neqc, N, c->neql, c->neqr);
}

/* full_mc.f90:240-317: abs_dens = dens - 2 neg_dens, neg_dens = (dens - sum|f|)/2;
   compute_num_coll_pairs (collision_utilities.f90:106-140) with both density
   ratios abs_dens/dens (or the explicit count); factor
   (dt/Kn) sigma abs_dens^2 / num_colls, halved for like species (:319-321).
   Every sample is an integral-3 sample. */
static void fmc_counts(orc_ctx* c, int cell, cell_coll* cc) {
  const dvm_config* cfg = &c->cfg;
  const double* p = c->props + (size_t)cell * DVM_NPROP;
  int N = c->g.num_points;
  memset(cc, 0, sizeof *cc);
  cc->fmc = 1;
  cc->dens = p[DVM_P_DENS];
  cc->temp = p[DVM_P_TEMP];
  cc->neq_dens = c->neqc[N];
  double neg_dens = 0.5 * (cc->dens - c->neqc[N]);
  double abs_dens = cc->dens - 2.0 * neg_dens;
  long long num;
  if (cfg->sample_mode == DVM_SAMPLES_EXPLICIT) {
    num = cfg->samples_per_cell;
  } else {
    double ratio = abs_dens / cc->dens;
    double rms_sq = cfg->coln_rms * cfg->coln_rms;
    double b3a = c->g.beta3_min * pow(cfg->mass, 3.0 / 2.0);
    double b3b = c->g.beta3_min * pow(cfg->mass, 3.0 / 2.0);
    double beta3 = 0.5 * (b3a + b3b);
    double ct = 0.5 * (cc->temp + cc->temp);
    ct = pow(ct, 2.0 / 3.0);
    double x = (0.5 * cfg->dt * ct * ratio * ratio) / (cfg->kn * rms_sq * beta3);
    num = (x > 0.0) ? fnint(x) : 0; /* the reference's nint of NaN (empty cell) is left undefined */
  }
  if (!(abs_dens > 0.0) || num < 0) num = 0;
  if (num > (1LL << 30)) num = 1LL << 30; /* the library's bound on a cell's samples */
  cc->num[2] = (int)num;
  if (num > 0) {
    double factor = (cfg->dt / cfg->kn) * c->sigma * abs_dens * abs_dens / (double)num;
    cc->factor[2] = 0.5 * factor;
  }
}

/* One VR sample (variance_reduction.f90:686-816 / 838-961 / 983-1116 with
   vr_depletion_routine :1127-1237, replenish_collision
   replenishing_collisions.f90:134-354). Scatters into fd; when rec is
   non-NULL records the sample for the parity trace. Returns 1 if skipped. */
static int vr_sample(orc_ctx* c, int cell, int ntime, int64_t q, int t, int s, const cell_coll* cc,
                     double* fd, int32_t* irec, double* drec) {
  const dvm_config* cfg = &c->cfg;
  const ogrid* g = &c->g;
  int N = g->num_points;
  int nx = g->n[0], nxy = g->n[0] * g->n[1];
  const double* f = c->f + (size_t)cell * N;
  const double* eq_df = c->eq_df + (size_t)cell * (N + 1);
  /* the sample's uniforms: uA, uB, then (phi, theta) per replenish location,
     from one Philox block per location or from the explicit list (sample q of
     the cell at su_base + q * (2 + 2 NRepl)) */
  const int per = 2 + 2 * cfg->num_repl;
  const double* us = c->su ? c->su + c->su_base + q * per : NULL;
  double u0[4];
  if (us) { u0[0] = us[0]; u0[1] = us[1]; u0[2] = us[2]; u0[3] = us[3]; }
  else uniforms4(cfg->seed, (uint32_t)s, (uint32_t)(t + 1), (uint32_t)cell, (uint32_t)ntime, u0);
  const double uA = u0[0], uB = u0[1];
  /* pick_collision_partner (collision_utilities.f90:142-186) */
  int lA, lB;
  if (t == 0) lA = orc_referenced_search(c->eqc, N, c->eql, c->eqr, uA);
  else        lA = orc_referenced_search(c->neqc, N, c->neql, c->neqr, uA);
  if (t == 1) lB = orc_referenced_search(c->eqc, N, c->eql, c->eqr, uB);
  else        lB = orc_referenced_search(c->neqc, N, c->neql, c->neqr, uB);
  /* global2local_map (velocity_grid.f90:691-712) */
  int kA = lA / nxy, jA = lA / nx - kA * g->n[1], iA = lA - jA * nx - kA * nxy;
  int kB = lB / nxy, jB = lB / nx - kB * g->n[1], iB = lB - jB * nx - kB * nxy;
  double x1 = g->ax[0][iA], y1 = g->ax[1][jA], z1 = g->ax[2][kA];
  double x2 = g->ax[0][iB], y2 = g->ax[1][jB], z2 = g->ax[2][kB];
  if (irec) {
    memset(irec, 0, sizeof(int32_t) * 32);
    memset(drec, 0, sizeof(double) * 16);
    irec[0] = t + 1; irec[1] = s; irec[2] = lA; irec[3] = lB;
  }
  /* compute_relative_velocity (collision_utilities.f90:188-218) */
  double gx = x2 - x1, gy = y2 - y1, gz = z2 - z1;
  double gg = sqrt(gx * gx + gy * gy + gz * gz);
  if (gg < DOUBLE_TOL) {
    if (irec) irec[4] = 1;
    return 1;
  }
  double g_sigma = pow(gg, c->vhs_exponent);
  /* signs of deviation partners: cdf_neq%sign(gl) = dsgn(f - phi_eq) at that node */
  /* full Monte Carlo (full_mc.f90:337-338, 434-437): dsgn of phi_old at the picks */
  double signA = cc->fmc ? dsgn(f[lA]) : dsgn(f[lA] - (eq_df[lA + 1] - eq_df[lA]));
  double signB = cc->fmc ? dsgn(f[lB]) : dsgn(f[lB] - (eq_df[lB + 1] - eq_df[lB]));
  double sumA = (t == 0) ? 1.0 : signA;
  double sumB = (t == 1) ? 1.0 : signB;
  double factor = cc->factor[t];
  double dep1 = sumB * 1.0 * factor * g_sigma;
  double dep2 = sumA * 1.0 * factor * g_sigma;
  double kd1 = sumA * dep1, kd2 = sumB * dep2;
  if (fd) {
    fd[lA] = fd[lA] - kd1;
    fd[lB] = fd[lB] - kd2;
  }
  double kr1 = kd1 / (double)cfg->num_repl, kr2 = kd2 / (double)cfg->num_repl;
  /* compute_center_of_mass_velocity (collision_utilities.f90:220-238) */
  double m1 = cfg->mass, m2 = cfg->mass;
  double mf1 = m1 / (m1 + m2), mf2 = m2 / (m1 + m2);
  double vc0 = mf1 * x1 + mf2 * x2, vc1 = mf1 * y1 + mf2 * y2, vc2 = mf1 * z1 + mf2 * z2;
  double alpha_ab = 0.5 * (cfg->alpha + cfg->alpha);
  if (irec) { drec[0] = kd1; drec[1] = kd2; }
  for (int r = 1; r <= cfg->num_repl; ++r) {
    double uphi = u0[2], uth = u0[3];
    if (r > 1 && us) {
      uphi = us[2 * r];
      uth = us[2 * r + 1];
    } else if (r > 1) {
      double ur[4];
      uniforms4(cfg->seed, (uint32_t)s, (uint32_t)((t + 1) | (r << 2)), (uint32_t)cell, (uint32_t)ntime, ur);
      uphi = ur[0];
      uth = ur[1];
    }
    /* find_post_collision_velocities (collision_utilities.f90:240-282), g_ratio = 1 */
    double gx1 = fabs(x1 - vc0), gy1 = fabs(y1 - vc1), gz1 = fabs(z1 - vc2);
    double gx2 = fabs(x2 - vc0), gy2 = fabs(y2 - vc1), gz2 = fabs(z2 - vc2);
    gx1 = gx1 * 1.0; gy1 = gy1 * 1.0; gz1 = gz1 * 1.0;
    gx2 = gx2 * 1.0; gy2 = gy2 * 1.0; gz2 = gz2 * 1.0;
    /* pick_random_velocity_on_sphere (collision_utilities.f90:284-352) */
    double ur = gx1 + gx2, vr = gy1 + gy2, wr = gz1 + gz2;
    double gp1 = sqrt(gx1 * gx1 + gy1 * gy1 + gz1 * gz1);
    double gp2 = sqrt(gx2 * gx2 + gy2 * gy2 + gz2 * gz2);
    double ph = 2.0 * PI_REF * uphi;
    double cph = cos(ph), sph = sin(ph);
    double ralpha = 1.0 / alpha_ab;
    double rx, ry, rz;
    if (fabs(ralpha - 1.0) > DOUBLE_TOL) {
      double ct = 2.0 * pow(uth, ralpha) - 1.0;
      double stt = sqrt(1.0 - ct * ct);
      double den = sqrt(vr * vr + wr * wr);
      if (den > DOUBLE_TOL * gg) {
        rx = ct * ur + stt * sph * den;
        ry = ct * vr + stt * (gg * wr * cph - ur * vr * sph) / den;
        rz = ct * wr - stt * (gg * vr * cph + ur * wr * sph) / den;
      } else {
        rx = ct * ur;
        ry = stt * cph * ur;
        rz = stt * sph * ur;
      }
    } else {
      double ct = 2.0 * uth - 1.0;
      double stt = sqrt(1.0 - ct * ct);
      rx = gg * stt * cph;
      ry = gg * stt * sph;
      rz = gg * ct;
    }
    double gr1 = gp1 / gg, gr2 = gp2 / gg;
    double xi0 = vc0 + gr1 * rx, xi1 = vc1 + gr1 * ry, xi2 = vc2 + gr1 * rz;
    double et0 = vc0 + -(gr2 * rx), et1 = vc1 + -(gr2 * ry), et2 = vc2 + -(gr2 * rz);
    int st1[7], st2[7];
    double ms1[7], ms2[7];
    int fl1 = remap1(g, xi0, xi1, xi2, 1.0, st1, ms1);
    int fl2 = remap1(g, et0, et1, et2, 1.0, st2, ms2);
    if (irec && r == 1) {
      for (int q = 0; q < 7; ++q) {
        irec[5 + q] = st1[q]; irec[12 + q] = st2[q];
        drec[2 + q] = ms1[q]; drec[9 + q] = ms2[q];
      }
    }
    if (fl1 == 0 || fl2 == 0) return -1;
    if (fd) {
      scatter_stencil(fd, g, st1, ms1, kr1);
      scatter_stencil(fd, g, st2, ms2, kr2);
    }
  }
  (void)nx;
  return 0;
}

/* The cell's CDFs and sample counts for the configured solver (the
   equilibrium cache is committed when commit = 1). */
static void cell_cdf_counts(orc_ctx* c, int cell, int commit, double* eq_df_out, cell_coll* cc) {
  if (c->cfg.coll_solver == DVM_COLL_FMC) {
    cdf_fmc(c, cell);
    fmc_counts(c, cell, cc);
  } else {
    cdf_vr(c, cell, commit, eq_df_out);
    pair_counts(c, cell, cc);
    cc->fmc = 0;
  }
}

static int collide_cell(orc_ctx* c, int cell, int ntime, int64_t* nsamp) {
  int N = c->g.num_points;
  cell_coll cc;
  cell_cdf_counts(c, cell, 1, NULL, &cc);
  if (c->su) {
    /* the cell's samples read the list from the cursor on */
    const int64_t need = ((int64_t)cc.num[0] + cc.num[1] + cc.num[2]) * (2 + 2 * c->cfg.num_repl);
    if (need > c->su_n - c->su_cursor)
      return set_err(c, DVM_E_ARG, "uniform stream: cell %d reads %lld values, %lld are left", cell,
                     (long long)need, (long long)(c->su_n - c->su_cursor));
    c->su_base = c->su_cursor;
    c->su_cursor += need;
  }
  double* fd = c->fd + (size_t)cell * N;
  memcpy(fd, c->f + (size_t)cell * N, sizeof(double) * N);
  int64_t tot = 0, q = 0;
  for (int t = 0; t < 3; ++t) {
    tot += cc.num[t];
    for (int s = 0; s < cc.num[t]; ++s, ++q) {
      if (vr_sample(c, cell, ntime, q, t, s, &cc, fd, NULL, NULL) < 0)
        return set_err(c, DVM_E_REMAP, "invalid remap exterior flag (cell %d)", cell);
    }
  }
  /* perform_dist_func_update: phi := delta_phi */
  memcpy(c->f + (size_t)cell * N, fd, sizeof(double) * N);
  if (nsamp) *nsamp = tot;
  return DVM_OK;
}

int orc_collide_cells(orc_ctx* c, int32_t ntime, int32_t c0, int32_t c1) {
  if (!c->props_valid) return set_err(c, DVM_E_STATE, "collide before moments");
  for (int cell = c0; cell < c1; ++cell) {
    int rc = collide_cell(c, cell, ntime, NULL);
    if (rc) return rc;
  }
  return DVM_OK;
}

int orc_collide(orc_ctx* c, int32_t ntime, int64_t* samples_out) {
  if (!c->props_valid) return set_err(c, DVM_E_STATE, "collide before moments");
  if (c->cfg.coll_solver != DVM_COLL_VR && c->cfg.coll_solver != DVM_COLL_FMC) return DVM_OK;
  for (int cell = 0; cell < c->ncell; ++cell) {
    int rc = collide_cell(c, cell, ntime, samples_out ? samples_out + cell : NULL);
    if (rc) return rc;
  }
  return DVM_OK;
}

int orc_collide_trace(orc_ctx* c, int32_t ntime, int32_t cell, int64_t max_records,
                      int32_t* irec, double* drec, int64_t* n_out) {
  if (!c->props_valid) return set_err(c, DVM_E_STATE, "trace before moments");
  if (cell < 0 || cell >= c->ncell) return set_err(c, DVM_E_ARG, "cell out of range");
  int N = c->g.num_points;
  /* compute with a scratch copy of the equilibrium cache */
  double* save = (double*)malloc(sizeof(double) * (N + 1));
  double par[5];
  memcpy(save, c->eq_df + (size_t)cell * (N + 1), sizeof(double) * (N + 1));
  memcpy(par, c->eq_par + (size_t)cell * 5, sizeof par);
  cell_coll cc;
  cell_cdf_counts(c, cell, 1, NULL, &cc);
  int64_t q = 0;
  if (c->su) {
    /* the traced cell reads the list from the cursor, which does not move */
    const int64_t need = ((int64_t)cc.num[0] + cc.num[1] + cc.num[2]) * (2 + 2 * c->cfg.num_repl);
    if (need > c->su_n - c->su_cursor) {
      memcpy(c->eq_df + (size_t)cell * (N + 1), save, sizeof(double) * (N + 1));
      memcpy(c->eq_par + (size_t)cell * 5, par, sizeof par);
      free(save);
      return set_err(c, DVM_E_ARG, "uniform stream: the trace reads %lld values, %lld are left", (long long)need,
                     (long long)(c->su_n - c->su_cursor));
    }
    c->su_base = c->su_cursor;
  }
  for (int t = 0; t < 3; ++t)
    for (int s = 0; s < cc.num[t]; ++s, ++q)
      if (q < max_records) vr_sample(c, cell, ntime, q, t, s, &cc, NULL, irec + 32 * q, drec + 16 * q);
  memcpy(c->eq_df + (size_t)cell * (N + 1), save, sizeof(double) * (N + 1));
  memcpy(c->eq_par + (size_t)cell * 5, par, sizeof par);
  free(save);
  *n_out = q;
  return DVM_OK;
}

/* Distance from equilibrium of each cell (the measure of dvm_l1; no
   counterpart in the reference): M_l = exp(-m |c_l - u|^2 / T) beta3_l
   scaled so that sum M = n, then sum |f - M| / n, in node order. */
int orc_l1(orc_ctx* c, double* l1) {
  if (!c->f_valid) return set_err(c, DVM_E_STATE, "L1 before upload");
  int rc = orc_moments(c, NULL);
  if (rc) return rc;
  const ogrid* g = &c->g;
  int N = g->num_points;
  for (int cell = 0; cell < c->ncell; ++cell) {
    const double* p = c->props + (size_t)cell * DVM_NPROP;
    const double* f = c->f + (size_t)cell * N;
    double n = p[DVM_P_DENS], u = p[DVM_P_U], v = p[DVM_P_V], w = p[DVM_P_W], T = p[DVM_P_TRTEMP];
    if (!(n > 0.0) || !(T > 0.0)) { l1[cell] = -1.0; continue; }
    double a = c->cfg.mass / T, msum = 0.0, e = 0.0;
    for (int pass = 0; pass < 2; ++pass) {
      double scale = (pass == 1) ? n / msum : 0.0;
      int l = 0;
      for (int k = 0; k < g->n[2]; ++k)
        for (int j = 0; j < g->n[1]; ++j)
          for (int i = 0; i < g->n[0]; ++i, ++l) {
            double dx = g->ax[0][i] - u, dy = g->ax[1][j] - v, dz = g->ax[2][k] - w;
            double m = exp(-(dx * dx + dy * dy + dz * dz) * a) * g->beta[0][i] * g->beta[1][j] * g->beta[2][k];
            if (pass == 0) msum += m;
            else e += fabs(f[l] - scale * m);
          }
    }
    l1[cell] = e / n;
  }
  return DVM_OK;
}

int orc_set_uniform_stream(orc_ctx* c, const double* u, int64_t n) {
  if (n < 0 || (n > 0 && !u)) return set_err(c, DVM_E_ARG, "uniform stream: bad length or null pointer");
  for (int64_t q = 0; q < n; ++q)
    if (!(u[q] >= 0.0 && u[q] < 1.0))
      return set_err(c, DVM_E_ARG, "uniform stream: u[%lld] = %g is not in [0, 1)", (long long)q, u[q]);
  free(c->su);
  c->su = NULL;
  c->su_n = c->su_cursor = c->su_base = 0;
  if (n == 0) return DVM_OK;
  c->su = (double*)malloc(sizeof(double) * (size_t)n);
  if (!c->su) return set_err(c, DVM_E_NOMEM, "uniform stream: out of memory");
  memcpy(c->su, u, sizeof(double) * (size_t)n);
  c->su_n = n;
  return DVM_OK;
}

int orc_uniform_stream_position(orc_ctx* c, int64_t* consumed) {
  *consumed = c->su_cursor;
  return DVM_OK;
}

int orc_cell_cdf(orc_ctx* c, int32_t cell, double* eq_cumul, double* neq_cumul, int32_t* eq_left,
                 int32_t* eq_right, int32_t* neq_left, int32_t* neq_right, int32_t* counts) {
  if (!c->props_valid) return set_err(c, DVM_E_STATE, "cdf before moments");
  int N = c->g.num_points;
  double* tmp = (double*)malloc(sizeof(double) * (N + 1));
  cell_coll cc;
  cell_cdf_counts(c, cell, 0, tmp, &cc);
  free(tmp);
  memcpy(eq_cumul, c->eqc, sizeof(double) * (N + 1));
  memcpy(neq_cumul, c->neqc, sizeof(double) * (N + 1));
  memcpy(eq_left, c->eql, sizeof c->eql);
  memcpy(eq_right, c->eqr, sizeof c->eqr);
  memcpy(neq_left, c->neql, sizeof c->neql);
  memcpy(neq_right, c->neqr, sizeof c->neqr);
  for (int t = 0; t < 3; ++t) counts[t] = cc.num[t];
  return DVM_OK;
}

/* ------------------------------------------------------------ convection */
/* dist_func.f90:317-344 with the wall coefficient sqrt(m^3/(pi^3 T^3)) of
   boundary_conditions.f90:791 (no density factor). */
static double maxwell(double x, double y, double z, double coeff, double mass, double u, double v,
                      double w, double temp) {
  double cx = (x - u); cx = cx * cx;
  double cy = (y - v); cy = cy * cy;
  double cz = (z - w); cz = cz * cz;
  double csq = cx + cy + cz;
  return coeff * exp(-csq * mass / temp);
}

/* Heat flux into the wall on `side` of a wall cell (boundary_conditions.f90:
   872-923 left, 1069-1115 right, 1262-1307 bottom, 1450-1495 top), in the
   reference's loop order: incoming half of f against the cell's bulk velocity,
   outgoing half of the ghost gh against the wall velocity; pc the cell's
   moments from the start of the step. */
static void wall_heat(orc_ctx* c, int side, int pos, const double* f0, const double* gh, const double* pc) {
  const ogrid* g = &c->g;
  int nvx = g->n[0], nvy = g->n[1], nvz = g->n[2];
  const double *X = g->ax[0], *Y = g->ax[1], *Z = g->ax[2];
  double xb = pc[DVM_P_U], yb = pc[DVM_P_V], zb = pc[DVM_P_W];
  const double* wl = c->cfg.wall[side];
  double u = wl[DVM_W_U], v = wl[DVM_W_V], w = wl[DVM_W_W];
  /* index ranges of the incoming and outgoing halves along the wall normal */
  int in0, in1, out0, out1;
  if (side == 0 || side == 2) {
    int neg = g->neg[side / 2], posi = g->pos[side / 2], mx = g->n[side / 2] - 1;
    in0 = 0; in1 = neg; out0 = posi; out1 = mx;
  } else {
    int neg = g->neg[side / 2], posi = g->pos[side / 2], mx = g->n[side / 2] - 1;
    in0 = posi; in1 = mx; out0 = 0; out1 = neg;
  }
  double hin = 0.0, hout = 0.0;
  for (int pass = 0; pass < 2; ++pass) {
    double ub = pass ? u : xb, vb = pass ? v : yb, wb = pass ? w : zb;
    const double* src = pass ? gh : f0;
    int a0 = pass ? out0 : in0, a1 = pass ? out1 : in1;
    double h = 0.0;
    for (int k = 0; k < nvz; ++k) {
      double z = Z[k], czsq = (z - wb) * (z - wb);
      if (side < 2) {
        for (int j = 0; j < nvy; ++j) {
          double y = Y[j], cysq = (y - vb) * (y - vb);
          for (int i = a0; i <= a1; ++i) {
            double x = X[i], cxsq = (x - ub) * (x - ub);
            double csq = cxsq + cysq + czsq;
            h = h + x * src[i + nvx * (j + nvy * k)] * csq;
          }
        }
      } else {
        for (int j = a0; j <= a1; ++j) {
          double y = Y[j], cysq = (y - vb) * (y - vb);
          for (int i = 0; i < nvx; ++i) {
            double x = X[i], cxsq = (x - ub) * (x - ub);
            double csq = cxsq + cysq + czsq;
            h = h + y * src[i + nvx * (j + nvy * k)] * csq;
          }
        }
      }
    }
    if (pass) hout = h; else hin = h;
  }
  int along = c->cfg.nx > c->cfg.ny ? c->cfg.nx : c->cfg.ny;
  c->heat[side * along + pos] = 0.5 * c->cfg.mass * (hin + hout);
}

int orc_set_wall_heat_flux_freq(orc_ctx* c, int32_t every) {
  if (every < 0) return set_err(c, DVM_E_ARG, "every must be >= 0");
  if (every > 0 && c->cfg.dim != DVM_DIM_2D) return set_err(c, DVM_E_ARG, "wall heat fluxes need a 2-D grid");
  int along = c->cfg.nx > c->cfg.ny ? c->cfg.nx : c->cfg.ny;
  if (every > 0 && !c->heat) c->heat = (double*)calloc(4 * (size_t)along, sizeof(double));
  c->heat_every = every;
  return DVM_OK;
}

int orc_wall_heat_flux(orc_ctx* c, double* q, int32_t* ntime_out) {
  int along = c->cfg.nx > c->cfg.ny ? c->cfg.nx : c->cfg.ny;
  for (int i = 0; i < 4 * along; ++i) q[i] = (c->heat && c->heat_ntime > 0) ? c->heat[i] : 0.0;
  if (ntime_out) *ntime_out = (c->heat && c->heat_ntime > 0) ? c->heat_ntime : -1;
  return DVM_OK;
}

/* boundary_conditions.f90:630-1656 set_boundary_nodes: builds the 4 ghosts of
   boundary cell (ix,iy) (0-based) and writes its upwind update into fc. */
static void boundary_cell_2d(orc_ctx* c, int ix, int iy, double* fc, double* ghost /* [4][N] */) {
  const dvm_config* cfg = &c->cfg;
  const ogrid* g = &c->g;
  int N = g->num_points;
  int nvx = g->n[0], nvy = g->n[1], nvz = g->n[2];
  int nxs = cfg->nx, nys = cfg->ny;
  const double *X = g->ax[0], *Y = g->ax[1], *Z = g->ax[2];
  const double *BX = g->beta[0], *BY = g->beta[1], *BZ = g->beta[2];
  double mass = cfg->mass, acc = cfg->acc_coeff;
  const double* f0 = c->f + (size_t)(ix + nxs * iy) * N;
  double* out = fc + (size_t)(ix + nxs * iy) * N;
  int i_zero = g->zero[0], i_neg = g->neg[0], i_pos = g->pos[0], i_max = nvx - 1;
  int j_zero = g->zero[1], j_neg = g->neg[1], j_pos = g->pos[1], j_max = nvy - 1;
  int i_zero_flag = (i_zero != 0 && i_zero != i_neg && i_zero != i_pos && i_zero != i_max);
  int j_zero_flag = (j_zero != 0 && j_zero != j_neg && j_zero != j_pos && j_zero != j_max);
  double x_factor = cfg->dt / cfg->dx, y_factor = cfg->dt / cfg->dy;
#define IDX(i, j, k) ((i) + nvx * ((j) + nvy * (k)))
  for (int l = 0; l < N; ++l) out[l] = 0.0;
  memset(ghost, 0, sizeof(double) * 4 * (size_t)N);
  /* ---- left ghost (1) */
  double* g1 = ghost;
  if (ix == 0) {
    if (cfg->bc_left == DVM_BC_WALL) {
      const double* wl = cfg->wall[0];
      double kin_temp = wl[DVM_W_TEMP], u = wl[DVM_W_U], v = wl[DVM_W_V], w = wl[DVM_W_W];
      double flux_in = 0.0;
      for (int k = 0; k < nvz; ++k)
        for (int j = 0; j < nvy; ++j)
          for (int i = 0; i <= i_neg; ++i) flux_in = flux_in + fabs(X[i]) * acc * f0[IDX(i, j, k)];
      if (acc > 0.0) {
        double mc = sqrt(mass * mass * mass / (PI_REF * PI_REF * PI_REF * kin_temp * kin_temp * kin_temp));
        double flux_out = 0.0;
        for (int k = 0; k < nvz; ++k)
          for (int j = 0; j < nvy; ++j)
            for (int i = i_pos; i <= i_max; ++i) {
              double beta3 = BX[i] * BY[j] * BZ[k];
              double mv = maxwell(X[i], Y[j], Z[k], mc, mass, u, v, w, kin_temp);
              g1[IDX(i, j, k)] = mv * beta3;
              flux_out = flux_out + fabs(X[i]) * mv * beta3;
            }
        double ratio = flux_in / flux_out;
        for (int l = 0; l < N; ++l) g1[l] = ratio * g1[l];
      }
      if (acc < 1.0) {
        for (int k = 0; k < nvz; ++k)
          for (int j = 0; j < nvy; ++j)
            for (int i = 0; i <= i_neg; ++i) {
              int st[7]; double m[7];
              double dens = (1.0 - acc) * f0[IDX(i, j, k)];
              remap1(g, -X[i], Y[j], Z[k], 1.0, st, m);
              scatter_stencil(g1, g, st, m, dens);
            }
      }
    }
    if (cfg->bc_left == DVM_BC_COUETTE) memcpy(g1, f0, sizeof(double) * N);
  } else {
    memcpy(g1, c->f + (size_t)(ix - 1 + nxs * iy) * N, sizeof(double) * N);
    if (i_zero_flag)
      for (int k = 0; k < nvz; ++k) for (int j = 0; j < nvy; ++j) g1[IDX(i_zero, j, k)] = 0.0;
  }
  /* ---- right ghost (2) */
  double* g2 = ghost + N;
  if (ix == nxs - 1) {
    if (cfg->bc_right == DVM_BC_WALL) {
      const double* wl = cfg->wall[1];
      double kin_temp = wl[DVM_W_TEMP], u = wl[DVM_W_U], v = wl[DVM_W_V], w = wl[DVM_W_W];
      double flux_in = 0.0;
      for (int k = 0; k < nvz; ++k)
        for (int j = 0; j < nvy; ++j)
          for (int i = i_pos; i <= i_max; ++i) flux_in = flux_in + fabs(X[i]) * acc * f0[IDX(i, j, k)];
      if (acc > 0.0) {
        double mc = sqrt(mass * mass * mass / (PI_REF * PI_REF * PI_REF * kin_temp * kin_temp * kin_temp));
        double flux_out = 0.0;
        for (int k = 0; k < nvz; ++k)
          for (int j = 0; j < nvy; ++j)
            for (int i = 0; i <= i_neg; ++i) {
              double beta3 = BX[i] * BY[j] * BZ[k];
              double mv = maxwell(X[i], Y[j], Z[k], mc, mass, u, v, w, kin_temp);
              g2[IDX(i, j, k)] = mv * beta3;
              flux_out = flux_out + fabs(X[i]) * mv * beta3;
            }
        double ratio = flux_in / flux_out;
        for (int l = 0; l < N; ++l) g2[l] = ratio * g2[l];
      }
      if (acc < 1.0) {
        for (int k = 0; k < nvz; ++k)
          for (int j = 0; j < nvy; ++j)
            for (int i = i_pos; i <= i_max; ++i) {
              int st[7]; double m[7];
              double dens = (1.0 - acc) * f0[IDX(i, j, k)];
              remap1(g, -X[i], Y[j], Z[k], 1.0, st, m);
              scatter_stencil(g2, g, st, m, dens);
            }
      }
    }
    if (cfg->bc_right == DVM_BC_COUETTE) memcpy(g2, c->f + (size_t)(0 + nxs * iy) * N, sizeof(double) * N);
  } else {
    memcpy(g2, c->f + (size_t)(ix + 1 + nxs * iy) * N, sizeof(double) * N);
    if (i_zero_flag)
      for (int k = 0; k < nvz; ++k) for (int j = 0; j < nvy; ++j) g2[IDX(i_zero, j, k)] = 0.0;
  }
  /* ---- bottom ghost (3): always a wall when iy == 0 (boundary_conditions.f90:1145) */
  double* g3 = ghost + 2 * (size_t)N;
  if (iy == 0) {
    const double* wl = cfg->wall[2];
    double kin_temp = wl[DVM_W_TEMP], u = wl[DVM_W_U], v = wl[DVM_W_V], w = wl[DVM_W_W];
    double flux_in = 0.0;
    for (int k = 0; k < nvz; ++k)
      for (int i = 0; i <= i_max; ++i)
        for (int j = 0; j <= j_neg; ++j) flux_in = flux_in + fabs(Y[j]) * acc * f0[IDX(i, j, k)];
    if (acc > 0.0) {
      double mc = sqrt(mass * mass * mass / (PI_REF * PI_REF * PI_REF * kin_temp * kin_temp * kin_temp));
      double flux_out = 0.0;
      for (int k = 0; k < nvz; ++k)
        for (int i = 0; i <= i_max; ++i)
          for (int j = j_pos; j <= j_max; ++j) {
            double beta3 = BX[i] * BY[j] * BZ[k];
            double mv = maxwell(X[i], Y[j], Z[k], mc, mass, u, v, w, kin_temp);
            g3[IDX(i, j, k)] = mv * beta3;
            flux_out = flux_out + fabs(Y[j]) * mv * beta3;
          }
      double ratio = flux_in / flux_out;
      for (int l = 0; l < N; ++l) g3[l] = ratio * g3[l];
    }
    if (acc < 1.0) {
      for (int k = 0; k < nvz; ++k)
        for (int j = 0; j <= j_neg; ++j)
          for (int i = 0; i <= i_max; ++i) {
            int st[7]; double m[7];
            double dens = (1.0 - acc) * f0[IDX(i, j, k)];
            remap1(g, X[i], -Y[j], Z[k], 1.0, st, m);
            scatter_stencil(g3, g, st, m, dens);
          }
    }
  } else {
    memcpy(g3, c->f + (size_t)(ix + nxs * (iy - 1)) * N, sizeof(double) * N);
    if (j_zero_flag)
      for (int k = 0; k < nvz; ++k) for (int i = 0; i < nvx; ++i) g3[IDX(i, j_zero, k)] = 0.0;
  }
  /* ---- top ghost (4) */
  double* g4 = ghost + 3 * (size_t)N;
  if (iy == nys - 1) {
    const double* wl = cfg->wall[3];
    double kin_temp = wl[DVM_W_TEMP], u = wl[DVM_W_U], v = wl[DVM_W_V], w = wl[DVM_W_W];
    double flux_in = 0.0;
    for (int k = 0; k < nvz; ++k)
      for (int i = 0; i <= i_max; ++i)
        for (int j = j_pos; j <= j_max; ++j) flux_in = flux_in + fabs(Y[j]) * acc * f0[IDX(i, j, k)];
    if (acc > 0.0) {
      double mc = sqrt(mass * mass * mass / (PI_REF * PI_REF * PI_REF * kin_temp * kin_temp * kin_temp));
      double flux_out = 0.0;
      for (int k = 0; k < nvz; ++k)
        for (int i = 0; i <= i_max; ++i)
          for (int j = 0; j <= j_neg; ++j) {
            double beta3 = BX[i] * BY[j] * BZ[k];
            double mv = maxwell(X[i], Y[j], Z[k], mc, mass, u, v, w, kin_temp);
            g4[IDX(i, j, k)] = mv * beta3;
            flux_out = flux_out + fabs(Y[j]) * mv * beta3;
          }
      double ratio = flux_in / flux_out;
      for (int l = 0; l < N; ++l) g4[l] = ratio * g4[l];
    }
    if (acc < 1.0) {
      for (int k = 0; k < nvz; ++k)
        for (int j = j_pos; j <= j_max; ++j)
          for (int i = 0; i <= i_max; ++i) {
            int st[7]; double m[7];
            double dens = (1.0 - acc) * f0[IDX(i, j, k)];
            remap1(g, X[i], -Y[j], Z[k], 1.0, st, m);
            scatter_stencil(g4, g, st, m, dens);
          }
    }
  } else {
    memcpy(g4, c->f + (size_t)(ix + nxs * (iy + 1)) * N, sizeof(double) * N);
    if (j_zero_flag)
      for (int k = 0; k < nvz; ++k) for (int i = 0; i < nvx; ++i) g4[IDX(i, j_zero, k)] = 0.0;
  }
  if (c->heat_on) {
    const double* pc = c->props + (size_t)(ix + nxs * iy) * DVM_NPROP;
    if (ix == 0 && cfg->bc_left == DVM_BC_WALL) wall_heat(c, 0, iy, f0, g1, pc);
    if (ix == nxs - 1 && cfg->bc_right == DVM_BC_WALL) wall_heat(c, 1, iy, f0, g2, pc);
    if (iy == 0) wall_heat(c, 2, ix, f0, g3, pc);
    if (iy == nys - 1) wall_heat(c, 3, ix, f0, g4, pc);
  }
  /* ---- boundary-cell upwind update (boundary_conditions.f90:1548-1643) */
  for (int k = 0; k < nvz; ++k) {
    for (int j = 0; j <= j_neg; ++j) {
      double y_cfl = y_factor * Y[j];
      for (int i = 0; i <= i_neg; ++i) {
        double x_cfl = x_factor * X[i];
        int l = IDX(i, j, k);
        out[l] = f0[l] + fabs(x_cfl) * (g2[l] - f0[l]) + fabs(y_cfl) * (g4[l] - f0[l]);
      }
      if (i_zero_flag) {
        int l = IDX(i_zero, j, k);
        out[l] = f0[l] + fabs(y_cfl) * (g4[l] - f0[l]);
      }
      for (int i = i_pos; i <= i_max; ++i) {
        double x_cfl = x_factor * X[i];
        int l = IDX(i, j, k);
        out[l] = f0[l] + fabs(x_cfl) * (g1[l] - f0[l]) + fabs(y_cfl) * (g4[l] - f0[l]);
      }
    }
    if (j_zero_flag) {
      for (int i = 0; i <= i_neg; ++i) {
        double x_cfl = x_factor * X[i];
        int l = IDX(i, j_zero, k);
        out[l] = f0[l] + fabs(x_cfl) * (g2[l] - f0[l]);
      }
      if (i_zero_flag) out[IDX(i_zero, j_zero, k)] = f0[IDX(i_zero, j_zero, k)];
      for (int i = i_pos; i <= i_max; ++i) {
        double x_cfl = x_factor * X[i];
        int l = IDX(i, j_zero, k);
        out[l] = f0[l] + fabs(x_cfl) * (g1[l] - f0[l]);
      }
    }
    for (int j = j_pos; j <= j_max; ++j) {
      double y_cfl = y_factor * Y[j];
      for (int i = 0; i <= i_neg; ++i) {
        double x_cfl = x_factor * X[i];
        int l = IDX(i, j, k);
        out[l] = f0[l] + fabs(x_cfl) * (g2[l] - f0[l]) + fabs(y_cfl) * (g3[l] - f0[l]);
      }
      if (i_zero_flag) {
        int l = IDX(i_zero, j, k);
        out[l] = f0[l] + fabs(y_cfl) * (g3[l] - f0[l]);
      }
      for (int i = i_pos; i <= i_max; ++i) {
        double x_cfl = x_factor * X[i];
        int l = IDX(i, j, k);
        out[l] = f0[l] + fabs(x_cfl) * (g1[l] - f0[l]) + fabs(y_cfl) * (g3[l] - f0[l]);
      }
    }
  }
#undef IDX
  (void)i_zero; (void)j_zero;
}

/* convection.f90:119-276 first_order_scheme: interior cells of the 2-D grid. */
static void interior_2d(orc_ctx* c, int ix, int iy, double* fc) {
  const dvm_config* cfg = &c->cfg;
  const ogrid* g = &c->g;
  int N = g->num_points, nxs = cfg->nx;
  double factor_x = cfg->dt / cfg->dx, factor_y = cfg->dt / cfg->dy;
  const double* f0 = c->f + (size_t)(ix + nxs * iy) * N;
  const double* fxp = c->f + (size_t)(ix + 1 + nxs * iy) * N;
  const double* fxm = c->f + (size_t)(ix - 1 + nxs * iy) * N;
  const double* fyp = c->f + (size_t)(ix + nxs * (iy + 1)) * N;
  const double* fym = c->f + (size_t)(ix + nxs * (iy - 1)) * N;
  double* out = fc + (size_t)(ix + nxs * iy) * N;
  int l = 0;
  for (int k = 0; k < g->n[2]; ++k)
    for (int j = 0; j < g->n[1]; ++j)
      for (int i = 0; i < g->n[0]; ++i, ++l) {
        double xv = g->ax[0][i], yv = g->ax[1][j];
        double cfl_x = factor_x * xv, cfl_y = factor_y * yv;
        double p0 = f0[l];
        double px = (xv < 0.0) ? fxp[l] : fxm[l];
        double py = (yv < 0.0) ? fyp[l] : fym[l];
        out[l] = p0 + fabs(cfl_x) * (px - p0) + fabs(cfl_y) * (py - p0);
      }
}

/* 1-D upwind and boundaries of the zone-grid solver
   (grid_ref_code convection.f90:111-256, boundary_conditions.f90:319-444,783-973). */
/* Limited slope of 1-D cell ix for one node (stride N): minmod of the two
   one-sided differences; zero in the boundary cells. Second-order upwind is
   not in the reference (its 1-D solver has first-order upwind only,
   grid_ref_code convection.f90:111-256); this is the library's definition. */
static double minmod(double a, double b) {
  if (a * b <= 0.0) return 0.0;
  return fabs(a) < fabs(b) ? a : b;
}

static double slope1d(const double* f, int ix, int nxs, size_t N, int l) {
  if (ix <= 0 || ix >= nxs - 1) return 0.0;
  const double p0 = f[(size_t)ix * N + l];
  return minmod(f[(size_t)(ix + 1) * N + l] - p0, p0 - f[(size_t)(ix - 1) * N + l]);
}

/* Interior 1-D update of node l of cell ix: the reference's donor-cell
   upwind (convection.f90:111-256 of the zone-grid solver), plus with order 2
   the limited Lax-Wendroff correction
     x >= 0: - nu (1 - nu) / 2 (s_i - s_(i-1)),   nu = dt/dx x
     x <  0: + mu (1 - mu) / 2 (s_i - s_(i+1)),   mu = -nu. */
static double upwind1d(const double* f, int ix, int nxs, size_t N, int l, double cfl, double xv, int order) {
  const double p0 = f[(size_t)ix * N + l];
  if (xv < 0.0) {
    double out = p0 - cfl * (f[(size_t)(ix + 1) * N + l] - p0);
    if (order == 2) {
      const double mu = -cfl;
      out = out + (mu * (1.0 - mu) * 0.5) * (slope1d(f, ix, nxs, N, l) - slope1d(f, ix + 1, nxs, N, l));
    }
    return out;
  }
  double out = p0 - cfl * (p0 - f[(size_t)(ix - 1) * N + l]);
  if (order == 2) out = out - (cfl * (1.0 - cfl) * 0.5) * (slope1d(f, ix, nxs, N, l) - slope1d(f, ix - 1, nxs, N, l));
  return out;
}

/* 1-D wall cell, apply_diffuse_reflection of the zone-grid solver
   (grid_ref_code boundary_conditions.f90:566-781): the cell is zeroed; its
   outgoing half becomes the wall Maxwellian times beta^3, scaled to
   acc * (sum of f over the incoming half, not flux-weighted); then, node by
   node over the incoming half in (k, j, i) order, the node takes the upwind
   value from the neighbour cell (dt/dx |x| (f - f_nb)) and (1 - acc) f of the
   node is remapped from (-x, y, z) and added (apply_bc_remapping :975-1053),
   so a later node's assignment overwrites what earlier nodes added to it.
   The zone-grid solver never sets the i_neg / i_pos it reads here; the
   trunk's zero points (velocity_grid.f90:653-689) are used. */
static int wall_1d(orc_ctx* c, int side, double* out) {
  const dvm_config* cfg = &c->cfg;
  const ogrid* g = &c->g;
  const int N = g->num_points, nxs = cfg->nx, n0 = g->n[0];
  const int ix = side == 0 ? 0 : nxs - 1, nb = side == 0 ? 1 : nxs - 2;
  const double* f0 = c->f + (size_t)ix * N;
  const double* fn = c->f + (size_t)nb * N;
  const double acc = cfg->acc_coeff, mass = cfg->mass;
  const double* wl = cfg->wall[side];
  const double kt = wl[DVM_W_TEMP], u = wl[DVM_W_U], v = wl[DVM_W_V], w = wl[DVM_W_W];
  int in0, in1, out0, out1;
  if (side == 0) { in0 = 0; in1 = g->neg[0]; out0 = g->pos[0]; out1 = n0 - 1; }
  else { in0 = g->pos[0]; in1 = n0 - 1; out0 = 0; out1 = g->neg[0]; }
  memset(out, 0, sizeof(double) * (size_t)N);
  double dens = 0.0;
  for (int k = 0; k < g->n[2]; ++k)
    for (int j = 0; j < g->n[1]; ++j)
      for (int i = in0; i <= in1; ++i) dens = dens + f0[i + n0 * (j + g->n[1] * k)];
  dens = acc * dens;
  const double mc = sqrt(mass * mass * mass / (PI_REF * PI_REF * PI_REF * kt * kt * kt));
  double computed = 0.0;
  for (int k = 0; k < g->n[2]; ++k)
    for (int j = 0; j < g->n[1]; ++j)
      for (int i = out0; i <= out1; ++i) {
        const double beta3 = g->beta[0][i] * g->beta[1][j] * g->beta[2][k];
        const double mv = maxwell(g->ax[0][i], g->ax[1][j], g->ax[2][k], mc, mass, u, v, w, kt);
        out[i + n0 * (j + g->n[1] * k)] = mv * beta3;
        computed = computed + mv * beta3;
      }
  const double ratio = dens / computed;
  for (int l = 0; l < N; ++l) out[l] = ratio * out[l];
  const double factor = cfg->dt / cfg->dx;
  for (int k = 0; k < g->n[2]; ++k)
    for (int j = 0; j < g->n[1]; ++j)
      for (int i = in0; i <= in1; ++i) {
        const int l = i + n0 * (j + g->n[1] * k);
        const double x = g->ax[0][i];
        out[l] = f0[l] - fabs(factor * x) * (f0[l] - fn[l]);
        const double ds = (1.0 - acc) * f0[l];
        int st[7];
        double m[7];
        remap1(g, -x, g->ax[1][j], g->ax[2][k], 1.0, st, m);
        if (st[6] == 0) return set_err(c, DVM_E_REMAP, "1-D wall: remap found no exterior point");
        scatter_stencil(out, g, st, m, ds);
      }
  return DVM_OK;
}

static int convect_1d(orc_ctx* c, double* fc) {
  const dvm_config* cfg = &c->cfg;
  const ogrid* g = &c->g;
  int N = g->num_points, nxs = cfg->nx;
  const int order = cfg->conv_order == 2 ? 2 : 1;
  double factor = cfg->dt / cfg->dx;
  for (int ix = 1; ix < nxs - 1; ++ix) {
    double* out = fc + (size_t)ix * N;
    int l = 0;
    for (int k = 0; k < g->n[2]; ++k)
      for (int j = 0; j < g->n[1]; ++j)
        for (int i = 0; i < g->n[0]; ++i, ++l) {
          double xv = g->ax[0][i];
          out[l] = upwind1d(c->f, ix, nxs, (size_t)N, l, factor * xv, xv, order);
        }
  }
  for (int side = 0; side < 2; ++side) {
    int bc = side == 0 ? cfg->bc_left : cfg->bc_right;
    int ix = side == 0 ? 0 : nxs - 1;
    double* out = fc + (size_t)ix * N;
    if (bc == DVM_BC_INOUT) {
      /* apply_fixed_in_out_flow: Maxwellian at the wall state with v = w = 0 */
      const double* wl = cfg->wall[side];
      double dens = wl[DVM_W_DENS], temp = wl[DVM_W_TEMP], u = wl[DVM_W_U], mass = cfg->mass;
      double mc = dens * sqrt(mass * mass * mass / (PI_REF * PI_REF * PI_REF * temp * temp * temp));
      int l = 0;
      for (int k = 0; k < g->n[2]; ++k)
        for (int j = 0; j < g->n[1]; ++j)
          for (int i = 0; i < g->n[0]; ++i, ++l) {
            double beta3 = g->beta[0][i] * g->beta[1][j] * g->beta[2][k];
            out[l] = maxwell(g->ax[0][i], g->ax[1][j], g->ax[2][k], mc, mass, u, 0.0, 0.0, temp) * beta3;
          }
    } else if (bc == DVM_BC_ZEROGRAD) {
      /* apply_zero_gradient_flow: boundary cell takes its neighbour's new values */
      int nb = side == 0 ? 1 : nxs - 2;
      memcpy(out, fc + (size_t)nb * N, sizeof(double) * N);
    } else if (bc == DVM_BC_WALL) {
      int rc = wall_1d(c, side, out);
      if (rc) return rc;
    } else {
      /* other 1-D boundary types keep the cell unchanged */
      memcpy(out, c->f + (size_t)ix * N, sizeof(double) * N);
    }
  }
  return DVM_OK;
}

int orc_convect(orc_ctx* c, int32_t species, int32_t ntime) {
  if (species != 0) return set_err(c, DVM_E_ARG, "only species 0 is supported");
  if (!c->cfg.enable_convection || c->cfg.dim == DVM_DIM_0D) return DVM_OK;
  c->heat_on = c->heat && c->heat_every > 0 && ntime % c->heat_every == 0;
  if (c->heat_on) c->heat_ntime = ntime;
  int N = c->g.num_points;
  double* fc = c->fd;
  if (c->cfg.dim == DVM_DIM_1D) {
    if (c->cfg.nx < 3) return set_err(c, DVM_E_ARG, "1-D convection needs nx >= 3");
    int rc = convect_1d(c, fc);
    if (rc) return rc;
  } else {
    int nxs = c->cfg.nx, nys = c->cfg.ny;
    double* ghost = (double*)malloc(sizeof(double) * 4 * (size_t)N);
    for (int iy = 0; iy < nys; ++iy)
      for (int ix = 0; ix < nxs; ++ix) {
        if (ix == 0 || iy == 0 || ix == nxs - 1 || iy == nys - 1) boundary_cell_2d(c, ix, iy, fc, ghost);
        else interior_2d(c, ix, iy, fc);
      }
    free(ghost);
  }
  memcpy(c->f, fc, sizeof(double) * (size_t)N * c->ncell);
  return DVM_OK;
}

int orc_step(orc_ctx* c, int32_t ntime) {
  int rc = orc_collide(c, ntime, NULL);
  if (rc) return rc;
  rc = orc_convect(c, 0, ntime);
  if (rc) return rc;
  return orc_moments(c, NULL);
}

/* Scalar building blocks for the golden tests. */
double orc_maxwellian(double x, double y, double z, double coeff, double mass, double u, double v,
                      double w, double temp) {
  return maxwell(x, y, z, coeff, mass, u, v, w, temp);
}

/* calculate_cross_section (scaling.f90:163-201) for a like pair of diameter d. */
double orc_cross_section(int32_t coeff_kind, double m_red, double omega, double d, double visc,
                         double tv) {
  double sigma_vhs = 0.25 * (d + d) * (d + d);
  if (coeff_kind == DVM_CS_BIRD) return sigma_vhs * pow(m_red, 0.5 - omega) / gamma_lanczos(2.5 - omega);
  if (coeff_kind == DVM_CS_VKI)
    return (3.32335097045 * pow(m_red, 1.0 - omega) * pow(tv, omega)) / (gamma_lanczos(4.5 - omega) * visc);
  return sigma_vhs;
}
