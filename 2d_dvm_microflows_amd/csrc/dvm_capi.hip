// dvm_capi.hip -- context management and the extern "C" entry points of
// include/dvm_hip.h. One translation unit: the kernels are included here.
#include "dvm_kernels.hip"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <type_traits>
#include <vector>

using namespace dvm;

struct dvm_ctx {
  dvm_config cfg;
  int ncell = 0, N = 0;
  hipStream_t stream = nullptr;
  // phase events: per slot, boundaries of the DVM_NPHASE phases
  // (cdf, prep, collide, conserve, walls, sweep, moments)
  static constexpr int kSlots = 16;
  hipEvent_t ev[kSlots][DVM_NPHASE + 1] = {};
  bool slot_used[kSlots] = {};
  int slot = 0;
  double phase_ms[DVM_NPHASE] = {};
  long long timed_steps = 0;
  // host copy of the grid
  std::vector<double> ax[3], beta[3];
  Grid g{};                      // device pointers
  double* d_axes = nullptr;      // [ax0 ax1 ax2 beta0 beta1 beta2]
  // resident state
  double *f = nullptr, *fd = nullptr, *eqv = nullptr, *cdf = nullptr;  // cdf: [C][2][N+1]
  int2* tabs = nullptr;                  // [C][2][1000]
  double *eqpar = nullptr, *props = nullptr, *factors = nullptr;
  int *counts = nullptr, *err = nullptr;
  // walls
  double *wallM = nullptr, *ratio = nullptr, *spec = nullptr, *inflow = nullptr;
  CdfParams cdfp{};
  CollParams collp{};
  ConvParams convp{};
  bool lds_collide = false;      // the FP64 slice fits LDS (window = grid)
  bool uniform_axes = false;     // equally spaced axes (the FP32 remap rounds to the nearest node)
  WinDims wd{};                  // LDS window extents
  size_t coll_lds = 0;
  int coll_bs = 1024;            // lanes per collision workgroup (256, 512, 768 or 1024)
  int pick_k = 1;                // samples per lane picked in lockstep (1, 2 or 4)
  int split = 1;                 // workgroups per cell in the collision kernel
  int cus = 256;                 // compute units of the device
  int heat_every = 0;            // wall heat fluxes every heat_every steps (0: off)
  int heat_ntime = -1;           // step of the values in d_heat
  double* d_heat = nullptr;      // [4][max(nx,ny)] wall heat fluxes
  double* dacc = nullptr;        // [C][N] collision change of the FP32 path
  unsigned char* sgn = nullptr;  // [C][2 (N+1)] bytes: deviation sign bits (ceil(N/8) bytes used per cell)
  double* cmat = nullptr;        // [C][kCmat] FP32 path: conservation factor of the moments pass
  int fast = 0;                  // the FP32 collision takes collide_fast (see collide_async)
  unsigned* tick = nullptr;      // [2] ticket and finish counters of k_collide_fast_cells (zero between launches)
  int fast_wgs = 0;              // workgroups of k_collide_fast_cells (0: not used)
  size_t fast_lds = 0;           // its dynamic LDS bytes
  bool fast_tab_lds = false;     // the bucket tables staged in LDS beside the sign words
  bool cmat_valid = false;       // cmat was computed from the current f (by moments_async)
  // explicit uniform list (dvm_set_uniform_stream): su_n values, read from
  // su_cursor on; sbase [C] per-cell offsets of the current collision
  double* d_su = nullptr;
  long long su_n = 0, su_cursor = 0;
  long long *d_sbase = nullptr, *d_stotal = nullptr;
  bool f_valid = false, props_valid = false;
  int trace_ready = -1;
  // device error flags of earlier dvm_step calls, copied to pinned host
  // memory behind each step and read at a later step once that copy is done
  int* err_host = nullptr;
  hipEvent_t err_ev = nullptr;
  bool err_pending = false;
  double* scratch = nullptr;     // this device's address of the file-scope scratch array
  char msg[512] = {0};
};

// Scratch of the allocation-free entry point (dvm_l1_noalloc): a file-scope
// device array, one copy per GPU the library runs on, held until the process
// ends. Host code never passes `scratch` itself (in host code the name is a
// stand-in, not a device address): each device's address is looked up once,
// with that device current, and kept in g_scratch_ptr.
constexpr long long kScratchDoubles = 1LL << 20;  // 8 MiB: the L1 staging of up to 2^20 cells
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
static double* g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// The current device's address of `scratch` (looked up on the first call on
// that device); nullptr with a message on failure.
static double* scratch_address(char* msg, size_t n) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    snprintf(msg, n, "scratch: no current device below %d", kMaxDevices);
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      snprintf(msg, n, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_ptr[dev] = (double*)ptr;
  }
  return g_scratch_ptr[dev];
}

static char g_create_msg[512] = "";

static int fail(dvm_ctx* c, int code, const char* fmt, ...) {
  char* dst = c ? c->msg : g_create_msg;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(dst, 512, fmt, ap);
  va_end(ap);
  return code;
}

#define HIP_TRY(c, x)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) return fail((c), DVM_E_HIP, "%s: %s", #x, hipGetErrorString(e_)); \
  } while (0)

template <class T>
static int dalloc(dvm_ctx* c, T** p, size_t n) {
  if (hipMalloc((void**)p, n * sizeof(T) + 16) != hipSuccess)
    return fail(c, DVM_E_NOMEM, "hipMalloc of %zu bytes failed", n * sizeof(T));
  return DVM_OK;
}

// velocity_grid.f90:653-689 find_zero_point (host side, once per grid)
static void zero_point(const std::vector<double>& a, int& zero, int& pos, int& neg) {
  const int mx = (int)a.size() - 1;
  if (a[0] > 0.0) { zero = -1; pos = 0; neg = -1; return; }
  for (int i = 1; i <= mx; ++i) {
    if (a[i] > 0.0) {
      zero = i - 1; pos = i; neg = i - 1;
      if (a[i - 1] > -kDoubleTol) neg = i - 2;
      return;
    }
  }
  zero = mx + 1; pos = mx + 1; neg = mx;
}

// math_utils.f90:83-135 gamma_calculator (Lanczos coefficients as listed there)
static double gamma_ref(double x) {
  static const double a[9] = {0.99999999999980993, 676.5203681218851, -1259.1392167224028,
                              771.32342877765313, -176.61502916214059, 12.507343278686905,
                              -0.13857109526572012, 9.9843695780195716e-6, 1.5056327351493116e-7};
  if (x < 0.5) return kPiRef / (sin(kPiRef * x) * gamma_ref(1.0 - x));
  x = x - 1.0;
  double y = a[0];
  for (int i = 1; i <= 8; ++i) y = y + a[i] / (x + (double)i);
  const double z = x + 7.0 + 0.5;
  return sqrt(2.0 * kPiRef) * pow(z, x + 0.5) * exp(-z) * y;
}

static double maxwell_host(double x, double y, double z, double coeff, double mass, double u, double v, double w,
                           double temp) {
  double cx = (x - u); cx = cx * cx;
  double cy = (y - v); cy = cy * cy;
  double cz = (z - w); cz = cz * cz;
  const double csq = cx + cy + cz;
  return coeff * exp(-csq * mass / temp);
}

static int validate(const dvm_config* cfg) {
  if (!cfg) return fail(nullptr, DVM_E_ARG, "null config");
  if (cfg->dim < DVM_DIM_0D || cfg->dim > DVM_DIM_2D) return fail(nullptr, DVM_E_ARG, "dim must be 0, 1 or 2");
  if (cfg->nx < 1 || cfg->ny < 1) return fail(nullptr, DVM_E_ARG, "nx, ny must be >= 1");
  if (cfg->dim == DVM_DIM_0D && (cfg->nx != 1 || cfg->ny != 1))
    return fail(nullptr, DVM_E_ARG, "space-homogeneous runs have one cell");
  if (cfg->dim == DVM_DIM_1D && cfg->ny != 1) return fail(nullptr, DVM_E_ARG, "1-D runs need ny = 1");
  if (cfg->nx > 65535) return fail(nullptr, DVM_E_ARG, "nx must be <= 65535");
  if (cfg->dim == DVM_DIM_1D && cfg->enable_convection && cfg->nx < 3)
    return fail(nullptr, DVM_E_ARG, "1-D convection needs nx >= 3");
  const int n[3] = {cfg->nvx, cfg->nvy, cfg->nvz};
  const double* a[3] = {cfg->vx, cfg->vy, cfg->vz};
  for (int d = 0; d < 3; ++d) {
    if (n[d] < 3 || n[d] > kMaxAxis) return fail(nullptr, DVM_E_ARG, "velocity axis %d needs 3..%d points", d, kMaxAxis);
    if (!a[d]) return fail(nullptr, DVM_E_ARG, "velocity axis %d is null", d);
    for (int i = 1; i < n[d]; ++i)
      if (!(a[d][i] > a[d][i - 1])) return fail(nullptr, DVM_E_ARG, "velocity axis %d is not increasing", d);
  }
  if (cfg->num_repl < 1) return fail(nullptr, DVM_E_ARG, "num_repl must be >= 1");
  if (cfg->coll_solver != DVM_COLL_NONE && cfg->coll_solver != DVM_COLL_FMC && cfg->coll_solver != DVM_COLL_VR)
    return fail(nullptr, DVM_E_ARG, "coll_solver must be 0 (none), 2 (full Monte Carlo) or 3 (variance reduction)");
  if (cfg->precision != DVM_PREC_FP64 && cfg->precision != DVM_PREC_FP32 && cfg->precision != DVM_PREC_FP32_REMAP32)
    return fail(nullptr, DVM_E_ARG, "precision must be DVM_PREC_FP64, DVM_PREC_FP32 or DVM_PREC_FP32_REMAP32");
  if (cfg->sample_mode != DVM_SAMPLES_REFERENCE && cfg->sample_mode != DVM_SAMPLES_EXPLICIT)
    return fail(nullptr, DVM_E_ARG, "invalid sample_mode");
  if (cfg->sample_mode == DVM_SAMPLES_REFERENCE && !(cfg->coln_rms > 0.0))
    return fail(nullptr, DVM_E_ARG, "coln_rms must be > 0 for the reference sample-count rule");
  if (cfg->sample_mode == DVM_SAMPLES_EXPLICIT && (cfg->samples_per_cell < 0 || cfg->samples_per_cell > kMaxCellSamples))
    return fail(nullptr, DVM_E_ARG, "samples_per_cell must be in [0, 2^30]");
  if (!(cfg->dt > 0.0) || !(cfg->kn > 0.0) || !(cfg->mass > 0.0)) return fail(nullptr, DVM_E_ARG, "dt, kn, mass must be > 0");
  if (cfg->enable_convection && cfg->dim != DVM_DIM_0D && (!(cfg->dx > 0.0) || (cfg->dim == DVM_DIM_2D && !(cfg->dy > 0.0))))
    return fail(nullptr, DVM_E_ARG, "dx (and dy in 2-D) must be > 0");
  if (!(cfg->acc_coeff >= 0.0 && cfg->acc_coeff <= 1.0)) return fail(nullptr, DVM_E_ARG, "acc_coeff must be in [0,1]");
  if (cfg->conv_order < 0 || cfg->conv_order > 2) return fail(nullptr, DVM_E_ARG, "conv_order must be 1 or 2");
  if (cfg->conv_order == 2 && cfg->dim != DVM_DIM_1D)
    return fail(nullptr, DVM_E_ARG, "the second-order sweep is available for 1-D runs");
  return DVM_OK;
}

extern "C" const char* dvm_last_error(const dvm_ctx* ctx) { return ctx ? ctx->msg : g_create_msg; }

extern "C" void dvm_ctx_destroy(dvm_ctx* c) {
  if (!c) return;
  void* ptrs[] = {c->d_heat, c->d_axes, c->f, c->fd, c->eqv, c->cdf, c->tabs, c->eqpar, c->props,
                  c->factors, c->counts, c->err, c->wallM, c->ratio, c->spec, c->inflow, c->dacc, c->sgn, c->cmat,
                  c->tick, c->d_su, c->d_sbase, c->d_stotal};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (auto& row : c->ev)
    for (auto& e : row)
      if (e) (void)hipEventDestroy(e);
  if (c->err_ev) (void)hipEventDestroy(c->err_ev);
  if (c->err_host) (void)hipHostFree(c->err_host);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

// uni: the FP32 path on equally spaced axes, whose remap takes the nearest
// node by rounding (the FP64 path always searches, as the reference); r32:
// with uni, the remap in FP32 as well (DVM_PREC_FP32_REMAP32).
template <int BS, int K>
static const void* collide_kernel_bsk(bool lds, bool f32, bool uni, bool r32) {
  if (f32 && uni && r32)
    return lds ? (const void*)k_collide<true, true, BS, K, false, true, true>
               : (const void*)k_collide<false, true, BS, K, false, true, true>;
  if (f32 && uni)
    return lds ? (const void*)k_collide<true, true, BS, K, false, true>
               : (const void*)k_collide<false, true, BS, K, false, true>;
  if (!lds)
    return f32 ? (const void*)k_collide<false, true, BS, K, false> : (const void*)k_collide<false, false, BS, K, false>;
  return f32 ? (const void*)k_collide<true, true, BS, K, false> : (const void*)k_collide<true, false, BS, K, false>;
}

// The collision kernel of the uniform-stream parity mode (kCollBS lanes, one
// pick per lane).
static const void* collide_stream_kernel(const dvm_ctx* c) {
  const bool f32 = prec_f32(c->cfg.precision);
  if (!c->lds_collide)
    return f32 ? (const void*)k_collide<false, true, kCollBS, 1, true>
               : (const void*)k_collide<false, false, kCollBS, 1, true>;
  return f32 ? (const void*)k_collide<true, true, kCollBS, 1, true> : (const void*)k_collide<true, false, kCollBS, 1, true>;
}

// Uniforms per sample: the two partner uniforms, then (phi, theta) per
// replenish location.
static int stream_per_sample(const dvm_ctx* c) { return 2 + 2 * c->cfg.num_repl; }

template <int BS>
static const void* collide_kernel_bs(int k, bool lds, bool f32, bool uni, bool r32) {
  switch (k) {
    case 1: return collide_kernel_bsk<BS, 1>(lds, f32, uni, r32);
    case 4: return collide_kernel_bsk<BS, 4>(lds, f32, uni, r32);
    default: return collide_kernel_bsk<BS, 2>(lds, f32, uni, r32);
  }
}

static const void* collide_kernel(const dvm_ctx* c) {
  // the fast FP32 path on kernels of its own (1024 lanes, one pick per lane)
  if (c->fast && c->coll_bs == 1024 && c->pick_k == 1)
    return c->lds_collide ? (const void*)k_collide<true, true, 1024, 1, false, true, false, true>
                          : (const void*)k_collide<false, true, 1024, 1, false, true, false, true>;
  const bool f32 = prec_f32(c->cfg.precision);
  const bool uni = c->uniform_axes;
  const bool r32 = c->cfg.precision == DVM_PREC_FP32_REMAP32;
  switch (c->coll_bs) {
    case 256: return collide_kernel_bs<256>(c->pick_k, c->lds_collide, f32, uni, r32);
    case 768: return collide_kernel_bs<768>(c->pick_k, c->lds_collide, f32, uni, r32);
    case 1024: return collide_kernel_bs<1024>(c->pick_k, c->lds_collide, f32, uni, r32);
    default: return collide_kernel_bs<512>(c->pick_k, c->lds_collide, f32, uni, r32);
  }
}

// LDS window nodes available to a collision workgroup next to the staged axes
// and the kernels' static LDS (the conservation reduction of the FP32 path).
constexpr size_t kStaticLds = 3072;
static_assert(sizeof(ConserveShared<1024>) + 64 <= kStaticLds, "static LDS of the collide kernels");
constexpr int kWindowCap = (int)((kLdsBytes - kStaticLds) / sizeof(double)) - kAxisSlots;

// Extents of the collision LDS window for a grid of n[0] x n[1] x n[2] nodes:
// the grid itself when it fits (returns true), else the largest box, near
// cubic, within kWindowCap nodes.
static bool window_dims(const int n[3], int w[3], int cap = kWindowCap) {
  const long long N = (long long)n[0] * n[1] * n[2];
  if (N <= cap) {
    for (int d = 0; d < 3; ++d) w[d] = n[d];
    return true;
  }
  int e = 1;
  while ((long long)(e + 1) * (e + 1) * (e + 1) <= cap) ++e;
  for (int d = 0; d < 3; ++d) w[d] = n[d] < e ? n[d] : e;
  for (bool grew = true; grew;) {
    grew = false;
    for (int d = 0; d < 3; ++d) {
      const long long rest = (long long)w[(d + 1) % 3] * w[(d + 2) % 3];
      if (w[d] < n[d] && (w[d] + 1) * rest <= cap) { ++w[d]; grew = true; }
    }
  }
  return false;
}

// Bytes of a cell's deviation sign words in LDS (the fast FP32 sampling).
static int sign_lds_bytes(long long N) { return 4 * (int)((N + 31) >> 5); }

// The LDS window cap of a fast FP32 collision, whose sign words share the LDS.
static int fast_window_cap(long long N) { return kWindowCap - (sign_lds_bytes(N) + 7) / 8; }

// Host copy of one velocity grid (velocity_grid.f90:271-597): axes, the
// trapezoid weights beta, the zero points, and the packed device image
// [ax0 ax1 ax2 beta0 beta1 beta2]. Device pointers are set by the caller.
struct HostGrid {
  std::vector<double> ax[3], beta[3], packed;
  Grid g{};
  double b3min = 0.0;
  bool uniform = false;  // every axis equally spaced (to 1e-9 of its spacing)
};

static void build_host_grid(const dvm_config* cfg, HostGrid& h) {
  const int n[3] = {cfg->nvx, cfg->nvy, cfg->nvz};
  const double* src[3] = {cfg->vx, cfg->vy, cfg->vz};
  h.packed.clear();
  for (int d = 0; d < 3; ++d) {
    h.ax[d].assign(src[d], src[d] + n[d]);
    auto& a = h.ax[d];
    auto& b = h.beta[d];
    b.resize(n[d]);
    b[0] = a[1] - a[0];
    b[n[d] - 1] = a[n[d] - 1] - a[n[d] - 2];
    for (int i = 1; i < n[d] - 1; ++i) b[i] = (a[i + 1] - a[i]) / 2.0 + (a[i] - a[i - 1]) / 2.0;
    h.g.n[d] = n[d];
    h.g.vmin[d] = a[0];
    h.g.vmax[d] = a[n[d] - 1];
    h.g.factor[d] = (double)(n[d] - 1) / (a[n[d] - 1] - a[0]);
    zero_point(a, h.g.zero[d], h.g.pos[d], h.g.neg[d]);
  }
  for (int d = 0; d < 3; ++d) h.packed.insert(h.packed.end(), h.ax[d].begin(), h.ax[d].end());
  for (int d = 0; d < 3; ++d) h.packed.insert(h.packed.end(), h.beta[d].begin(), h.beta[d].end());
  h.g.N = n[0] * n[1] * n[2];
  h.g.nxy = n[0] * n[1];
  h.g.inv_n0 = 1.0 / (double)n[0];
  h.g.inv_nxy = 1.0 / (double)(n[0] * n[1]);
  double b3min = h.beta[0][0] * h.beta[1][0] * h.beta[2][0];
  for (int k = 0; k < n[2]; ++k)
    for (int j = 0; j < n[1]; ++j)
      for (int i = 0; i < n[0]; ++i) {
        const double b3 = h.beta[0][i] * h.beta[1][j] * h.beta[2][k];
        if (b3 < b3min) b3min = b3;
      }
  h.b3min = b3min;
  h.uniform = true;
  for (int d = 0; d < 3; ++d) {
    const auto& a = h.ax[d];
    const double sp = (a[n[d] - 1] - a[0]) / (double)(n[d] - 1);
    for (int i = 0; i < n[d]; ++i)
      if (!(fabs(a[i] - (a[0] + i * sp)) <= 1e-9 * fabs(sp))) h.uniform = false;
  }
}

// Point the grid at its packed image at device address p.
static void place_grid(Grid& g, double* p) {
  for (int d = 0; d < 3; ++d) { g.ax[d] = p; p += g.n[d]; }
  for (int d = 0; d < 3; ++d) { g.beta[d] = p; p += g.n[d]; }
}

// Species constants (scaling.f90:163-201 calculate_cross_section) and the
// sample-count / collision parameters derived from them.
static void species_params(const dvm_config* cfg, double b3min, CdfParams& cp, CollParams& co) {
  const double m = cfg->mass, d = cfg->diam;
  const double m_red = m * m / (m + m);
  const double omega_ab = 0.5 * (cfg->omega + cfg->omega);
  const double sigma_vhs = 0.25 * (d + d) * (d + d);
  double sigma = sigma_vhs;
  if (cfg->csection_coeff == DVM_CS_BIRD)
    sigma = sigma_vhs * pow(m_red, 0.5 - omega_ab) / gamma_ref(2.5 - omega_ab);
  else if (cfg->csection_coeff == DVM_CS_VKI)
    sigma = (3.32335097045 * pow(m_red, 1.0 - omega_ab) * pow(cfg->t_viscosity, omega_ab)) /
            (gamma_ref(4.5 - omega_ab) * cfg->viscosity);
  cp.mass = m;
  cp.cdf_tolerance = cfg->cdf_tolerance;
  cp.dt = cfg->dt;
  cp.kn = cfg->kn;
  cp.sigma = sigma;
  cp.rms_sq = cfg->coln_rms * cfg->coln_rms;
  const double b3a = b3min * pow(m, 3.0 / 2.0);
  cp.beta3 = 0.5 * (b3a + b3a);
  cp.sample_mode = cfg->sample_mode;
  cp.samples_per_cell = cfg->samples_per_cell;
  cp.fmc = cfg->coll_solver == DVM_COLL_FMC;
  co.seed = cfg->seed;
  co.mass = m;
  co.vhs_exponent = 2.0 - 2.0 * omega_ab;
  co.hard_sphere = (co.vhs_exponent == 1.0);
  const double alpha_ab = 0.5 * (cfg->alpha + cfg->alpha);
  co.ralpha = 1.0 / alpha_ab;
  co.vss = fabs(co.ralpha - 1.0) > kDoubleTol;
  co.num_repl = cfg->num_repl;
  co.inv_repl = 1.0f / (float)cfg->num_repl;
}

// apply_fixed_in_out_flow (grid_ref_code boundary_conditions.f90:783-913):
// the wall Maxwellian (n, u, T; v = w = 0) times beta^3 on both ends, [2][N].
static void inflow_maxwellians(const dvm_config* cfg, const HostGrid& h, std::vector<double>& inf) {
  const int n[3] = {h.g.n[0], h.g.n[1], h.g.n[2]};
  const int N = h.g.N;
  const double m = cfg->mass;
  inf.assign((size_t)2 * N, 0.0);
  for (int s = 0; s < 2; ++s) {
    const double* wl = cfg->wall[s];
    const double dens = wl[DVM_W_DENS], temp = wl[DVM_W_TEMP], u = wl[DVM_W_U];
    const double mc = dens * sqrt(m * m * m / (kPiRef * kPiRef * kPiRef * temp * temp * temp));
    for (int k = 0, l = 0; k < n[2]; ++k)
      for (int j = 0; j < n[1]; ++j)
        for (int i = 0; i < n[0]; ++i, ++l) {
          const double beta3 = h.beta[0][i] * h.beta[1][j] * h.beta[2][k];
          inf[(size_t)s * N + l] = maxwell_host(h.ax[0][i], h.ax[1][j], h.ax[2][k], mc, m, u, 0.0, 0.0, temp) * beta3;
        }
  }
}

// Convection constants of a run (the sweep's CFL factors, boundary types and
// the zero-plane flags of boundary_conditions.f90:662-673, kept per grid: the
// reference's SAVE makes them sticky across grids).
static void conv_params(const dvm_config* cfg, const HostGrid& hg, ConvParams& p) {
  const int* n = hg.g.n;
  p.nx = cfg->nx;
  p.ny = cfg->ny;
  p.max_along = cfg->nx > cfg->ny ? cfg->nx : cfg->ny;
  p.fx = cfg->dt / cfg->dx;
  p.fy = (cfg->dim == DVM_DIM_2D) ? cfg->dt / cfg->dy : 0.0;
  p.acc = cfg->acc_coeff;
  p.bc[0] = cfg->bc_left;
  p.bc[1] = cfg->bc_right;
  p.order = cfg->conv_order == 2 ? 2 : 1;
  for (int s = 0; s < 4; ++s) { p.wall_side[s] = 0; p.couette_side[s] = 0; p.flux_out[s] = 1.0; }
  const int izero = hg.g.zero[0], jzero = hg.g.zero[1];
  p.izflag = (izero != 0 && izero != hg.g.neg[0] && izero != hg.g.pos[0] && izero != n[0] - 1);
  p.jzflag = (jzero != 0 && jzero != hg.g.neg[1] && jzero != hg.g.pos[1] && jzero != n[1] - 1);
}

// 2-D walls: which sides are diffuse / specular walls or Couette sides, the
// wall Maxwellian * beta^3 on each wall's outgoing half-space (wm [4][N]) and
// its flux (boundary_conditions.f90:785-828 and analogues), summed in the
// reference's loop order.
static void conv2d_walls(const dvm_config* cfg, const HostGrid& hg, ConvParams& p, std::vector<double>& wm) {
  const int* n = hg.g.n;
  const int N = hg.g.N;
  const double m = cfg->mass;
  p.wall_side[0] = cfg->bc_left == DVM_BC_WALL;
  p.wall_side[1] = cfg->bc_right == DVM_BC_WALL;
  p.wall_side[2] = 1;  // bottom and top are walls unconditionally in trunk 2-D
  p.wall_side[3] = 1;
  p.couette_side[0] = cfg->bc_left == DVM_BC_COUETTE;
  p.couette_side[1] = cfg->bc_right == DVM_BC_COUETTE;
  wm.assign((size_t)4 * N, 0.0);
  for (int s = 0; s < 4; ++s) {
    if (!p.wall_side[s]) continue;
    const double* wl = cfg->wall[s];
    const double kt = wl[DVM_W_TEMP], u = wl[DVM_W_U], v = wl[DVM_W_V], w = wl[DVM_W_W];
    const double mc = sqrt(m * m * m / (kPiRef * kPiRef * kPiRef * kt * kt * kt));
    double flux_out = 0.0;
    double* gm = wm.data() + (size_t)s * N;
    auto put = [&](int i, int j, int k) {
      const double beta3 = hg.beta[0][i] * hg.beta[1][j] * hg.beta[2][k];
      const double mv = maxwell_host(hg.ax[0][i], hg.ax[1][j], hg.ax[2][k], mc, m, u, v, w, kt);
      gm[i + n[0] * (j + n[1] * k)] = mv * beta3;
      const double vn = (s < 2) ? hg.ax[0][i] : hg.ax[1][j];
      flux_out = flux_out + fabs(vn) * mv * beta3;
    };
    if (s == 0) { for (int k = 0; k < n[2]; ++k) for (int j = 0; j < n[1]; ++j) for (int i = hg.g.pos[0]; i < n[0]; ++i) put(i, j, k); }
    if (s == 1) { for (int k = 0; k < n[2]; ++k) for (int j = 0; j < n[1]; ++j) for (int i = 0; i <= hg.g.neg[0]; ++i) put(i, j, k); }
    if (s == 2) { for (int k = 0; k < n[2]; ++k) for (int i = 0; i < n[0]; ++i) for (int j = hg.g.pos[1]; j < n[1]; ++j) put(i, j, k); }
    if (s == 3) { for (int k = 0; k < n[2]; ++k) for (int i = 0; i < n[0]; ++i) for (int j = 0; j <= hg.g.neg[1]; ++j) put(i, j, k); }
    p.flux_out[s] = flux_out;
  }
}

static int create_impl(const dvm_config* cfg, dvm_ctx* c) {
  c->cfg = *cfg;
  c->ncell = cfg->nx * cfg->ny;
  const int n[3] = {cfg->nvx, cfg->nvy, cfg->nvz};
  c->N = n[0] * n[1] * n[2];
  const int N = c->N;
  HIP_TRY(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  for (auto& row : c->ev)
    for (auto& e : row) HIP_TRY(c, hipEventCreate(&e));
  HostGrid hg;
  build_host_grid(cfg, hg);
  for (int d = 0; d < 3; ++d) { c->ax[d] = hg.ax[d]; c->beta[d] = hg.beta[d]; }
  c->g = hg.g;
  c->uniform_axes = hg.uniform;
  if (int rc = dalloc(c, &c->d_axes, hg.packed.size())) return rc;
  HIP_TRY(c, hipMemcpy(c->d_axes, hg.packed.data(), hg.packed.size() * sizeof(double), hipMemcpyHostToDevice));
  place_grid(c->g, c->d_axes);
  species_params(cfg, hg.b3min, c->cdfp, c->collp);
  const double m = cfg->mass;
  // ---- resident buffers
  const size_t CN = (size_t)c->ncell * N, CN1 = (size_t)c->ncell * (N + 1);
  int rc;
  if ((rc = dalloc(c, &c->f, CN)) || (rc = dalloc(c, &c->fd, CN)) || (rc = dalloc(c, &c->eqv, CN)) ||
      (rc = dalloc(c, &c->cdf, 2 * CN1)) || (rc = dalloc(c, &c->tabs, (size_t)c->ncell * 2 * kNBucket)) ||
      (rc = dalloc(c, &c->eqpar, (size_t)c->ncell * 5)) ||
      (rc = dalloc(c, &c->props, (size_t)c->ncell * DVM_NPROP)) || (rc = dalloc(c, &c->factors, (size_t)c->ncell * 3)) ||
      (rc = dalloc(c, &c->counts, (size_t)c->ncell * 4)) || (rc = dalloc(c, &c->err, 4)) ||
      (rc = dalloc(c, &c->sgn, 2 * CN1)))
    return rc;
  HIP_TRY(c, hipMemset(c->err, 0, 4 * sizeof(int)));
  HIP_TRY(c, hipHostMalloc((void**)&c->err_host, 4 * sizeof(int), hipHostMallocDefault));
  c->err_host[0] = 0;
  HIP_TRY(c, hipEventCreateWithFlags(&c->err_ev, hipEventDisableTiming));
  HIP_TRY(c, hipMemset(c->counts, 0, (size_t)c->ncell * 4 * sizeof(int)));
  // the full Monte Carlo kernel samples with f_eq = 0: its partner signs are
  // dsgn(f - 0) = dsgn(f)
  HIP_TRY(c, hipMemset(c->eqv, 0, CN * sizeof(double)));
  // collision accumulator placement: the cell's slice in LDS when it fits
  const bool f32 = prec_f32(cfg->precision);
  // the fast FP32 sampling (collide_fast): equally spaced axes, the FP64
  // remap, isotropic scattering; its sign words share the LDS with the
  // window
  const bool fast_ok = cfg->precision == DVM_PREC_FP32 && c->uniform_axes && !c->collp.vss;
  c->lds_collide = window_dims(n, c->wd.w, fast_ok ? fast_window_cap(N) : kWindowCap);
  // FP64 with an LDS window (the grid exceeds LDS): 768 lanes, where the FP64
  // sampling code's registers spill less than at the 128-register cap of 1024
  // lanes (1-D shock, 31^3: collide 0.933 ms against 0.966; the cavity's
  // whole-grid FP64 slice and the FP32 path run faster with 1024)
  if (!f32 && !c->lds_collide) c->coll_bs = 768;
  if (const char* e = getenv("DVM_COLL_BS")) {
    const int v = atoi(e);
    if (v == 256 || v == 512 || v == 768 || v == 1024) c->coll_bs = v;
  }
  if (const char* e = getenv("DVM_PICK_K")) {
    const int v = atoi(e);
    if (v == 1 || v == 2 || v == 4) c->pick_k = v;
  }
  c->coll_lds = (size_t)(kAxisSlots + c->wd.w[0] * c->wd.w[1] * c->wd.w[2]) * sizeof(double);
  if (f32) {
    if (int rc2 = dalloc(c, &c->dacc, CN)) return rc2;
  }
  if (fast_ok) {
    c->coll_lds += (size_t)sign_lds_bytes(N);  // sign words after the slice or window
    c->fast = 1;
    if (c->lds_collide)
      if (int rc2 = dalloc(c, &c->cmat, (size_t)c->ncell * kCmat)) return rc2;
  }
  const void* kern = collide_kernel(c);
  HIP_TRY(c, hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)c->coll_lds));
  HIP_TRY(c, hipFuncSetAttribute(collide_stream_kernel(c), hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)c->coll_lds));
  {
    int dev = 0;
    c->cus = 256;
    if (hipGetDevice(&dev) == hipSuccess)
      (void)hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, dev);
  }
  // the device's scratch address, for dvm_l1_noalloc (looked up here, at set-up)
  c->scratch = scratch_address(c->msg, sizeof c->msg);
  if (!c->scratch) return DVM_E_HIP;
  // Few cells: split each cell's samples over several workgroups, at least
  // two rounds of workgroups over the chip (and at least 4096 samples per part
  // when the count is known up front). Among the splits up to four times that,
  // the one whose last round fills the chip best (fewest parts among equals):
  // 200 cells on 256 CUs take 5 parts (1000 workgroups in 1024 slots), not 3
  // (600 in 768).
  {
    int per_cu = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, c->coll_bs, c->coll_lds) != hipSuccess ||
        per_cu < 1)
      per_cu = 1;
    const long long slots = (long long)per_cu * c->cus;
    const long long want = (2LL * c->cus + c->ncell - 1) / c->ncell;
    long long lo = want, hi = want > 1 ? 4 * want : 1;
    if (cfg->sample_mode == DVM_SAMPLES_EXPLICIT) {
      const long long by_samples = cfg->samples_per_cell / 4096;
      lo = lo < by_samples ? lo : by_samples;
      hi = hi < by_samples ? hi : by_samples;
    }
    lo = lo < 1 ? 1 : (lo > 1024 ? 1024 : lo);
    hi = hi < lo ? lo : (hi > 1024 ? 1024 : hi);
    long long best = lo;
    double best_fill = -1.0;
    for (long long sp = lo; sp <= hi; ++sp) {
      const long long wg = (long long)c->ncell * sp, rounds = (wg + slots - 1) / slots;
      const double fill = (double)wg / (double)(rounds * slots);
      if (fill > best_fill + 1e-9) { best = sp; best_fill = fill; }
    }
    c->split = (int)best;
    if (const char* e = getenv("DVM_COLL_SPLIT")) {
      const int v = atoi(e);
      if (v >= 1 && v <= 1024) c->split = v;
    }
  }
  // The fast FP32 collision of whole cells on persistent workgroups, one per
  // CU (k_collide_fast_cells): cells each one part, the grid in LDS, 1024
  // lanes and one pick per lane. Its ticket counters start at zero here and
  // the kernel leaves them at zero.
  if (c->fast && c->lds_collide && c->split == 1 && c->coll_bs == kCollBS && c->pick_k == 1) {
    // the cell's bucket tables join the sign words in LDS when they fit
    const size_t tab_lds = ((c->coll_lds + 7) & ~(size_t)7) + 2 * kNBucket * sizeof(int2);
    // DVM_FAST_LDS_TABLES=0 keeps them in global memory (a test compares
    // the two)
    const char* lt = getenv("DVM_FAST_LDS_TABLES");
    c->fast_tab_lds = tab_lds <= kLdsBytes - kStaticLds && !(lt && atoi(lt) == 0);
    c->fast_lds = c->fast_tab_lds ? tab_lds : c->coll_lds;
    const void* pk = c->fast_tab_lds ? (const void*)k_collide_fast_cells<kCollBS, true>
                                     : (const void*)k_collide_fast_cells<kCollBS, false>;
    HIP_TRY(c, hipFuncSetAttribute(pk, hipFuncAttributeMaxDynamicSharedMemorySize, (int)c->fast_lds));
    int per_cu = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pk, kCollBS, c->fast_lds) != hipSuccess || per_cu < 1)
      per_cu = 1;
    const long long wgs = (long long)per_cu * c->cus;
    c->fast_wgs = (int)(wgs < c->ncell ? wgs : c->ncell);
    if (int rc2 = dalloc(c, &c->tick, 2)) return rc2;
    // zeroed in the context's stream order (set-up; never inside a captured step)
    HIP_TRY(c, hipMemsetAsync(c->tick, 0, 2 * sizeof(unsigned), c->stream));
  }
  // ---- convection constants
  ConvParams& p = c->convp;
  conv_params(cfg, hg, p);
  if (cfg->dim == DVM_DIM_2D && cfg->enable_convection) {
    std::vector<double> wm;
    conv2d_walls(cfg, hg, p, wm);
    if ((rc = dalloc(c, &c->wallM, (size_t)4 * N)) || (rc = dalloc(c, &c->ratio, (size_t)4 * p.max_along))) return rc;
    HIP_TRY(c, hipMemcpy(c->wallM, wm.data(), wm.size() * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemset(c->ratio, 0, (size_t)4 * p.max_along * sizeof(double)));
    if (cfg->acc_coeff < 1.0) {
      if ((rc = dalloc(c, &c->spec, (size_t)4 * p.max_along * N))) return rc;
    }
  }
  if (cfg->dim == DVM_DIM_1D && cfg->enable_convection) {
    std::vector<double> inf;
    inflow_maxwellians(cfg, hg, inf);
    if ((rc = dalloc(c, &c->inflow, (size_t)2 * N))) return rc;
    HIP_TRY(c, hipMemcpy(c->inflow, inf.data(), inf.size() * sizeof(double), hipMemcpyHostToDevice));
    // 1-D walls (apply_diffuse_reflection, grid_ref_code
    // boundary_conditions.f90:566-781): M_wall beta^3 on the outgoing half
    // and its sum, in the reference's loop order
    std::vector<double> wm((size_t)2 * N, 0.0);
    bool any_wall = false;
    for (int s = 0; s < 2; ++s) {
      if (p.bc[s] != DVM_BC_WALL) continue;
      any_wall = true;
      const double* wl = cfg->wall[s];
      const double kt = wl[DVM_W_TEMP], u = wl[DVM_W_U], v = wl[DVM_W_V], w = wl[DVM_W_W];
      const double mc = sqrt(m * m * m / (kPiRef * kPiRef * kPiRef * kt * kt * kt));
      const int o0 = s == 0 ? c->g.pos[0] : 0, o1 = s == 0 ? n[0] - 1 : c->g.neg[0];
      double computed = 0.0;
      for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
          for (int i = o0; i <= o1; ++i) {
            const double beta3 = c->beta[0][i] * c->beta[1][j] * c->beta[2][k];
            const double mv = maxwell_host(c->ax[0][i], c->ax[1][j], c->ax[2][k], mc, m, u, v, w, kt);
            wm[(size_t)s * N + i + n[0] * (j + n[1] * k)] = mv * beta3;
            computed = computed + mv * beta3;
          }
      p.flux_out[s] = computed;
    }
    if (any_wall) {
      if ((rc = dalloc(c, &c->wallM, (size_t)2 * N)) || (rc = dalloc(c, &c->ratio, 2)) ||
          (rc = dalloc(c, &c->spec, (size_t)2 * N)))
        return rc;
      HIP_TRY(c, hipMemcpy(c->wallM, wm.data(), wm.size() * sizeof(double), hipMemcpyHostToDevice));
      HIP_TRY(c, hipMemset(c->ratio, 0, 2 * sizeof(double)));
      HIP_TRY(c, hipMemset(c->spec, 0, (size_t)2 * N * sizeof(double)));
    }
  }
  return DVM_OK;
}

extern "C" int dvm_ctx_create(const dvm_config* cfg, dvm_ctx** out) {
  if (!out) return fail(nullptr, DVM_E_ARG, "null output pointer");
  *out = nullptr;
  if (int rc = validate(cfg)) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(nullptr, DVM_E_HIP, "no HIP device is visible");
  dvm_ctx* c = new dvm_ctx();
  const int rc = create_impl(cfg, c);
  if (rc) {
    snprintf(g_create_msg, sizeof g_create_msg, "%s", c->msg);
    dvm_ctx_destroy(c);
    return rc;
  }
  *out = c;
  return DVM_OK;
}

// Message and status of a device error word (the kErr* bits); 0 if clear.
static int device_error(int flags, char* text, size_t n) {
  if (!flags) return DVM_OK;
  snprintf(text, n, "device error flags 0x%x:%s%s%s%s", flags,
           (flags & kErrRemap) ? " a post-collision remap found no exterior point (the reference stops, "
                                 "remapping.f90:348-354);" : "",
           (flags & kErrSpecRemap) ? " a specular ghost remap found no exterior point;" : "",
           (flags & kErrSpecNode) ? " a specular ghost stencil point left the grid;" : "",
           (flags & kErrConserve) ? " the FP32 conservation system of a cell was not positive definite "
                                    "(the change was kept uncorrected);" : "");
  return (flags & (kErrRemap | kErrSpecRemap | kErrSpecNode)) ? DVM_E_REMAP : DVM_E_STATE;
}

static int check_err(dvm_ctx* c) {
  int h[4] = {0, 0, 0, 0};
  HIP_TRY(c, hipMemcpyAsync(h, c->err, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->err_pending = false;  // this read covers every step queued so far
  if (h[0]) {
    HIP_TRY(c, hipMemset(c->err, 0, sizeof h));
    char text[400];
    const int rc = device_error(h[0], text, sizeof text);
    return fail(c, rc, "%s", text);
  }
  return DVM_OK;
}

extern "C" int dvm_upload_f(dvm_ctx* c, int32_t species, const double* f) {
  if (!c || !f) return fail(c, DVM_E_ARG, "null argument");
  if (species != 0) return fail(c, DVM_E_ARG, "only species 0 is supported");
  const size_t bytes = (size_t)c->ncell * c->N * sizeof(double);
  HIP_TRY(c, hipMemcpyAsync(c->f, f, bytes, hipMemcpyHostToDevice, c->stream));
  c->cmat_valid = false;
  HIP_TRY(c, hipMemsetAsync(c->eqpar, 0, (size_t)c->ncell * 5 * sizeof(double), c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->f_valid = true;
  c->props_valid = false;
  return DVM_OK;
}

extern "C" int dvm_download_f(dvm_ctx* c, int32_t species, double* f) {
  if (!c || !f) return fail(c, DVM_E_ARG, "null argument");
  if (species != 0) return fail(c, DVM_E_ARG, "only species 0 is supported");
  if (!c->f_valid) return fail(c, DVM_E_STATE, "no distribution uploaded");
  const size_t bytes = (size_t)c->ncell * c->N * sizeof(double);
  HIP_TRY(c, hipMemcpyAsync(f, c->f, bytes, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return DVM_OK;
}

// ---- phase timing: a ring of event slots, one slot per entry-point call
static int timing_drain_slot(dvm_ctx* c, int k) {
  if (!c->slot_used[k]) return DVM_OK;
  HIP_TRY(c, hipEventSynchronize(c->ev[k][DVM_NPHASE]));
  for (int q = 0; q < DVM_NPHASE; ++q) {
    float ms = 0.0f;
    HIP_TRY(c, hipEventElapsedTime(&ms, c->ev[k][q], c->ev[k][q + 1]));
    c->phase_ms[q] += ms;
  }
  c->slot_used[k] = false;
  c->timed_steps++;
  return DVM_OK;
}

static int begin_slot(dvm_ctx* c) { return timing_drain_slot(c, c->slot); }

static int mark(dvm_ctx* c, int q) {
  HIP_TRY(c, hipEventRecord(c->ev[c->slot][q], c->stream));
  return DVM_OK;
}

static void end_slot(dvm_ctx* c) {
  c->slot_used[c->slot] = true;
  c->slot = (c->slot + 1) % dvm_ctx::kSlots;
}

// ---- phases
// Collision phases run for the variance-reduction and full Monte Carlo solvers.
static bool collides(const dvm_ctx* c) {
  return c->cfg.coll_solver == DVM_COLL_VR || c->cfg.coll_solver == DVM_COLL_FMC;
}

static int moments_async(dvm_ctx* c) {
  if (c->cmat) {
    // FP32 fast path: the moments and the next collision's conservation factor
    // two cells per workgroup when that still gives two workgroups per CU
    if (c->ncell >= 4 * c->cus)
      hipLaunchKernelGGL((k_moments_cmat<2, kMomBS>), dim3((c->ncell + 1) / 2), dim3(kMomBS), 0, c->stream, c->g,
                         c->f, c->props, c->cmat, c->cfg.mass, c->ncell);
    else if (c->ncell >= 2 * c->cus)
      hipLaunchKernelGGL((k_moments_cmat<1, kMomBS>), dim3(c->ncell), dim3(kMomBS), 0, c->stream, c->g, c->f,
                         c->props, c->cmat, c->cfg.mass, c->ncell);
    else
      hipLaunchKernelGGL((k_moments_cmat<1, 1024>), dim3(c->ncell), dim3(1024), 0, c->stream, c->g, c->f, c->props,
                         c->cmat, c->cfg.mass, c->ncell);
    HIP_TRY(c, hipGetLastError());
    c->props_valid = true;
    c->cmat_valid = true;
    return DVM_OK;
  }
  // kMomCells cells per workgroup when that still gives two workgroups per CU;
  // one cell per workgroup otherwise, on 1024 lanes when even that leaves the
  // chip short of two workgroups per CU (the 1-D shock, a homogeneous cell)
  if (c->ncell >= 2 * kMomCells * c->cus)
    hipLaunchKernelGGL((k_moments<kMomCells, kMomBS>), dim3((c->ncell + kMomCells - 1) / kMomCells), dim3(kMomBS),
                       0, c->stream, c->g, c->f, c->props, c->cfg.mass, c->ncell);
  else if (c->ncell >= 2 * c->cus)
    hipLaunchKernelGGL((k_moments<1, kMomBS>), dim3(c->ncell), dim3(kMomBS), 0, c->stream, c->g, c->f, c->props,
                       c->cfg.mass, c->ncell);
  else
    hipLaunchKernelGGL((k_moments<1, 1024>), dim3(c->ncell), dim3(1024), 0, c->stream, c->g, c->f, c->props,
                       c->cfg.mass, c->ncell);
  HIP_TRY(c, hipGetLastError());
  c->props_valid = true;
  return DVM_OK;
}

// refresh: the pass leaves f_eq stored for a reader of it (the trace and
// uniform-stream kernels sample with the general code, which reads f_eq).
static int cdf_async(dvm_ctx* c, bool refresh = false) {
  if (!collides(c)) return DVM_OK;
  // few cells (fewer than four workgroups per CU): 1024 lanes per cell
  if (c->ncell >= 4 * c->cus) {
    const int cap = cdf_sum_cap(c->g.N, kCdfV);
    // the fast FP32 collision reads the sign bits, never f_eq: the pass then
    // stores no f_eq (and computes it again on a cache hit)
    const int eq_mode = !c->fast ? 1 : ((refresh || c->d_su) ? 2 : 0);
    hipLaunchKernelGGL((k_cdf<kCdfCtxBS, kCdfV, true>), dim3(c->ncell), dim3(kCdfCtxBS), cap * sizeof(float), c->stream,
                       c->g, c->cdfp, c->f, c->props, c->eqv, c->cdf, c->tabs, c->eqpar, c->counts, c->factors, cap,
                       c->sgn, eq_mode);
  } else {
    const int cap = cdf_sum_cap(c->g.N, kCdfWideV);
    hipLaunchKernelGGL((k_cdf<kCdfWideBS, kCdfWideV, false>), dim3(c->ncell), dim3(kCdfWideBS), cap * sizeof(float),
                       c->stream, c->g, c->cdfp, c->f, c->props, c->eqv, c->cdf, c->tabs, c->eqpar, c->counts,
                       c->factors, cap, c->sgn, 1);
  }
  HIP_TRY(c, hipGetLastError());
  return DVM_OK;
}

// fd = f (FP64) or dacc = 0 (FP32) before a collision that spills from its
// LDS window to global atomics.
static int prep_async(dvm_ctx* c) {
  if (!collides(c) || (c->lds_collide && c->split == 1)) return DVM_OK;
  if (prec_f32(c->cfg.precision))
    hipLaunchKernelGGL(k_fill, dim3(2048), dim3(256), 0, c->stream, c->dacc, (long long)c->ncell * c->N, 0.0);
  else
    HIP_TRY(c, hipMemcpyAsync(c->fd, c->f, (size_t)c->ncell * c->N * sizeof(double), hipMemcpyDeviceToDevice,
                              c->stream));
  HIP_TRY(c, hipGetLastError());
  return DVM_OK;
}

// The FP32 path corrects conservation inside the collide kernel when the
// cell's slice fits LDS and one workgroup takes the cell, else in k_conserve.
static bool separate_conserve(const dvm_ctx* c) {
  return prec_f32(c->cfg.precision) && (!c->lds_collide || c->split > 1);
}

static int collide_async(dvm_ctx* c, int ntime) {
  if (!collides(c)) return DVM_OK;
  const bool stream = c->d_su != nullptr;
  int per = stream_per_sample(c);
  const double* su = c->d_su;
  const long long* sbase = c->d_sbase;
  if (stream) {
    // offsets of every cell's samples in the list; the step must fit in what
    // is left of it (a host sync: this is the parity mode)
    hipLaunchKernelGGL(k_stream_bases, dim3(1), dim3(64), 0, c->stream, c->counts, c->ncell, c->su_cursor, per,
                       c->d_sbase, c->d_stotal);
    long long need = 0;
    HIP_TRY(c, hipMemcpyAsync(&need, c->d_stotal, sizeof need, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (need > c->su_n - c->su_cursor)
      return fail(c, DVM_E_ARG, "uniform stream: the step reads %lld values, %lld are left of %lld", need,
                  c->su_n - c->su_cursor, c->su_n);
    c->su_cursor += need;
  }
  int fast = (c->fast && !stream) ? 1 : 0;
  // the conservation factor of the moments pass when it is that of this f
  // (else the collision kernel sums it itself)
  const double* cm = c->cmat_valid ? c->cmat : nullptr;
  void* args[] = {&c->g,    &c->collp, &ntime,  &c->split,  &c->wd,      &c->props, &c->f, &c->fd, &c->dacc,
                  &c->eqv,  &c->cdf,   &c->tabs, &c->counts, &c->factors, &c->err,   &su,   &sbase, &per,
                  &c->sgn,  &cm,       &fast};
  if (stream) {
    HIP_TRY(c, hipLaunchKernel(collide_stream_kernel(c), dim3((unsigned)c->ncell * c->split), dim3(kCollBS), args,
                               c->coll_lds, c->stream));
  } else if (fast && cm && c->fast_wgs > 0) {
    // whole cells on persistent workgroups (the factor is that of this f)
    if (c->fast_tab_lds)
      hipLaunchKernelGGL((k_collide_fast_cells<kCollBS, true>), dim3((unsigned)c->fast_wgs), dim3(kCollBS),
                         c->fast_lds, c->stream, c->g, c->collp, ntime, c->ncell, c->f, c->fd, c->eqv, c->cdf,
                         c->tabs, c->counts, c->factors, c->err, c->sgn, cm, c->tick);
    else
      hipLaunchKernelGGL((k_collide_fast_cells<kCollBS, false>), dim3((unsigned)c->fast_wgs), dim3(kCollBS),
                         c->fast_lds, c->stream, c->g, c->collp, ntime, c->ncell, c->f, c->fd, c->eqv, c->cdf,
                         c->tabs, c->counts, c->factors, c->err, c->sgn, cm, c->tick);
    HIP_TRY(c, hipGetLastError());
  } else {
    HIP_TRY(c, hipLaunchKernel(collide_kernel(c), dim3((unsigned)c->ncell * c->split), dim3(c->coll_bs), args,
                               c->coll_lds, c->stream));
  }
  if (!separate_conserve(c)) {
    double* t = c->f; c->f = c->fd; c->fd = t;  // phi := delta_phi
  }
  c->cmat_valid = false;
  return DVM_OK;
}

// FP32 path beyond LDS: fd = f + dacc with the FP64 conservation correction,
// then phi := delta_phi.
static int conserve_async(dvm_ctx* c) {
  if (!collides(c) || !separate_conserve(c)) return DVM_OK;
  hipLaunchKernelGGL(k_conserve<256>, dim3(c->ncell), dim3(256), 0, c->stream, c->g, c->f, c->dacc, c->fd, c->err);
  HIP_TRY(c, hipGetLastError());
  double* t = c->f; c->f = c->fd; c->fd = t;
  return DVM_OK;
}

static bool convects(const dvm_ctx* c) { return c->cfg.enable_convection && c->cfg.dim != DVM_DIM_0D; }

static int walls_async(dvm_ctx* c, int ntime) {
  if (!convects(c)) return DVM_OK;
  ConvParams& p = c->convp;
  const int N = c->N;
  if (c->cfg.dim == DVM_DIM_1D) {
    if (!c->wallM) return DVM_OK;  // no 1-D wall
    hipLaunchKernelGGL(k_wall1d_ratio, dim3(2), dim3(kMomBS), 0, c->stream, c->g, p, c->f, c->ratio);
    if (p.acc < 1.0) {
      hipLaunchKernelGGL(k_fill, dim3(64), dim3(256), 0, c->stream, c->spec, (long long)2 * N, 0.0);
      hipLaunchKernelGGL(k_spec1d, dim3((N + 255) / 256, 2), dim3(256), 0, c->stream, c->g, p, c->f, c->spec,
                         c->err);
    }
    HIP_TRY(c, hipGetLastError());
    return DVM_OK;
  }
  if (p.acc > 0.0)
    hipLaunchKernelGGL(k_wall_ratio, dim3(p.max_along, 4), dim3(kWallBS), 0, c->stream, c->g, p, c->f, c->ratio);
  if (p.acc < 1.0) {
    const long long ns = (long long)4 * p.max_along * N;
    hipLaunchKernelGGL(k_fill, dim3(1024), dim3(256), 0, c->stream, c->spec, ns, 0.0);
    hipLaunchKernelGGL(k_spec_ghost, dim3((N + 255) / 256, p.max_along, 4), dim3(256), 0, c->stream, c->g, p, c->f,
                       c->spec, c->err);
  }
  if (c->heat_every > 0 && ntime % c->heat_every == 0) {
    WallVel wv;
    for (int s = 0; s < 4; ++s)
      for (int d = 0; d < 3; ++d) wv.u[s][d] = c->cfg.wall[s][DVM_W_U + d];
    hipLaunchKernelGGL(k_wall_heat, dim3(p.max_along, 4), dim3(kMomBS), 0, c->stream, c->g, p, wv, c->cfg.mass,
                       c->f, c->props, c->wallM, c->ratio, c->spec, c->d_heat);
    c->heat_ntime = ntime;
  }
  HIP_TRY(c, hipGetLastError());
  return DVM_OK;
}

static int sweep_async(dvm_ctx* c) {
  if (!convects(c)) return DVM_OK;
  // node chunks padded to a multiple of 8 (one XCD per chunk index, see k_convect2d)
  const unsigned chunks = (unsigned)(((c->N + kConvBS - 1) / kConvBS + 7) / 8 * 8);
  const unsigned ys = (unsigned)(c->ncell < 65535 ? c->ncell : 65535);
  if (c->cfg.dim == DVM_DIM_2D)
    launch_convect2d(c->g, c->convp, c->f, c->fd, c->wallM, c->ratio, c->spec, c->stream);
  else
    hipLaunchKernelGGL(k_convect1d, dim3(chunks, ys), dim3(kConvBS), 0, c->stream, c->g, c->convp, c->f, c->fd,
                       c->inflow, c->wallM, c->ratio, c->spec);
  HIP_TRY(c, hipGetLastError());
  double* t = c->f; c->f = c->fd; c->fd = t;  // phi := phi_conv
  c->cmat_valid = false;
  return DVM_OK;
}

// Runs the phases [first, last] of the step (0 cdf, 1 prep, 2 collide,
// 3 conserve, 4 walls, 5 sweep, 6 moments) with timing marks around each.
static int run_phases(dvm_ctx* c, int ntime, int first, int last) {
  if (int rc = begin_slot(c)) return rc;
  int rc = DVM_OK;
  for (int q = 0; q < DVM_NPHASE && rc == DVM_OK; ++q) {
    if ((rc = mark(c, q))) break;
    if (q < first || q > last) continue;
    switch (q) {
      case 0: rc = cdf_async(c); break;
      case 1: rc = prep_async(c); break;
      case 2: rc = collide_async(c, ntime); break;
      case 3: rc = conserve_async(c); break;
      case 4: rc = walls_async(c, ntime); break;
      case 5: rc = sweep_async(c); break;
      case 6: rc = moments_async(c); break;
    }
  }
  if (rc == DVM_OK) rc = mark(c, DVM_NPHASE);
  if (rc == DVM_OK) end_slot(c);
  return rc;
}

extern "C" int dvm_moments(dvm_ctx* c, double* props_out) {
  if (!c) return fail(c, DVM_E_ARG, "null context");
  if (!c->f_valid) return fail(c, DVM_E_STATE, "moments before upload");
  if (int rc = run_phases(c, 0, 6, 6)) return rc;
  if (props_out) {
    HIP_TRY(c, hipMemcpyAsync(props_out, c->props, (size_t)c->ncell * DVM_NPROP * sizeof(double),
                              hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
  }
  return DVM_OK;
}

extern "C" int dvm_moments_ext(dvm_ctx* c, double* ext) {
  if (!c || !ext) return fail(c, DVM_E_ARG, "null argument");
  if (!c->f_valid) return fail(c, DVM_E_STATE, "moments before upload");
  if (int rc = run_phases(c, 0, 6, 6)) return rc;
  double* d_ext = nullptr;
  if (int rc = dalloc(c, &d_ext, (size_t)c->ncell * DVM_NEXT)) return rc;
  hipLaunchKernelGGL(k_moments_ext, dim3(c->ncell), dim3(kMomBS), 0, c->stream, c->g, c->f, c->props, c->cfg.mass,
                     d_ext);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(ext, d_ext, (size_t)c->ncell * DVM_NEXT * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d_ext);
  if (e != hipSuccess) return fail(c, DVM_E_HIP, "dvm_moments_ext: %s", hipGetErrorString(e));
  return DVM_OK;
}

extern "C" int dvm_l1(dvm_ctx* c, double* l1) {
  if (!c || !l1) return fail(c, DVM_E_ARG, "null argument");
  if (!c->f_valid) return fail(c, DVM_E_STATE, "L1 before upload");
  if (int rc = run_phases(c, 0, 6, 6)) return rc;
  double* d_l1 = nullptr;
  if (int rc = dalloc(c, &d_l1, (size_t)c->ncell)) return rc;
  hipLaunchKernelGGL(k_l1, dim3(c->ncell), dim3(kMomBS), 0, c->stream, c->g, c->f, c->props, c->cfg.mass, d_l1);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(l1, d_l1, (size_t)c->ncell * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d_l1);
  if (e != hipSuccess) return fail(c, DVM_E_HIP, "dvm_l1: %s", hipGetErrorString(e));
  return DVM_OK;
}

// dvm_l1 without any device allocation: the per-cell results are staged in
// the device's file-scope scratch array (every entry the kernel reads it
// writes first, so nothing is cleared). Calls that overlap on one GPU share
// that array and race.
extern "C" int dvm_l1_noalloc(dvm_ctx* c, double* l1) {
  if (!c || !l1) return fail(c, DVM_E_ARG, "null argument");
  if (!c->f_valid) return fail(c, DVM_E_STATE, "L1 before upload");
  if ((long long)c->ncell > kScratchDoubles)
    return fail(c, DVM_E_ARG, "dvm_l1_noalloc: %d cells exceed its scratch of %lld cells (use dvm_l1)", c->ncell,
                kScratchDoubles);
  if (!c->scratch) return fail(c, DVM_E_STATE, "dvm_l1_noalloc: no scratch address for this device");
  if (int rc = run_phases(c, 0, 6, 6)) return rc;
  hipLaunchKernelGGL(k_l1, dim3(c->ncell), dim3(kMomBS), 0, c->stream, c->g, c->f, c->props, c->cfg.mass,
                     c->scratch);
  HIP_TRY(c, hipGetLastError());
  HIP_TRY(c, hipMemcpyAsync(l1, c->scratch, (size_t)c->ncell * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return DVM_OK;
}

extern "C" int dvm_collide(dvm_ctx* c, int32_t ntime, int64_t* samples_out) {
  if (!c) return fail(c, DVM_E_ARG, "null context");
  if (!c->f_valid || !c->props_valid) return fail(c, DVM_E_STATE, "collide needs f and its moments");
  if (ntime < 0) return fail(c, DVM_E_ARG, "ntime must be >= 0");
  if (int rc = run_phases(c, ntime, 0, 3)) return rc;
  if (samples_out) {
    std::vector<int> h((size_t)c->ncell * 4, 0);
    if (collides(c))
      HIP_TRY(c, hipMemcpyAsync(h.data(), c->counts, h.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (int q = 0; q < c->ncell; ++q) samples_out[q] = h[(size_t)q * 4 + 3];
  }
  return check_err(c);
}

extern "C" int dvm_convect(dvm_ctx* c, int32_t species, int32_t ntime) {
  if (!c) return fail(c, DVM_E_ARG, "null context");
  if (species != 0) return fail(c, DVM_E_ARG, "only species 0 is supported");
  if (!c->f_valid) return fail(c, DVM_E_STATE, "convect before upload");
  if (int rc = run_phases(c, ntime, 4, 5)) return rc;
  return check_err(c);
}

// dvm_step does not wait for the device: the error flags that a step raises
// are copied to pinned host memory behind it and reported by the first later
// dvm_step that finds the copy complete (or by any call that synchronises,
// which checks the flags itself), so a failing step surfaces within a step or
// two, as the reference's stop would end the run.
extern "C" int dvm_step(dvm_ctx* c, int32_t ntime) {
  if (!c) return fail(c, DVM_E_ARG, "null context");
  if (!c->f_valid || !c->props_valid) return fail(c, DVM_E_STATE, "step needs f and its moments");
  if (ntime < 0) return fail(c, DVM_E_ARG, "ntime must be >= 0");
  if (c->err_pending && hipEventQuery(c->err_ev) == hipSuccess) {
    c->err_pending = false;
    if (c->err_host[0]) return check_err(c);
  }
  if (int rc = run_phases(c, ntime, 0, 6)) return rc;
  if (!c->err_pending) {
    HIP_TRY(c, hipMemcpyAsync(c->err_host, c->err, 4 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipEventRecord(c->err_ev, c->stream));
    c->err_pending = true;
  }
  return DVM_OK;
}

extern "C" int dvm_timing_reset(dvm_ctx* c) {
  if (!c) return DVM_E_ARG;
  for (int k = 0; k < dvm_ctx::kSlots; ++k)
    if (int rc = timing_drain_slot(c, k)) return rc;
  for (double& v : c->phase_ms) v = 0.0;
  c->timed_steps = 0;
  return check_err(c);
}

extern "C" int dvm_timing_read(dvm_ctx* c, double* phase_ms, int64_t* calls) {
  if (!c) return DVM_E_ARG;
  for (int k = 0; k < dvm_ctx::kSlots; ++k)
    if (int rc = timing_drain_slot(c, k)) return rc;
  if (phase_ms)
    for (int q = 0; q < DVM_NPHASE; ++q) phase_ms[q] = c->phase_ms[q];
  if (calls) *calls = c->timed_steps;
  return check_err(c);
}

extern "C" int dvm_remap(dvm_ctx* c, int64_t n, const double* uvw, double f_off, int32_t* stencil, double* mass) {
  if (!c || (n > 0 && (!uvw || !stencil || !mass))) return fail(c, DVM_E_ARG, "null argument");
  if (n <= 0) return DVM_OK;
  double *d_uvw = nullptr, *d_m = nullptr;
  int* d_st = nullptr;
  int rc;
  if ((rc = dalloc(c, &d_uvw, (size_t)3 * n)) || (rc = dalloc(c, &d_m, (size_t)7 * n)) ||
      (rc = dalloc(c, &d_st, (size_t)7 * n))) {
    if (d_uvw) (void)hipFree(d_uvw);
    if (d_m) (void)hipFree(d_m);
    return rc;
  }
  hipError_t e = hipMemcpyAsync(d_uvw, uvw, (size_t)3 * n * sizeof(double), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_remap_batch, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->g, (long long)n,
                       d_uvw, f_off, d_st, d_m);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(stencil, d_st, (size_t)7 * n * sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(mass, d_m, (size_t)7 * n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d_uvw);
  (void)hipFree(d_m);
  (void)hipFree(d_st);
  if (e != hipSuccess) return fail(c, DVM_E_HIP, "dvm_remap: %s", hipGetErrorString(e));
  return DVM_OK;
}

extern "C" int dvm_set_wall_heat_flux_freq(dvm_ctx* c, int32_t every) {
  if (!c) return fail(c, DVM_E_ARG, "null context");
  if (every < 0) return fail(c, DVM_E_ARG, "every must be >= 0");
  if (every > 0 && c->cfg.dim != DVM_DIM_2D) return fail(c, DVM_E_ARG, "wall heat fluxes need a 2-D grid");
  if (every > 0 && !c->d_heat) {
    const size_t n = (size_t)4 * c->convp.max_along;
    if (int rc = dalloc(c, &c->d_heat, n)) return rc;
    HIP_TRY(c, hipMemsetAsync(c->d_heat, 0, n * sizeof(double), c->stream));
  }
  c->heat_every = every;
  return DVM_OK;
}

extern "C" int dvm_wall_heat_flux(dvm_ctx* c, double* q, int32_t* ntime_out) {
  if (!c || !q) return fail(c, DVM_E_ARG, "null argument");
  const size_t n = (size_t)4 * (c->cfg.nx > c->cfg.ny ? c->cfg.nx : c->cfg.ny);
  if (!c->d_heat || c->heat_ntime < 0) {
    for (size_t i = 0; i < n; ++i) q[i] = 0.0;
    if (ntime_out) *ntime_out = -1;
    return DVM_OK;
  }
  HIP_TRY(c, hipMemcpyAsync(q, c->d_heat, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (ntime_out) *ntime_out = c->heat_ntime;
  return DVM_OK;
}

extern "C" int dvm_collide_trace(dvm_ctx* c, int32_t ntime, int32_t cell, int64_t max_records, int32_t* irec,
                                 double* drec, int64_t* n_out) {
  if (!c || !n_out) return fail(c, DVM_E_ARG, "null argument");
  if (cell < 0 || cell >= c->ncell) return fail(c, DVM_E_ARG, "cell out of range");
  if (!c->f_valid || !c->props_valid) return fail(c, DVM_E_STATE, "trace needs f and its moments");
  if (max_records > 0 && (!irec || !drec)) return fail(c, DVM_E_ARG, "null record buffers");
  // The CDF rebuild below commits the equilibrium cache exactly as the next
  // collide would (same f, same moments), so the trace leaves later steps
  // unchanged.
  if (!collides(c)) { *n_out = 0; return DVM_OK; }
  if (int rc = cdf_async(c, true)) return rc;
  int h[4];
  HIP_TRY(c, hipMemcpyAsync(h, c->counts + (size_t)cell * 4, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  const long long total = h[3];
  const long long nrec = max_records < total ? max_records : total;
  if (nrec > 0) {
    int* d_i = nullptr;
    double* d_d = nullptr;
    int rc;
    if ((rc = dalloc(c, &d_i, (size_t)32 * nrec)) || (rc = dalloc(c, &d_d, (size_t)16 * nrec))) {
      if (d_i) (void)hipFree(d_i);
      return rc;
    }
    const unsigned blocks = (unsigned)((nrec + kCollBS - 1) / kCollBS);
    hipError_t e = hipSuccess;
    if (c->d_su) {
      // the traced cell reads the list from the cursor, which does not move
      const int per = stream_per_sample(c);
      const long long base = c->su_cursor;
      e = hipMemcpyAsync(c->d_sbase + cell, &base, sizeof base, hipMemcpyHostToDevice, c->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
      if (e == hipSuccess && total * per > c->su_n - c->su_cursor) {
        (void)hipFree(d_i);
        (void)hipFree(d_d);
        return fail(c, DVM_E_ARG, "uniform stream: the trace reads %lld values, %lld are left", total * per,
                    c->su_n - c->su_cursor);
      }
      if (e == hipSuccess)
        hipLaunchKernelGGL(k_trace<true>, dim3(blocks), dim3(kCollBS), 0, c->stream, c->g, c->collp, ntime, cell, c->f,
                           c->eqv, c->cdf, c->tabs, c->counts, c->factors, (long long)nrec, d_i, d_d, c->d_su,
                           c->d_sbase, per);
    } else {
      hipLaunchKernelGGL(k_trace<false>, dim3(blocks), dim3(kCollBS), 0, c->stream, c->g, c->collp, ntime, cell, c->f,
                         c->eqv, c->cdf, c->tabs, c->counts, c->factors, (long long)nrec, d_i, d_d,
                         (const double*)nullptr, (const long long*)nullptr, 0);
    }
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(irec, d_i, (size_t)32 * nrec * sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(drec, d_d, (size_t)16 * nrec * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d_i);
    (void)hipFree(d_d);
    if (e != hipSuccess) return fail(c, DVM_E_HIP, "dvm_collide_trace: %s", hipGetErrorString(e));
  }
  *n_out = total;
  return DVM_OK;
}


extern "C" int dvm_set_uniform_stream(dvm_ctx* c, const double* u, int64_t n) {
  if (!c) return fail(c, DVM_E_ARG, "null context");
  if (n < 0 || (n > 0 && !u)) return fail(c, DVM_E_ARG, "uniform stream: bad length or null pointer");
  for (int64_t q = 0; q < n; ++q)
    if (!(u[q] >= 0.0 && u[q] < 1.0)) return fail(c, DVM_E_ARG, "uniform stream: u[%lld] = %g is not in [0, 1)",
                                                 (long long)q, u[q]);
  HIP_TRY(c, hipStreamSynchronize(c->stream));  // queued collisions may still read the old list
  if (c->d_su) (void)hipFree(c->d_su);
  c->d_su = nullptr;
  c->su_n = 0;
  c->su_cursor = 0;
  if (n == 0) return DVM_OK;  // back to Philox
  if (!c->d_sbase) {
    if (int rc = dalloc(c, &c->d_sbase, (size_t)c->ncell)) return rc;
    if (int rc = dalloc(c, &c->d_stotal, 1)) return rc;
  }
  if (int rc = dalloc(c, &c->d_su, (size_t)n)) return rc;
  HIP_TRY(c, hipMemcpy(c->d_su, u, (size_t)n * sizeof(double), hipMemcpyHostToDevice));
  c->su_n = n;
  return DVM_OK;
}

extern "C" int dvm_uniform_stream_position(dvm_ctx* c, int64_t* consumed) {
  if (!c || !consumed) return fail(c, DVM_E_ARG, "null argument");
  *consumed = c->su_cursor;
  return DVM_OK;
}

#include "dvm_batch.hip"
