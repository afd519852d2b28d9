// dvm_kernels.h -- kernel parameter blocks and declarations shared by the
// kernels (dvm_kernels.hip) and the context / C-ABI code (dvm_capi.hip).
#pragma once

#include "dvm_device.h"
#include "../../include/dvm_hip.h"

namespace dvm {

// The FP32-collision precisions (the remap in FP64 or, on equally spaced
// axes, in FP32).
inline bool prec_f32(int p) { return p == DVM_PREC_FP32 || p == DVM_PREC_FP32_REMAP32; }

constexpr int kMomBS = 256;      // moments / wall-flux workgroup
constexpr int kMomCells = 4;     // cells per k_moments workgroup
constexpr int kCdfBS = 256;      // CDF workgroup of the batch
constexpr int kCdfCtxBS = 256;   // CDF workgroup of a context with many cells (256: 0.80 ms on the
                                 // 64x64 x 25^3 cavity, 512: 1.05, 128: 0.96)
constexpr int kCdfV = 8;         // consecutive nodes per lane in the CDF scan
constexpr int kCdfWideBS = 1024; // CDF workgroup when the cells are too few to fill the chip
constexpr int kCdfWideV = 4;     //   (four consecutive nodes per lane)
// LDS floats for the CDF scan's group ends (both distributions); cells whose
// 2 ceil(N/V) group ends exceed it build their bucket tables from a full read
constexpr int kCdfSumCap = 12288;
// Dynamic LDS floats of a CDF launch whose largest grid has max_n nodes.
inline int cdf_sum_cap(int max_n, int V) {
  const int need = 2 * ((max_n + V - 1) / V);
  return need < kCdfSumCap ? need : kCdfSumCap;
}
constexpr int kCollBS = 1024;    // collision workgroup: 16 waves share one cell's LDS slice
constexpr int kConvBS = 256;     // convection lanes per workgroup
constexpr int kWallBS = 1024;    // wall flux-balance lanes per workgroup (one workgroup per wall cell)
#ifndef DVM_CONV_STRIP
#define DVM_CONV_STRIP 4
#endif
#ifndef DVM_CONV_DEPTH
#define DVM_CONV_DEPTH 2
#endif
#ifndef DVM_CONV_SEGROWS
#define DVM_CONV_SEGROWS 8
#endif
constexpr int kConvStrip = DVM_CONV_STRIP;      // 2-D sweep: adjacent columns per workgroup
constexpr int kConvDepth = DVM_CONV_DEPTH;      // 2-D sweep: rows loaded ahead
constexpr int kConvSegRows = DVM_CONV_SEGROWS;  // 2-D sweep: rows per workgroup
constexpr int kAxisSlots = 384;  // LDS doubles reserved for the three axes (<= 128 points each)
constexpr int kMaxAxis = 128;
// LDS bytes available to one workgroup on gfx950
constexpr size_t kLdsBytes = 160 * 1024;

struct CdfParams {
  double mass, cdf_tolerance, dt, kn, sigma;
  double rms_sq, beta3;          // reference sample-count rule
  int sample_mode;
  long long samples_per_cell;
  int fmc;                       // full Monte Carlo kernel: |f| CDF, one integral
};

struct CollParams {
  uint64_t seed;
  double mass, vhs_exponent, ralpha;
  int hard_sphere;               // vhs_exponent == 1: g^(2-2w) == g exactly
  int vss;                       // |1/alpha - 1| > 1e-13: Bird's VSS deflection
  int num_repl;
  float inv_repl;                // 1 / num_repl (FP32 path)
};

struct CellView {
  const double* f;
  const double* eqv;
  const double* eqc;
  const double* neqc;
  const int2* eqtab;
  const int2* neqtab;
  double eq_total, neq_total;    // eqc[N], neqc[N]
};

struct ConvParams {
  int nx, ny, max_along;
  double fx, fy;                 // dt/dx, dt/dy
  double acc;                    // accommodation coefficient
  int wall_side[4];              // diffuse/specular wall on side s
  int couette_side[4];
  int bc[2];                     // 1-D boundary types (left, right)
  int izflag, jzflag;            // zero-plane flags of boundary_conditions.f90:662-673
  double flux_out[4];            // 2-D: sum |v_n| M_wall beta^3 over the wall's outgoing half-space;
                                 // 1-D walls (sides 0, 1): sum M_wall beta^3 over the outgoing half
  int order;                     // 1-D sweep order (1 or 2)
};

template <int C, int BS>
__global__ void k_moments(Grid g, const double* __restrict__ f, double* __restrict__ props, double mass,
                          int ncell);
__global__ void k_moments_ext(Grid g, const double* __restrict__ f, const double* __restrict__ props, double mass,
                              double* __restrict__ ext);
// The CDF pass writes the deviation sign bytes beside the CDFs (ceil(N/8)
// bytes at byte offset 2 (N+1) per cell, the CDF slice's offset in doubles).
// The FP32 path's moments kernel writes the Cholesky factor of
// sum |f| phi phi^T per cell (kCmat doubles: 15 factor entries and a
// not-positive-definite flag).
constexpr int kCmat = 16;
template <int BS, int V, bool kFused>
__global__ void k_cdf(Grid g, CdfParams cp, const double* __restrict__ f, const double* __restrict__ props,
                      double* __restrict__ eqv, double* __restrict__ cdf, int2* __restrict__ tabs,
                      double* __restrict__ eqpar,
                      int* __restrict__ counts, double* __restrict__ factors, int sum_cap,
                      unsigned char* __restrict__ sgn, int eq_mode);
template <int C, int BS>
__global__ void k_moments_cmat(Grid g, const double* __restrict__ f, double* __restrict__ props,
                               double* __restrict__ cmat, double mass, int ncell);
template <bool kStream>
__global__ void k_trace(Grid g, CollParams cp, int ntime, int cell, const double* __restrict__ f,
                        const double* __restrict__ eqv, const double* __restrict__ cdf,
                        const int2* __restrict__ tabs, const int* __restrict__ counts,
                        const double* __restrict__ factors, long long max_records, int* __restrict__ irec,
                        double* __restrict__ drec, const double* __restrict__ su,
                        const long long* __restrict__ sbase, int sper);
__global__ void k_stream_bases(const int* __restrict__ counts, int ncell, long long cursor, int per,
                               long long* __restrict__ base, long long* __restrict__ total);
__global__ void k_remap_batch(Grid g, long long n, const double* __restrict__ uvw, double f_off,
                              int* __restrict__ st, double* __restrict__ m);
__global__ void k_wall_ratio(Grid g, ConvParams p, const double* __restrict__ f, double* __restrict__ ratio);
__global__ void k_spec_ghost(Grid g, ConvParams p, const double* __restrict__ f, double* __restrict__ spec,
                             int* __restrict__ err);
__global__ void k_fill(double* __restrict__ a, long long n, double v);
void launch_convect2d(const Grid& g, const ConvParams& p, const double* f, double* fo, const double* wallM,
                      const double* ratio, const double* spec, hipStream_t st);
__global__ void k_convect1d(Grid g, ConvParams p, const double* __restrict__ f, double* __restrict__ fo,
                            const double* __restrict__ inflow, const double* __restrict__ wallM,
                            const double* __restrict__ ratio, const double* __restrict__ spec);
__global__ void k_wall1d_ratio(Grid g, ConvParams p, const double* __restrict__ f, double* __restrict__ ratio);
__global__ void k_spec1d(Grid g, ConvParams p, const double* __restrict__ f, double* __restrict__ spec,
                         int* __restrict__ err);

}  // namespace dvm
