// dvm_device.h -- device-side building blocks of the gfx950 DVM time step.
//
// Every routine here is written for one lane of a 64-wide wavefront; the
// kernels in dvm_kernels.hip compose them. FP64 routines follow the reference's
// expression order (the library is built with -ffp-contract=off) so that the
// sampled nodes and remap stencils equal the reference's bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dvm {

// constants.f90:30 -- the reference's value of pi (a typo for 3.14159...),
// kept because it enters the equilibrium normalisation and the scattering
// azimuth.
constexpr double kPiRef = 3.1459265358979323;
constexpr double kDoubleTol = 1.0e-13;  // constants.f90:52
constexpr int kNBucket = 1000;          // search table size, variance_reduction.f90:471-493
constexpr int kSearchCap = 96;          // bound on bisection steps (>> log2 of any grid)
// Samples of one cell per collision phase: at most 2^30, each integral's
// count at most a third of that, so that sample indices, per-cell totals and
// the lanes' running indices stay within 32-bit integers.
constexpr int kMaxCellSamples = 1 << 30;
constexpr int kMaxIntegralSamples = kMaxCellSamples / 3;
// Device error flags (bit values of the context's / batch's error word):
constexpr int kErrRemap = 1;       // a post-collision remap found no exterior point (remapping.f90:348-354)
constexpr int kErrSpecRemap = 2;   // a specular ghost remap found no exterior point
constexpr int kErrSpecNode = 4;    // a specular ghost stencil point left the grid
constexpr int kErrConserve = 8;    // the FP32 path's conservation system was not positive definite

// Velocity grid as seen by a kernel. Axes and quadrature weights may live in
// global memory or be staged into LDS by the caller.
struct Grid {
  int n[3];        // points per axis
  int N;           // n0*n1*n2
  int nxy;         // n0*n1
  double vmin[3];  // first coordinate of each axis
  double vmax[3];  // last coordinate of each axis
  double factor[3];// (n-1)/(vmax-vmin): first guess of the nearest-node search
  int zero[3], neg[3], pos[3];  // find_zero_point results per axis
  double inv_n0, inv_nxy;       // 1/n0 and 1/(n0*n1): node index decoding without integer division
  const double* ax[3];
  const double* beta[3];
};

// ------------------------------------------------------------------ Philox
// Philox4x32-10 (Salmon et al., SC'11), the counter-based generator of
// Random123; one call gives four 32-bit words.
__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1, uint32_t out[4]) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    // each 32 x 32 -> 64-bit product in one v_mad_u64_u32
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
    const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
    const uint32_t n0 = hi1 ^ c1 ^ k0;
    const uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Philox block of sample `sample` for counter (sample, tag, cell, ntime):
// tag = integral (1..3) | slot << 2. Slot 0 gives (u_A, u_B, u_phi, u_theta)
// of replenish location 1, slot r gives (u_phi, u_theta) of location r >= 2.
// A uniform is word * 2^-32 in [0,1), the resolution of the reference's
// gsl_rng_uniform on mt19937.
__device__ __forceinline__ void philox_block(uint64_t seed, uint32_t sample, uint32_t tag, uint32_t cell,
                                             uint32_t ntime, uint32_t out[4]) {
  philox4x32_10(sample, tag, cell, ntime, (uint32_t)seed, (uint32_t)(seed >> 32), out);
}
__device__ __forceinline__ double u32_to_uniform(uint32_t x) { return (double)x * 0x1.0p-32; }

// ---------------------------------------------------------------- signs
// isgn / dsgn (math_utils.f90:38-60): sign(1, x) with sign(1, -0.0) = -1.
__device__ __forceinline__ int isgn(double x) { return signbit(x) ? -1 : 1; }
__device__ __forceinline__ double dsgn(double x) { return signbit(x) ? -1.0 : 1.0; }
__device__ __forceinline__ int isign1(int k) { return k >= 0 ? 1 : -1; }

// ---------------------------------------------------------- CDF sampling
// referenced_search (sort_and_search.f90:40-86): inverse-CDF pick of node l
// with u*cumul[N] in (cumul[l], cumul[l+1]], bracketed first by a
// 1000-bucket table (tab[b] = (left, right)).
__device__ __forceinline__ int referenced_search(const double* __restrict__ cumul, int N,
                                                 const int2* __restrict__ tab, double u) {
  if (u < kDoubleTol) return 0;
  int b = (int)ceil(u * 1000.0);
  b = b < 1 ? 1 : (b > kNBucket ? kNBucket : b);
  const int2 lr = tab[b - 1];
  int left = lr.x, right = lr.y;
  int index = left + (right - left) / 2;
  if (left == right) return index;
  const double adj = u * cumul[N];
  for (int it = 0; it < kSearchCap; ++it) {
    const double c0 = cumul[index], c1 = cumul[index + 1];
    if (adj > c0 && adj <= c1) break;
    if (adj > c0) left = index; else right = index;
    index = left + (right - left) / 2;
  }
  return index;
}

// ---------------------------------------------------------------- remap
// The velocity axes in the arithmetic R of a remap: FP64, whose stencils and
// weights equal the reference's bit for bit, or FP32 (the FP32-collision
// path, whose replenishment weights are FP32 like the rest of its collision
// arithmetic).
template <class R>
struct Axes {
  const R* a[3];
  int n[3];
  R vmin[3], vmax[3], factor[3];
  // Closed-form remap on equally spaced axes (remap_uniform_interior): the
  // spacing h of each axis, 1 / (2 (hx^2 + hy^2 + hz^2)), and the FP32 copies
  // of vmin and factor that locate the nearest node.
  R h[3], rden;
  float fvmin[3], ffactor[3];
};

// Spacings and closed-form constants of Axes x from its vmin, vmax and n.
template <class R>
__device__ __forceinline__ void set_uniform_consts(Axes<R>& x) {
  R s = 0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    x.h[d] = x.n[d] > 1 ? (x.vmax[d] - x.vmin[d]) / (R)(x.n[d] - 1) : (R)0;
    s += x.h[d] * x.h[d];
    x.fvmin[d] = (float)x.vmin[d];
    x.ffactor[d] = (float)x.factor[d];
  }
  x.rden = s > (R)0 ? (R)1 / ((R)2 * s) : (R)0;
}

__device__ __forceinline__ Axes<double> grid_axes(const Grid& g) {
  Axes<double> x;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    x.a[d] = g.ax[d];
    x.n[d] = g.n[d];
    x.vmin[d] = g.vmin[d];
    x.vmax[d] = g.vmax[d];
    x.factor[d] = g.factor[d];
  }
  set_uniform_consts(x);
  return x;
}

// a / b: correctly rounded in FP64; in FP32 a times the hardware reciprocal
// (v_rcp_f32, 1 ulp), two instructions instead of the scaled Newton sequence.
template <class R>
__device__ __forceinline__ R quot(R a, R b) {
  if constexpr (sizeof(R) == 4) return a * __builtin_amdgcn_rcpf(b); else return a / b;
}
template <class R>
__device__ __forceinline__ int isgn_r(R x) { return signbit(x) ? -1 : 1; }

// binary_search_with_first_guess (sort_and_search.f90:130-183): nearest axis
// node to val, ties to the lower index, clamped to the axis ends. The
// reference's bisection can only stop at the unique index with
// a[index] < val <= a[index+1] (then it keeps the closer end), so its result
// does not depend on the first guess. A uniform grid's guess is within one
// node of that bracket, so the bracket is settled by one comparison at the
// guess; a walk and then the bisection run only when it is not (non-uniform
// axes).
template <class R>
__device__ __forceinline__ int nearest_node(const R* __restrict__ a, int mx, R val, int index, R lo, R hi) {
  // lo = a[0], hi = a[mx], held in registers
  if (lo > val) return 0;
  if (hi < val) return mx;
  index = index < 0 ? 0 : (index > mx ? mx : index);
  const R ag = a[index];
  if (fabs(val - ag) < (R)kDoubleTol) return index;
  // the bracket is [index, index+1] or [index-1, index] when the guess is
  // within one node (uniform axes): one compare and two independent reads
  int i = (val > ag) ? index : index - 1;
  i = i < 0 ? 0 : (i > mx - 1 ? mx - 1 : i);
  R a0 = a[i], a1 = a[i + 1];
  if (!(val > a0 && val <= a1)) {
    // farther guesses (non-uniform axes): walk, then bisect
    index = i;
#pragma unroll
    for (int w = 0; w < 2; ++w) {
      if (val > a1 && index + 1 < mx) { ++index; a0 = a1; a1 = a[index + 1]; }
      else if (!(val > a0) && index > 0) { --index; a1 = a0; a0 = a[index]; }
    }
    if (!(val > a0 && val <= a1)) {
      int left = 0, right = mx;
      for (int it = 0; it < kSearchCap; ++it) {
        a0 = a[index]; a1 = a[index + 1];
        if (val > a0 && val <= a1) break;
        if (val > a0) left = index; else right = index;
        index = left + (right - left) / 2;
      }
    }
    i = index;
  }
  return (fabs(val - a0) > fabs(val - a1)) ? i + 1 : i;
}

// First guess of nearest_node, clamped to [-1, mx + 1]. The reference rounds
// with nint (halves away from zero); this rounds halves to even, one
// instruction. The two differ only when val lies exactly midway between two
// nodes of a uniform axis, where either guess is one of the two bracketing
// nodes and, val being a half-spacing from both, neither takes the guess's
// early exit: the bracket and so the result are the same.
template <class R>
__device__ __forceinline__ int first_guess(R factor, R val, R vmin, int mx) {
  R g = rint(factor * (val - vmin));
  g = g < (R)-1 ? (R)-1 : (g > (R)(mx + 1) ? (R)(mx + 1) : g);
  return (int)g;
}

// find_boundary_stencil (remapping.f90:396-438)
template <class R>
__device__ __forceinline__ void boundary_stencil(int& sa, int& sb, int& sc, R abc, int oi, int oj, int ok,
                                                 const R* X, const R* Y, const R* Z, int imax, int jmax, int kmax) {
  while (oi + sa > 0 && oi + sa < imax) {
    if (fabs(abc) > (R)0.5 * fabs(X[oi + sa] - X[oi])) sa = sa + isign1(sa); else break;
  }
  while (oj + sb > 0 && oj + sb < jmax && oj - sb > 0 && oj - sb < jmax) {
    if (fabs(abc) > (R)0.5 * fabs(Y[oj + sb] - Y[oj])) sb = sb + isign1(sb); else break;
  }
  while (ok + sc > 0 && ok + sc < kmax && ok - sc > 0 && ok - sc < kmax) {
    if (fabs(abc) > (R)0.5 * fabs(Z[ok + sc] - Z[ok])) sc = sc + isign1(sc); else break;
  }
}

// Result of a conservative remap: origin node, interior offsets, exterior
// flag (bit 0 x, bit 1 y, bit 2 z) and the seven weights.
template <class R>
struct StencilT {
  int oi, oj, ok, a, b, c, flag;
  R m[7];  // f_o, f_ix, f_iy, f_iz, f_ex, f_ey, f_ez
};
using Stencil = StencilT<double>;

// Nearest node of axis d. kUniform (the FP32 path on equally spaced axes): the
// rounded position (v - vmin) (n - 1) / (vmax - vmin), clamped to the axis --
// the nearest node up to FP32 rounding at a midpoint, without the bracket
// search; otherwise the reference's search (nearest_node).
template <class R, bool kUniform>
__device__ __forceinline__ int nearest_on_axis(const Axes<R>& g, int d, R val) {
  const int mx = g.n[d] - 1;
  const int guess = first_guess<R>(g.factor[d], val, g.vmin[d], mx);
  if constexpr (kUniform) return guess < 0 ? 0 : (guess > mx ? mx : guess);
  else return nearest_node<R>(g.a[d], mx, val, guess, g.vmin[d], g.vmax[d]);
}

// compute_remapping_quantities (remapping.f90:92-394), the non-uniform branch
// that trunk always takes. Moments 0, 1 and 2 of the 7-point stencil equal
// f_off at (u,v,w) in exact arithmetic. flag == 0 is the reference's fatal case.
template <class R, bool kUniform = false>
__device__ __forceinline__ void remap_t(const Axes<R>& g, R u, R v, R w, R f_off, StencilT<R>& s) {
  const R *X = g.a[0], *Y = g.a[1], *Z = g.a[2];
  const int imax = g.n[0] - 1, jmax = g.n[1] - 1, kmax = g.n[2] - 1;
  int corner = 0;
  int oi = nearest_on_axis<R, kUniform>(g, 0, u);
  if (oi >= imax) { oi = imax; ++corner; } else if (oi <= 0) { oi = 0; ++corner; }
  int oj = nearest_on_axis<R, kUniform>(g, 1, v);
  if (oj >= jmax) { oj = jmax; ++corner; } else if (oj <= 0) { oj = 0; ++corner; }
  int ok = nearest_on_axis<R, kUniform>(g, 2, w);
  if (ok >= kmax) { ok = kmax; ++corner; } else if (ok <= 0) { ok = 0; ++corner; }
  if (corner == 3) {
    oi -= isign1(oi - 1);
    oj -= isign1(oj - 1);
    ok -= isign1(ok - 1);
  }
  const R ox = X[oi], oy = Y[oj], oz = Z[ok];
  const R a = u - ox, b = v - oy, c = w - oz;
  int sa = isgn_r<R>(a), sb = isgn_r<R>(b), sc = isgn_r<R>(c);
  if (oi + sa > imax || oi + sa < 0) { sa = -sa; boundary_stencil<R>(sa, sb, sc, a, oi, oj, ok, X, Y, Z, imax, jmax, kmax); }
  if (oj + sb > jmax || oj + sb < 0) { sb = -sb; boundary_stencil<R>(sb, sa, sc, b, oj, oi, ok, Y, X, Z, jmax, imax, kmax); }
  if (ok + sc > kmax || ok + sc < 0) { sc = -sc; boundary_stencil<R>(sc, sb, sa, c, ok, oj, oi, Z, Y, X, kmax, jmax, imax); }
  const R ix = X[oi + sa] - ox, iy = Y[oj + sb] - oy, iz = Z[ok + sc] - oz;
  int flag = 0;
  R ex = 0, ey = 0, ez = 0;
  if (oi - sa <= imax && oi - sa >= 0) { flag |= 1; ex = X[oi - sa] - ox; }
  if (oj - sb <= jmax && oj - sb >= 0) { flag |= 2; ey = Y[oj - sb] - oy; }
  if (ok - sc <= kmax && ok - sc >= 0) { flag |= 4; ez = Z[ok - sc] - oz; }
  const R num = f_off * (a * a + b * b + c * c - a * ix - b * iy - c * iz);
  const R dx = ex * (ex - ix), dy = ey * (ey - iy), dz = ez * (ez - iz);
  R den;
  switch (flag) {
    case 7: den = dx + dy + dz; break;
    case 1: den = dx; break;
    case 2: den = dy; break;
    case 4: den = dz; break;
    case 3: den = dx + dy; break;
    case 5: den = dx + dz; break;
    case 6: den = dy + dz; break;
    default: den = 1; break;
  }
  const R fe = (flag != 0) ? quot<R>(num, den) : (R)0;
  const R fex = (flag & 1) ? fe : (R)0;
  const R fey = (flag & 2) ? fe : (R)0;
  const R fez = (flag & 4) ? fe : (R)0;
  const R fix = quot<R>(f_off * a - fex * ex, ix);
  const R fiy = quot<R>(f_off * b - fey * ey, iy);
  const R fiz = quot<R>(f_off * c - fez * ez, iz);
  s.oi = oi; s.oj = oj; s.ok = ok; s.a = sa; s.b = sb; s.c = sc; s.flag = flag;
  s.m[0] = f_off - fix - fiy - fiz - fex - fey - fez;
  s.m[1] = fix; s.m[2] = fiy; s.m[3] = fiz; s.m[4] = fex; s.m[5] = fey; s.m[6] = fez;
  if (flag == 0) {
#pragma unroll
    for (int q = 0; q < 7; ++q) s.m[q] = 0;
  }
}

// Conservative remap of a unit mass at (u, v, w) on equally spaced axes, for
// the FP32-collision path, when the nearest node o lies off every axis end
// (1 <= o_d <= n_d - 2); returns false, scattering nothing, otherwise. The
// stencil is remap_t's (interior points o + s x, exterior o - s x, s = sign of
// the offset, flag 7), and with ix = s hx, ex = -s hx its weight solve has the
// closed form
//   f_e  = (a^2 + b^2 + c^2 - |a| hx - |b| hy - |c| hz) / (2 (hx^2 + hy^2 + hz^2)),
//   f_ix = |a| / hx + f_e  (likewise y, z),   f_o = 1 - f_ix - f_iy - f_iz - 3 f_e,
// equal to remap_t's in exact arithmetic (moments 0, 1, 2 conserved) and
// evaluated in R without divisions. The nearest node is located in FP32 (the
// node closest to the FP32 velocity up to a rounding at a midpoint, where
// either node gives a valid stencil). scatter(i, j, k, node, r * weight).
template <class R, class Scatter>
__device__ __forceinline__ bool remap_uniform_interior(const Axes<R>& g, float u, float v, float w, R r, int nx,
                                                       int nxy, Scatter&& scatter) {
  const float pu = fminf(fmaxf((u - g.fvmin[0]) * g.ffactor[0], -1.0f), (float)g.n[0]);
  const float pv = fminf(fmaxf((v - g.fvmin[1]) * g.ffactor[1], -1.0f), (float)g.n[1]);
  const float pw = fminf(fmaxf((w - g.fvmin[2]) * g.ffactor[2], -1.0f), (float)g.n[2]);
  const int oi = (int)rintf(pu), oj = (int)rintf(pv), ok = (int)rintf(pw);
  if ((unsigned)(oi - 1) >= (unsigned)(g.n[0] - 2) || (unsigned)(oj - 1) >= (unsigned)(g.n[1] - 2) ||
      (unsigned)(ok - 1) >= (unsigned)(g.n[2] - 2))
    return false;
  const R a = (R)u - (g.vmin[0] + (R)oi * g.h[0]);
  const R b = (R)v - (g.vmin[1] + (R)oj * g.h[1]);
  const R c = (R)w - (g.vmin[2] + (R)ok * g.h[2]);
  const R fa = fabs(a), fb = fabs(b), fc = fabs(c);
  const R fe = (a * a + b * b + c * c - fa * g.h[0] - fb * g.h[1] - fc * g.h[2]) * g.rden;
  const R fix = fa * g.factor[0] + fe, fiy = fb * g.factor[1] + fe, fiz = fc * g.factor[2] + fe;
  const R fo = (R)1 - fix - fiy - fiz - (R)3 * fe;
  const int sa = signbit(a) ? -1 : 1, sb = signbit(b) ? -1 : 1, sc = signbit(c) ? -1 : 1;
  const int L = oi + nx * oj + nxy * ok;
  const R re = r * fe;
  scatter(oi, oj, ok, L, (double)(r * fo));
  scatter(oi + sa, oj, ok, L + sa, (double)(r * fix));
  scatter(oi, oj + sb, ok, L + sb * nx, (double)(r * fiy));
  scatter(oi, oj, ok + sc, L + sc * nxy, (double)(r * fiz));
  scatter(oi - sa, oj, ok, L - sa, (double)re);
  scatter(oi, oj - sb, ok, L - sb * nx, (double)re);
  scatter(oi, oj, ok - sc, L - sc * nxy, (double)re);
  return true;
}

__device__ __forceinline__ void remap(const Grid& g, double u, double v, double w, double f_off, Stencil& s) {
  remap_t<double>(grid_axes(g), u, v, w, f_off, s);
}

// Node index of stencil point q (0 origin, 1..3 interior, 4..6 exterior).
template <class S>
__device__ __forceinline__ void stencil_ijk(const S& s, int q, int& i, int& j, int& k) {
  i = s.oi;
  j = s.oj;
  k = s.ok;
  switch (q) {
    case 1: i += s.a; break;
    case 2: j += s.b; break;
    case 3: k += s.c; break;
    case 4: i -= s.a; break;
    case 5: j -= s.b; break;
    case 6: k -= s.c; break;
    default: break;
  }
}

template <class S>
__device__ __forceinline__ int stencil_node(const S& s, int q, int nx, int nxy) {
  int i, j, k;
  stencil_ijk(s, q, i, j, k);
  return i + nx * j + nxy * k;
}

// ------------------------------------------------------- wave reductions
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Inclusive prefix sum across the 64 lanes of a wavefront.
__device__ __forceinline__ double wave_inclusive_scan(double v, int lane) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const double t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}

}  // namespace dvm
