// dvm_kernels.hip -- gfx950 kernels of the DVM time step.
//
// Layout in HBM (all FP64 unless noted), N = velocity nodes, C = cells:
//   f, fd        [C][N]      distribution and its double buffer, node fastest
//   eqv          [C][N]      cached discrete equilibrium per node (f_eq * beta^3)
//   eqc, neqc    [C][N+1]    cumulative equilibrium / |deviation| distributions
//   eqtab,neqtab [C][1000]   int2 bucket brackets of the inverse-CDF search
//   eqpar        [C][5]      (dens,u,v,w,T) the equilibrium was built for
//   props        [C][8]      moments (DVM_P_*)
//   counts       [C][4]      int32 samples of the 3 integrals + their sum
//   factors      [C][3]      depletion factor of each integral
// Kernels:
//   k_moments    one or four cells per workgroup; one pass, FP64 wave reductions
//   k_cdf        one workgroup per cell, one pass over f: cache test,
//                equilibrium rebuild on a miss, both prefix scans, bucket
//                tables, pair counts
//   k_collide    one workgroup per cell; samples spread over the lanes, the
//                cell's increments accumulated in LDS (ds_add_f64) when the
//                slice fits, else global atomics into fd
//   k_convect2d  one lane per node of a strip of adjacent columns, walking y;
//                interior donor-cell upwind and the trunk boundary-cell update
//                with wall / neighbour ghosts
//   k_convect1d  one lane per (cell,node); zone-grid 1-D upwind and in/out flow
//   k_wall_ratio, k_spec_ghost  diffuse flux balance and specular remap ghosts
#include "dvm_kernels.h"

#include <type_traits>

namespace dvm {

// ------------------------------------------------------------ reductions
template <int BS>
__device__ __forceinline__ double block_sum(double v, double* red) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  v = wave_sum(v);
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < BS / 64; ++w) t += red[w];
  __syncthreads();
  return t;
}

// Inclusive scan over the workgroup; returns the lane's prefix and the
// workgroup total in `total`.
template <int BS>
__device__ __forceinline__ double block_scan(double v, double* red, double& total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const double inc = wave_inclusive_scan(v, lane);
  if (lane == 63) red[wid] = inc;
  __syncthreads();
  double off = 0.0, tot = 0.0;
#pragma unroll
  for (int w = 0; w < BS / 64; ++w) {
    const double s = red[w];
    if (w < wid) off += s;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return inc + off;
}

// Node l -> (i, j, k), i fastest. floor((l + 0.5) / n) is computed as
// (l + 0.5) * (1/n) in FP64: for l < 2^24 the product is within 1e-9 of the
// quotient, which itself is at least 0.5/n away from an integer, so the floor
// is exact (no integer division, which is a long instruction sequence).
__device__ __forceinline__ void node_ijk(const Grid& g, int l, int& i, int& j, int& k) {
  k = (int)(((double)l + 0.5) * g.inv_nxy);
  const int r = l - k * g.nxy;
  j = (int)(((double)r + 0.5) * g.inv_n0);
  i = r - j * g.n[0];
}

// --------------------------------------------------------------- moments
// calculate_physical_properties (physical_props.f90:285-485) for one
// monatomic species: n, u, v, w, <|v|>, T and the derived fields, in one pass
// over f. The reference's second pass sum f |c - u|^2 is formed as
// sum f |c|^2 - u . sum f c (algebraically equal; the distribution is read once).
// Moments of one cell from its six sums {n, n u, n v, n w, sum f |c|^2,
// sum f |c|}.
__device__ __forceinline__ void moments_finish(const double* tot, double* __restrict__ p, double mass) {
  const double dens = tot[0];
  if (dens == 0.0) {
    for (int q = 0; q < DVM_NPROP; ++q) p[q] = 0.0;
    return;
  }
  const double u = tot[1] / dens, v = tot[2] / dens, w = tot[3] / dens;
  const double st = tot[4] - (u * tot[1] + v * tot[2] + w * tot[3]);
  const double t = 2.0 * mass * st / (3.0 * dens);
  p[DVM_P_DENS] = dens;
  p[DVM_P_U] = u;
  p[DVM_P_V] = v;
  p[DVM_P_W] = w;
  p[DVM_P_TRTEMP] = t;
  p[DVM_P_TEMP] = (3.0 * t + 0.0 + 0.0) / 3.0;  // compute_overall_temperature, no internal dof
  p[DVM_P_PRESSURE] = 0.5 * dens * t;
  p[DVM_P_AVGSPEED] = tot[5] / dens;
}

// Sums of C cells f[c * stride + l] over the nodes, one workgroup of BS
// lanes: lane t takes nodes t, t + BS, ..., so the node's coordinates and
// speed are computed once for the C cells and the C loads of an iteration are
// independent. Each cell's sums are added in the same order whatever C is.
template <int C, int BS = kMomBS>
__device__ __forceinline__ void moments_sums(const Grid& g, const double* __restrict__ f, size_t stride, int nc,
                                             double (*red)[6][BS / 64]) {
  const double *X = g.ax[0], *Y = g.ax[1], *Z = g.ax[2];
  double acc[C][6];
#pragma unroll
  for (int c = 0; c < C; ++c)
#pragma unroll
    for (int q = 0; q < 6; ++q) acc[c][q] = 0.0;
  for (int l = threadIdx.x; l < g.N; l += BS) {
    int i, j, k;
    node_ijk(g, l, i, j, k);
    const double x = X[i], y = Y[j], z = Z[k];
    const double c2 = x * x + y * y + z * z;
    const double sp = sqrt(c2);
    double v[C];
#pragma unroll
    for (int c = 0; c < C; ++c) v[c] = (c < nc) ? f[(size_t)c * stride + l] : 0.0;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      acc[c][0] += v[c];
      acc[c][1] += v[c] * x;
      acc[c][2] += v[c] * y;
      acc[c][3] += v[c] * z;
      acc[c][4] += v[c] * c2;
      acc[c][5] += v[c] * sp;
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < C; ++c)
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const double t = wave_sum(acc[c][q]);
      if (lane == 0) red[c][q][wid] = t;
    }
  __syncthreads();
}

template <int BS = kMomBS>
__device__ __forceinline__ void moments_total(const double (*red)[BS / 64], double* tot) {
  for (int q = 0; q < 6; ++q) {
    double t = 0.0;
    for (int w = 0; w < BS / 64; ++w) t += red[q][w];
    tot[q] = t;
  }
}

// calculate_physical_properties (physical_props.f90:285-485) for one
// monatomic species: n, u, v, w, <|v|>, T and the derived fields, in one pass
// over f. The reference's second pass sum f |c - u|^2 is formed as
// sum f |c|^2 - u . sum f c (algebraically equal; the distribution is read once).
__device__ __forceinline__ void moments_cell(const Grid& g, const double* __restrict__ fc, double* __restrict__ p,
                                             double mass, double (*red)[6][kMomBS / 64]) {
  moments_sums<1>(g, fc, 0, 1, red);
  if (threadIdx.x == 0) {
    double tot[6];
    moments_total(red[0], tot);
    moments_finish(tot, p, mass);
  }
}

// Moments of C cells (moments_cell) together with the Cholesky factor of
// M = sum |f| phi phi^T, phi = (1, x, y, z, |v|^2), per cell, which the FP32
// path's conservation correction of the next collision solves with
// (conserve_cell_factored): one pass over f of 6 + 15 sums per cell, the
// node's coordinates formed once for the C cells (f of cell c at
// fc + c * stride; nc <= C cells present). cm[c * kCmat ..] receives the
// factor L (row-major packed lower triangle, 15 values) and a 0, or 1 and
// zeros when M is not numerically positive definite.
template <int C, int BS>
struct MomCmatShared {
  double red[C][21][BS / 64];
  double M[C][25];
};

template <int C, int BS>
__device__ __forceinline__ void moments_cmat_cells(const Grid& g, const double* __restrict__ fc, size_t stride,
                                                   int nc, double* __restrict__ p, double* __restrict__ cm,
                                                   double mass, MomCmatShared<C, BS>& sh) {
  const double *X = g.ax[0], *Y = g.ax[1], *Z = g.ax[2];
  double a[C][21];
#pragma unroll
  for (int c = 0; c < C; ++c)
#pragma unroll
    for (int q = 0; q < 21; ++q) a[c][q] = 0.0;
  // with one cell the loads of U nodes are issued together
  constexpr int U = (C == 1) ? 4 : 1;
  for (int l0 = threadIdx.x; l0 < g.N; l0 += U * BS) {
    double fv[U][C];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int c = 0; c < C; ++c)
        fv[u][c] = (l0 + u * BS < g.N && c < nc) ? fc[(size_t)c * stride + l0 + u * BS] : 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int l = l0 + u * BS;
      if (l >= g.N) break;
      int i, j, k;
      node_ijk(g, l, i, j, k);
      const double x = X[i], y = Y[j], z = Z[k];
      const double c2 = x * x + y * y + z * z;
      const double sp = sqrt(c2);
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const double v = fv[u][c];
        // moments_sums' six, in its order
        a[c][0] += v;
        a[c][1] += v * x;
        a[c][2] += v * y;
        a[c][3] += v * z;
        a[c][4] += v * c2;
        a[c][5] += v * sp;
        // |f| phi_a phi_b, a <= b
        const double w = fabs(v);
        const double wx = w * x, wy = w * y, wz = w * z, we = w * c2;
        a[c][6] += w; a[c][7] += wx; a[c][8] += wy; a[c][9] += wz; a[c][10] += we;
        a[c][11] += wx * x; a[c][12] += wx * y; a[c][13] += wx * z; a[c][14] += wx * c2;
        a[c][15] += wy * y; a[c][16] += wy * z; a[c][17] += wy * c2;
        a[c][18] += wz * z; a[c][19] += wz * c2;
        a[c][20] += we * c2;
      }
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < C; ++c)
#pragma unroll
    for (int q = 0; q < 21; ++q) {
      const double t = wave_sum(a[c][q]);
      if (lane == 0) sh.red[c][q][wid] = t;
    }
  __syncthreads();
  const int c = threadIdx.x;
  if (c < nc) {
    double tot[6];
    for (int q = 0; q < 6; ++q) {
      double t = 0.0;
      for (int w = 0; w < BS / 64; ++w) t += sh.red[c][q][w];
      tot[q] = t;
    }
    moments_finish(tot, p + (size_t)c * DVM_NPROP, mass);
    double* M = sh.M[c];
    int o = 6;
    for (int r = 0; r < 5; ++r)
      for (int e = r; e < 5; ++e) {
        double t = 0.0;
        for (int w = 0; w < BS / 64; ++w) t += sh.red[c][o][w];
        M[r * 5 + e] = t;
        M[e * 5 + r] = t;
        ++o;
      }
    bool ok = true;
    for (int j = 0; j < 5 && ok; ++j) {
      double d = M[j * 5 + j];
      for (int k = 0; k < j; ++k) d -= M[j * 5 + k] * M[j * 5 + k];
      if (!(d > 0.0)) {
        ok = false;
        break;
      }
      d = sqrt(d);
      M[j * 5 + j] = d;
      for (int i = j + 1; i < 5; ++i) {
        double v = M[i * 5 + j];
        for (int k = 0; k < j; ++k) v -= M[i * 5 + k] * M[j * 5 + k];
        M[i * 5 + j] = v / d;
      }
    }
    double* out = cm + (size_t)c * kCmat;
    o = 0;
    for (int i = 0; i < 5; ++i)
      for (int j = 0; j <= i; ++j) out[o++] = ok ? M[i * 5 + j] : 0.0;
    out[15] = ok ? 0.0 : 1.0;
  }
}

// Moments and conservation factors of the cells of an FP32 context whose
// collisions take collide_fast: C cells per workgroup.
template <int C, int BS>
__global__ void __launch_bounds__(BS) k_moments_cmat(Grid g, const double* __restrict__ f,
                                                     double* __restrict__ props, double* __restrict__ cmat,
                                                     double mass, int ncell) {
  __shared__ MomCmatShared<C, BS> sh;
  const int c0 = blockIdx.x * C;
  const int nc = (ncell - c0 < C) ? ncell - c0 : C;
  moments_cmat_cells<C, BS>(g, f + (size_t)c0 * g.N, (size_t)g.N, nc, props + (size_t)c0 * DVM_NPROP,
                            cmat + (size_t)c0 * kCmat, mass, sh);
}

// Distance of one cell's f from equilibrium: sum |f - M| / n, where M is the
// discrete Maxwellian (times beta^3) at the cell's own density, velocity and
// temperature p, scaled so that sum M = n; -1 for an empty or cold cell. The
// reference has no such measure (it is the stopping rule of relaxation runs).
// One workgroup of kMomBS lanes.
__device__ __forceinline__ double l1_cell(const Grid& g, const double* __restrict__ f, const double* __restrict__ p,
                                          double mass, double* red) {
  const double n = p[DVM_P_DENS], u = p[DVM_P_U], v = p[DVM_P_V], w = p[DVM_P_W], T = p[DVM_P_TRTEMP];
  if (!(n > 0.0) || !(T > 0.0)) return -1.0;
  const double a = mass / T;
  double sm = 0.0;
  for (int l = threadIdx.x; l < g.N; l += kMomBS) {
    int i, j, k;
    node_ijk(g, l, i, j, k);
    const double dx = g.ax[0][i] - u, dy = g.ax[1][j] - v, dz = g.ax[2][k] - w;
    sm += exp(-(dx * dx + dy * dy + dz * dz) * a) * g.beta[0][i] * g.beta[1][j] * g.beta[2][k];
  }
  const double msum = block_sum<kMomBS>(sm, red);
  const double scale = n / msum;
  double e = 0.0;
  for (int l = threadIdx.x; l < g.N; l += kMomBS) {
    int i, j, k;
    node_ijk(g, l, i, j, k);
    const double dx = g.ax[0][i] - u, dy = g.ax[1][j] - v, dz = g.ax[2][k] - w;
    const double m = scale * exp(-(dx * dx + dy * dy + dz * dz) * a) * g.beta[0][i] * g.beta[1][j] * g.beta[2][k];
    e += fabs(f[l] - m);
  }
  e = block_sum<kMomBS>(e, red);
  return e / n;
}

__global__ void __launch_bounds__(kMomBS) k_l1(Grid g, const double* __restrict__ f,
                                               const double* __restrict__ props, double mass,
                                               double* __restrict__ out) {
  __shared__ double red[kMomBS / 64];
  const int cell = blockIdx.x;
  const double e = l1_cell(g, f + (size_t)cell * g.N, props + (size_t)cell * DVM_NPROP, mass, red);
  if (threadIdx.x == 0) out[cell] = e;
}

// Output moments of one cell, from f and its moments p (density, velocity,
// pressure): heat flux q_i = m/2 sum (c_i - u_i)|c - u|^2 f
// (physical_props.f90:1914-1986), viscous stress
// tau_ij = -m sum (c_i - u_i)(c_j - u_j) f, plus the pressure on the diagonal
// (:1988-2065), and entropy -sum_{f != 0} |f| log |f| (:1871-1912, C = 1).
// ext = {qx, qy, qz, txx, tyy, tzz, txy, txz, tyz, entropy}.
__global__ void __launch_bounds__(kMomBS) k_moments_ext(Grid g, const double* __restrict__ f,
                                                        const double* __restrict__ props, double mass,
                                                        double* __restrict__ ext) {
  __shared__ double red[kMomBS / 64];
  const int cell = blockIdx.x;
  const double* fc = f + (size_t)cell * g.N;
  const double* p = props + (size_t)cell * DVM_NPROP;
  const double dens = p[DVM_P_DENS], u = p[DVM_P_U], v = p[DVM_P_V], w = p[DVM_P_W];
  double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int l = threadIdx.x; l < g.N; l += kMomBS) {
    int i, j, k;
    node_ijk(g, l, i, j, k);
    const double cx = g.ax[0][i] - u, cy = g.ax[1][j] - v, cz = g.ax[2][k] - w;
    const double fl = fc[l];
    const double csq = cx * cx + cy * cy + cz * cz;
    acc[0] += cx * csq * fl;
    acc[1] += cy * csq * fl;
    acc[2] += cz * csq * fl;
    acc[3] -= cx * cx * fl;
    acc[4] -= cy * cy * fl;
    acc[5] -= cz * cz * fl;
    acc[6] -= cx * cy * fl;
    acc[7] -= cx * cz * fl;
    acc[8] -= cy * cz * fl;
    if (fl != 0.0) acc[9] -= fabs(fl) * log(fabs(fl));
  }
  double* out = ext + (size_t)cell * DVM_NEXT;
#pragma unroll 1
  for (int q = 0; q < 10; ++q) {
    const double t = block_sum<kMomBS>(acc[q], red);
    if (threadIdx.x == 0) {
      double o = t;
      if (q < 3) o = (dens == 0.0) ? 0.0 : t * (0.5 * mass);
      else if (q < 6) o = p[DVM_P_PRESSURE] + mass * t;
      else if (q < 9) o = mass * t;
      out[q] = o;
    }
  }
}

// C cells per workgroup of BS lanes (the last workgroup may hold fewer).
template <int C, int BS>
__global__ void __launch_bounds__(BS) k_moments(Grid g, const double* __restrict__ f,
                                                double* __restrict__ props, double mass, int ncell) {
  __shared__ double red[C][6][BS / 64];
  const int c0 = blockIdx.x * C;
  const int nc = (ncell - c0 < C) ? ncell - c0 : C;
  moments_sums<C, BS>(g, f + (size_t)c0 * g.N, (size_t)g.N, nc, red);
  if ((int)threadIdx.x < nc) {
    double tot[6];
    moments_total<BS>(red[threadIdx.x], tot);
    moments_finish(tot, props + (size_t)(c0 + threadIdx.x) * DVM_NPROP, mass);
  }
}

// ------------------------------------------------------------------- CDF
// Bucket brackets of the inverse-CDF search (variance_reduction.f90:468-493).
// The reference walks l upward over thresholds r*(0.001*total), so that
// right[r] = first l >= 1 with cumul[l] >= r*mx (N+1 if none), left[r] =
// right[r-1] - 1, left[1] = 0, right[1000] = N. For a non-decreasing CDF each
// threshold is crossed at exactly one node, found here in one pass over the
// nodes (each node tests the few thresholds in (cumul[l-1], cumul[l]]).
template <int BS>
__device__ __forceinline__ void build_buckets(const double* __restrict__ cumul, int N, int2* __restrict__ tab,
                                              int* __restrict__ rights) {
  for (int r = threadIdx.x; r < kNBucket; r += BS) rights[r] = N + 1;
  __syncthreads();
  const double mx = 0.001 * cumul[N];
  if (mx > 0.0) {
    const double inv = 1.0 / mx;
    for (int l = threadIdx.x + 1; l <= N; l += BS) {
      const double c0 = cumul[l - 1], c1 = cumul[l];
      if (!(c1 > c0)) continue;
      int r0 = (int)fmin(c0 * inv, 2000.0), r1 = (int)fmin(c1 * inv, 2000.0) + 1;
      r0 = r0 < 1 ? 1 : r0;
      r1 = r1 > kNBucket ? kNBucket : r1;
      for (int r = r0; r <= r1; ++r) {
        const double t = (double)r * mx;
        if (t > c0 && t <= c1) rights[r - 1] = l;
      }
    }
  } else {
    for (int r = threadIdx.x; r < kNBucket; r += BS) rights[r] = 1;
  }
  __syncthreads();
  for (int r = threadIdx.x + 1; r <= kNBucket; r += BS) {
    const int left = (r == 1) ? 0 : rights[r - 2] - 1;
    const int right = (r == kNBucket) ? N : rights[r - 1];
    tab[r - 1] = make_int2(left, right);
  }
  __syncthreads();
}

// build_buckets from the group ends G[gi] ~ cumul[min((gi + 1) V, N)] of the
// ceil(N / V) groups of V consecutive nodes, held in LDS by the scan as
// floats: a threshold in (cumul[gi V], cumul[min((gi + 1) V, N)]] is crossed
// at a node of group gi, so only the groups whose float bounds, widened by
// 2^-22 against the float rounding, admit a threshold are read back and given
// the node test of build_buckets, which decides on the FP64 values; for a
// non-decreasing CDF the brackets are those of build_buckets. A total below
// 1e-30 (where float ends lose their relative accuracy) takes build_buckets.
template <int BS, int V>
__device__ __forceinline__ void build_buckets_grouped(const double* __restrict__ cumul, const float* __restrict__ G,
                                                      int N, int2* __restrict__ tab, int* __restrict__ rights) {
  const double total = cumul[N];
  if (!(total >= 1e-30)) {
    build_buckets<BS>(cumul, N, tab, rights);
    return;
  }
  const int ng = (N + V - 1) / V;
  for (int r = threadIdx.x; r < kNBucket; r += BS) rights[r] = N + 1;
  __syncthreads();
  const double mx = 0.001 * total;
  const double inv = 1.0 / mx;
  constexpr double kLo = 1.0 - 0x1.0p-22, kHi = 1.0 + 0x1.0p-22;
  for (int gi = threadIdx.x; gi < ng; gi += BS) {
    const double g0 = (gi > 0) ? (double)G[gi - 1] * kLo : 0.0, g1 = (double)G[gi] * kHi;
    if (!(g1 > g0)) continue;
    int r0 = (int)fmin(g0 * inv, 2000.0), r1 = (int)fmin(g1 * inv, 2000.0) + 1;
    r0 = r0 < 1 ? 1 : r0;
    r1 = r1 > kNBucket ? kNBucket : r1;
    bool any = false;
    for (int r = r0; r <= r1; ++r) {
      const double th = (double)r * mx;
      any = any || (th > g0 && th <= g1);
    }
    if (!any) continue;
    const int lo = gi * V;
    const int nq = (N - lo < V) ? N - lo : V;
    double c[V + 1];
#pragma unroll
    for (int q = 0; q <= V; ++q) c[q] = (q <= nq) ? cumul[lo + q] : 0.0;
#pragma unroll
    for (int q = 1; q <= V; ++q) {
      const double c0 = c[q - 1], c1 = c[q];
      if (q > nq || !(c1 > c0)) continue;
      int s0 = (int)fmin(c0 * inv, 2000.0), s1 = (int)fmin(c1 * inv, 2000.0) + 1;
      s0 = s0 < 1 ? 1 : s0;
      s1 = s1 > kNBucket ? kNBucket : s1;
      for (int r = s0; r <= s1; ++r) {
        const double th = (double)r * mx;
        if (th > c0 && th <= c1) rights[r - 1] = lo + q;
      }
    }
  }
  __syncthreads();
  for (int r = threadIdx.x + 1; r <= kNBucket; r += BS) {
    const int left = (r == 1) ? 0 : rights[r - 2] - 1;
    const int right = (r == kNBucket) ? N : rights[r - 1];
    tab[r - 1] = make_int2(left, right);
  }
  __syncthreads();
}

// Position in the LDS chunk tile of node m of a chunk: lane m / V owns V
// consecutive nodes, rows padded to V + 1 doubles against bank conflicts.
template <int V>
__device__ __forceinline__ int tile_pos(int m) { return (m / V) * (V + 1) + (m % V); }

// Workgroup-wide scan of V consecutive values per lane: v[q] becomes
// carry + (prefix of everything before this lane in the chunk) + the lane's
// own running sum. Returns the chunk total.
template <int BS, int V>
__device__ __forceinline__ double chunk_scan(double (&v)[V], double* red, double carry) {
#pragma unroll
  for (int q = 1; q < V; ++q) v[q] = v[q - 1] + v[q];
  double tot;
  const double inc = block_scan<BS>(v[V - 1], red, tot);
  const double excl = carry + (inc - v[V - 1]);
#pragma unroll
  for (int q = 0; q < V; ++q) v[q] = excl + v[q];
  return tot;
}

// chunk_scan that also gives each lane the value its first node's predecessor
// took (the previous lane's last, or prev_last for lane 0) and the chunk's
// last value `last`, bit for bit, without another round of barriers: lane 0 of
// a wave repeats the arithmetic by which lane 63 of the wave before formed its
// last value, from that lane's inclusive prefix and local total (red[0], red[1]).
template <int BS, int V>
__device__ __forceinline__ double chunk_scan_prev(double (&v)[V], double (*red)[BS / 64], double carry,
                                                  double prev_last, double& before, double& last) {
#pragma unroll
  for (int q = 1; q < V; ++q) v[q] = v[q - 1] + v[q];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const double loc = v[V - 1];
  const double inc = wave_inclusive_scan(loc, lane);
  if (lane == 63) { red[0][wid] = inc; red[1][wid] = loc; }
  __syncthreads();
  double off = 0.0, tot = 0.0, offp = 0.0, offl = 0.0;
#pragma unroll
  for (int w = 0; w < BS / 64; ++w) {
    const double s = red[0][w];
    if (w < wid) off += s;
    if (w < wid - 1) offp += s;
    if (w < BS / 64 - 1) offl += s;
    tot += s;
  }
  // lane 63 of wave x: v[V-1] = (carry + ((red[0][x] + off_x) - red[1][x])) + red[1][x]
  const int px = wid > 0 ? wid - 1 : 0;
  const double lp = (carry + ((red[0][px] + offp) - red[1][px])) + red[1][px];
  const double ll = (carry + ((red[0][BS / 64 - 1] + offl) - red[1][BS / 64 - 1])) + red[1][BS / 64 - 1];
  __syncthreads();
  const double excl = carry + ((inc + off) - loc);
#pragma unroll
  for (int q = 0; q < V; ++q) v[q] = excl + v[q];
  const double up = __shfl_up(v[V - 1], 1, 64);
  before = (lane > 0) ? up : (wid > 0 ? lp : prev_last);
  last = ll;
  return tot;
}

// chunk_scan of two sequences in one round of barriers; the chunk totals are
// added to the carries.
template <int BS, int V>
__device__ __forceinline__ void chunk_scan2(double (&a)[V], double (&b)[V], double (*red)[BS / 64], double& ca,
                                            double& cb) {
#pragma unroll
  for (int q = 1; q < V; ++q) {
    a[q] = a[q - 1] + a[q];
    b[q] = b[q - 1] + b[q];
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const double ia = wave_inclusive_scan(a[V - 1], lane);
  const double ib = wave_inclusive_scan(b[V - 1], lane);
  if (lane == 63) { red[0][wid] = ia; red[1][wid] = ib; }
  __syncthreads();
  double oa = 0.0, ob = 0.0, ta = 0.0, tb = 0.0;
#pragma unroll
  for (int w = 0; w < BS / 64; ++w) {
    const double x = red[0][w], y = red[1][w];
    if (w < wid) { oa += x; ob += y; }
    ta += x;
    tb += y;
  }
  __syncthreads();
  const double ea = ca + ((ia + oa) - a[V - 1]), eb = cb + ((ib + ob) - b[V - 1]);
#pragma unroll
  for (int q = 0; q < V; ++q) {
    a[q] = ea + a[q];
    b[q] = eb + b[q];
  }
  ca = ca + ta;
  cb = cb + tb;
}

// compute_cdf_vr (variance_reduction.f90:234-496) for the like-species pair,
// plus compute_num_coll_pairs (collision_utilities.f90:106-140) and the
// depletion factors (variance_reduction.f90:598-604,679,831,976).
// One workgroup of BS lanes per cell, one pass over the nodes in chunks of
// BS*V: lane t owns V consecutive nodes for the scans; loads go coalesced
// through an LDS tile, and so do stores. On an equilibrium rebuild the pass
// computes f_eq*beta^3, its prefix eq_df, the differences eqv and both
// cumulative distributions (eqv and |f - eqv|), so f is read once and eqv is
// not read back; otherwise it reads f and the cached eqv. With gsum (LDS,
// 2 ceil(N/V) floats) the lanes' group ends are kept for the bucket tables,
// which then read back only the groups a threshold falls in.
template <int BS, int V>
struct CdfShared {
  double red[2][BS / 64];
  double tile[BS * (V + 1)];
  int rights[kNBucket];
  int s_new;
};

// Sign bits of the deviation f - f_eq at V consecutive nodes from l0 (bit q
// for node l0 + q, set when the difference has its sign bit set, so -0 reads
// as negative like dsgn): one byte per 8 nodes, byte l / 8 bit l % 8. V = 8:
// each lane writes its byte; V = 4: lane pairs combine their halves.
template <int V>
__device__ __forceinline__ void store_sign_bits(const double (&dev)[V], int l0, int N,
                                                unsigned char* __restrict__ sgn) {
  static_assert(V == 8 || V == 4, "sign bytes need 4 or 8 nodes per lane");
  unsigned b = 0;
#pragma unroll
  for (int q = 0; q < V; ++q) b |= (signbit(dev[q]) ? 1u : 0u) << q;
  if constexpr (V == 8) {
    if (l0 < N) sgn[l0 >> 3] = (unsigned char)b;
  } else {
    const unsigned hi = (unsigned)__shfl_xor((int)b, 1, 64);
    if (!(threadIdx.x & 1) && l0 < N) sgn[l0 >> 3] = (unsigned char)(b | (hi << 4));
  }
}

// Coalesced store of a chunk held as V consecutive values per lane, through
// the tile (whose last reads precede a barrier the caller has passed).
template <int BS, int V>
__device__ __forceinline__ void tile_store(const double (&x)[V], double* tile, double* __restrict__ dst, int base,
                                           int N) {
  const int own = threadIdx.x * (V + 1);
#pragma unroll
  for (int q = 0; q < V; ++q) tile[own + q] = x[q];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < V; ++q) {
    const int m = q * BS + threadIdx.x, l = base + m;
    if (l < N) dst[l] = tile[tile_pos<V>(m)];
  }
}

// f_eq*beta^3 of V consecutive nodes from l0 (zero beyond N): no loads, each
// lane steps (i, j, k) along the node order (dist_func.f90:317-344).
template <int V>
__device__ __forceinline__ void eq_values(const Grid& g, int l0, double coeff, double u, double v, double w,
                                          double temp, double mass, double (&pv)[V]) {
  const int N = g.N, nx = g.n[0], ny = g.n[1];
  int i, j, k;
  node_ijk(g, l0 < N ? l0 : 0, i, j, k);
#pragma unroll
  for (int q = 0; q < V; ++q) {
    double val = 0.0;
    if (l0 + q < N) {
      const double x = g.ax[0][i], y = g.ax[1][j], z = g.ax[2][k];
      const double beta3 = g.beta[0][i] * g.beta[1][j] * g.beta[2][k];
      double phi_eq = 0.0;
      if (temp > 0.0) {
        double cx = (x - u); cx = cx * cx;
        double cy = (y - v); cy = cy * cy;
        double cz = (z - w); cz = cz * cz;
        const double csq = cx + cy + cz;
        phi_eq = coeff * exp(-csq * mass / temp);
      }
      val = phi_eq * beta3;
    }
    pv[q] = val;
    if (++i == nx) { i = 0; if (++j == ny) { j = 0; ++k; } }
  }
}

// One cell: fc its distribution, p its moments, ev / ec / tab2 / par / cnt4 /
// fac3 its slices of eqv, cdf ([2][N+1]), tabs ([2][1000]), eqpar, counts and
// factors.
//   kFused: an equilibrium rebuild shares the one pass over f (many cells in
//           flight, where the pass is bandwidth-bound); otherwise (few cells,
//           where every barrier of a chunk is on the critical path) the
//           equilibrium is built by a pass of its own with no loads, and the
//           |f - f_eq| pass reads it back.
//   eq_mode: 1 -- f_eq is stored on a rebuild and read back on a cache hit;
//           0 -- no collision kernel reads f_eq: the fused pass does not
//           store it, and a cache hit computes it again from the cached
//           parameters, bit for bit as the rebuild did (the few-cell pass,
//           whose |f - f_eq| pass reads f_eq back, computes and stores it
//           on a hit too); 2 -- computed again on a hit and stored either
//           way (f_eq current for a reader).
template <int BS, int V, bool kFused>
__device__ __forceinline__ void cdf_cell(const Grid& g, const CdfParams& cp, const double* __restrict__ fc,
                                         const double* __restrict__ p, double* __restrict__ ev,
                                         double* __restrict__ ec, int2* __restrict__ tab2, double* __restrict__ par,
                                         int* __restrict__ cnt4, double* __restrict__ fac3, CdfShared<BS, V>& sh,
                                         float* gsum, unsigned char* __restrict__ sgn, int eq_mode = 1) {
  constexpr int CH = BS * V;
  int* rights = sh.rights;
  double* tile = sh.tile;
  const int N = g.N;
  const int ng = (N + V - 1) / V;
  double* nc = ec + (N + 1);
  const double dens = p[DVM_P_DENS], u = p[DVM_P_U], v = p[DVM_P_V], w = p[DVM_P_W];
  double temp = p[DVM_P_TEMP];
  const double mass = cp.mass;
  const bool fmc = cp.fmc != 0;
  if (threadIdx.x == 0) {
    const double tol = cp.cdf_tolerance;
    // the equilibrium cache test: w is not compared; a negative u or v makes
    // its test always true
    sh.s_new = (!cp.fmc && (fabs(dens - par[0]) > tol * dens || fabs(u - par[1]) > tol * u ||
                            fabs(v - par[2]) > tol * v || fabs(temp - par[4]) > tol * temp)) ? 1 : 0;
  }
  __syncthreads();
  const bool new_eq = sh.s_new != 0;
  __syncthreads();
  // the equilibrium is computed on a rebuild, and on a cache hit unless f_eq
  // is read back (eq_mode 1), then from the cached parameters
  const bool recompute = !new_eq && !fmc && eq_mode != 1;
  const bool fuse = kFused && (new_eq || recompute);
  const bool store_ev = fuse && eq_mode != 0;
  const double edens = new_eq ? dens : par[0], eu = new_eq ? u : par[1], evv = new_eq ? v : par[2];
  const double ew = new_eq ? w : par[3], etemp = new_eq ? temp : par[4];
  double coeff = 0.0;
  if ((new_eq || recompute) && edens > 0.0)
    coeff = edens * sqrt(mass * mass * mass / (kPiRef * kPiRef * kPiRef * etemp * etemp * etemp));
  const int own = threadIdx.x * (V + 1);
  // carries: eq_df prefix, eqv prefix, |f - eqv| prefix; last eq_df of the previous chunk
  double cpv = 0.0, cec = 0.0, cnc = 0.0, prev_last = 0.0;
  if (threadIdx.x == 0) {
    nc[0] = 0.0;
    if (new_eq) ec[0] = 0.0;
  }
  if (!kFused && (new_eq || recompute)) {
    // the equilibrium's own pass: eq_df = prefix of f_eq*beta^3,
    // eqv(l) = eq_df(l) - eq_df(l-1), eqc = prefix of eqv (the equilibrium
    // CDF and its table only on a rebuild)
    for (int base = 0; base < N; base += CH) {
      double pv[V], e[V];
      eq_values<V>(g, base + (int)threadIdx.x * V, coeff, eu, evv, ew, etemp, mass, pv);
      double before, last;
      cpv = cpv + chunk_scan_prev<BS, V>(pv, sh.red, cpv, prev_last, before, last);
      prev_last = last;
#pragma unroll
      for (int q = 0; q < V; ++q) e[q] = pv[q] - (q == 0 ? before : pv[q - 1]);
      tile_store<BS, V>(e, tile, ev, base, N);
      if (new_eq) {
        cec = cec + chunk_scan<BS, V>(e, sh.red[0], cec);
        tile_store<BS, V>(e, tile, ec + 1, base, N);
      }
      // the next chunk writes the tile only after its first scan's barriers
    }
    if (new_eq) {
      build_buckets<BS>(ec, N, tab2, rights);  // its barriers also publish ev
      if (threadIdx.x == 0) {
        par[0] = dens; par[1] = u; par[2] = v; par[3] = w; par[4] = temp;
      }
    } else {
      __syncthreads();  // publishes ev (and ends the tile's reads)
    }
  }
  // f (and, unless the equilibrium is built in this pass, eqv) of the next
  // chunk are loaded while this one is scanned
  double nf[V], ne[V];
#pragma unroll
  for (int q = 0; q < V; ++q) {
    const int l = q * BS + threadIdx.x;
    nf[q] = (l < N) ? fc[l] : 0.0;
    ne[q] = (l < N && !fuse && !fmc) ? ev[l] : 0.0;
  }
  for (int base = 0; base < N; base += CH) {
    const int l0 = base + (int)threadIdx.x * V;
    double fv[V], e[V], d[V];
    if (fuse) {
#pragma unroll
      for (int q = 0; q < V; ++q) tile[tile_pos<V>(q * BS + threadIdx.x)] = nf[q];
    } else {
      // the signed deviation (its magnitude is the CDF's term, its sign bit
      // the deviation sign)
#pragma unroll
      for (int q = 0; q < V; ++q) tile[tile_pos<V>(q * BS + threadIdx.x)] = nf[q] - ne[q];
    }
#pragma unroll
    for (int q = 0; q < V; ++q) {
      const int l = base + CH + q * BS + threadIdx.x;
      if (l < N) {
        nf[q] = fc[l];
        ne[q] = (fuse || fmc) ? 0.0 : ev[l];
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < V; ++q) fv[q] = tile[own + q];
    if (fuse) {
      // eq_df = prefix of f_eq*beta^3; eqv(l) = eq_df(l) - eq_df(l-1); eqc = prefix of eqv
      double pv[V];
      eq_values<V>(g, l0, coeff, eu, evv, ew, etemp, mass, pv);
      // eq_df of the node before this lane's first: the previous lane's last
      double before, last;
      cpv = cpv + chunk_scan_prev<BS, V>(pv, sh.red, cpv, prev_last, before, last);
      prev_last = last;
      double dv[V];
#pragma unroll
      for (int q = 0; q < V; ++q) {
        e[q] = pv[q] - (q == 0 ? before : pv[q - 1]);
        dv[q] = (l0 + q < N) ? fv[q] - e[q] : 0.0;
        d[q] = fabs(dv[q]);
      }
      if (sgn) store_sign_bits<V>(dv, l0, N, sgn);
      if (store_ev) tile_store<BS, V>(e, tile, ev, base, N);
      if (new_eq) {
        chunk_scan2<BS, V>(e, d, sh.red, cec, cnc);
        tile_store<BS, V>(e, tile, ec + 1, base, N);
        __syncthreads();
      } else {
        // a cache hit keeps the stored equilibrium CDF
        cnc = cnc + chunk_scan<BS, V>(d, sh.red[0], cnc);
      }
      tile_store<BS, V>(d, tile, nc + 1, base, N);
      if (gsum && l0 < N) {
        if (new_eq) gsum[l0 / V] = (float)e[V - 1];
        gsum[ng + l0 / V] = (float)d[V - 1];
      }
    } else {
      // |f - f_eq| prefix with the cached eqv; |f| prefix for the full Monte
      // Carlo kernel (full_mc.f90:110-180), which takes f_eq = 0
      if (sgn) store_sign_bits<V>(fv, l0, N, sgn);
#pragma unroll
      for (int q = 0; q < V; ++q) d[q] = fabs(fv[q]);
      cnc = cnc + chunk_scan<BS, V>(d, sh.red[0], cnc);
      tile_store<BS, V>(d, tile, nc + 1, base, N);
      if (gsum && l0 < N) gsum[ng + l0 / V] = (float)d[V - 1];
    }
    __syncthreads();  // the tile's reads end before the next chunk writes it
  }
  if (fuse && new_eq) {
    if (gsum) build_buckets_grouped<BS, V>(ec, gsum, N, tab2, rights);
    else build_buckets<BS>(ec, N, tab2, rights);
    if (threadIdx.x == 0) {
      par[0] = dens; par[1] = u; par[2] = v; par[3] = w; par[4] = temp;
    }
  }
  if (gsum) build_buckets_grouped<BS, V>(nc, gsum + ng, N, tab2 + kNBucket, rights);
  else build_buckets<BS>(nc, N, tab2 + kNBucket, rights);
  if (fmc) {
    // full_monte_carlo_kernel (full_mc.f90:240-317): abs_dens = dens - 2 neg_dens
    // with neg_dens = (dens - sum |f|) / 2; compute_num_coll_pairs with both
    // density ratios abs_dens / dens; factor halved for like species. All
    // samples are integral-3 samples (both partners from the |f| CDF).
    if (threadIdx.x == 0) {
      const double neg_dens = 0.5 * (dens - nc[N]);
      const double abs_dens = dens - 2.0 * neg_dens;
      double num;
      if (cp.sample_mode == DVM_SAMPLES_EXPLICIT) {
        num = (double)cp.samples_per_cell;
      } else {
        const double ratio = abs_dens / dens;
        const double ct = pow(0.5 * (p[DVM_P_TEMP] + p[DVM_P_TEMP]), 2.0 / 3.0);
        num = round((0.5 * cp.dt * ct * ratio * ratio) / (cp.kn * cp.rms_sq * cp.beta3));
      }
      if (!(num > 0.0) || !(abs_dens > 0.0)) num = 0.0;
      if (num > (double)kMaxCellSamples) num = (double)kMaxCellSamples;
      const int n = (int)num;
      cnt4[0] = 0; cnt4[1] = 0; cnt4[2] = n; cnt4[3] = n;
      fac3[0] = 0.0; fac3[1] = 0.0;
      fac3[2] = (n > 0) ? 0.5 * ((cp.dt / cp.kn) * cp.sigma * abs_dens * abs_dens / (double)n) : 0.0;
    }
    return;
  }
  if (threadIdx.x == 0) {
    const double neq_dens = nc[N], eq_dens = ec[N];
    const double h = 0.5 * neq_dens / dens;
    const double dA[3] = {1.0, h, h}, dB[3] = {h, 1.0, h};
    const double wA[3] = {eq_dens, neq_dens, neq_dens}, wB[3] = {neq_dens, eq_dens, neq_dens};
    const double factor_coeff = cp.dt * 0.5 / cp.kn;
    const double tot = dA[0] * dB[0] + dA[1] * dB[1] + dA[2] * dB[2];
    int sum = 0;
    for (int t = 0; t < 3; ++t) {
      double num;
      if (cp.sample_mode == DVM_SAMPLES_EXPLICIT) {
        num = (tot > 0.0) ? round((double)cp.samples_per_cell * (dA[t] * dB[t]) / tot) : 0.0;
      } else {
        const double ct = pow(0.5 * (p[DVM_P_TEMP] + p[DVM_P_TEMP]), 2.0 / 3.0);
        num = round((0.5 * cp.dt * ct * dA[t] * dB[t]) / (cp.kn * cp.rms_sq * cp.beta3));
      }
      if (!(num > 0.0)) num = 0.0;
      if (num > (double)kMaxIntegralSamples) num = (double)kMaxIntegralSamples;
      const int n = (int)num;
      cnt4[t] = n;
      sum += n;
      fac3[t] = (n > 0) ? factor_coeff * cp.sigma * wA[t] * wB[t] / (double)n : 0.0;
    }
    cnt4[3] = sum;
  }
}

template <int BS, int V, bool kFused>
__global__ void __launch_bounds__(BS) k_cdf(Grid g, CdfParams cp, const double* __restrict__ f,
                                                const double* __restrict__ props, double* __restrict__ eqv,
                                                double* __restrict__ cdf, int2* __restrict__ tabs,
                                                double* __restrict__ eqpar,
                                                int* __restrict__ counts, double* __restrict__ factors,
                                                int sum_cap, unsigned char* __restrict__ sgn, int eq_mode) {
  __shared__ CdfShared<BS, V> sh;
  extern __shared__ float gdyn[];  // sum_cap floats: the scan's group ends
  const int cell = blockIdx.x;
  const int N = g.N;
  float* gsum = (2 * ((N + V - 1) / V) <= sum_cap) ? gdyn : nullptr;
  cdf_cell<BS, V, kFused>(g, cp, f + (size_t)cell * N, props + (size_t)cell * DVM_NPROP, eqv + (size_t)cell * N,
           cdf + (size_t)cell * 2 * (N + 1), tabs + (size_t)cell * 2 * kNBucket, eqpar + (size_t)cell * 5,
           counts + cell * 4, factors + cell * 3, sh, gsum, sgn ? sgn + (size_t)cell * 2 * (N + 1) : nullptr,
           eq_mode);
}

// --------------------------------------------------------------- collide
struct SampleRec {
  int lA, lB, skipped, st1[7], st2[7];
  double kd1, kd2, m1[7], m2[7];
};

// FP32 collision arithmetic uses the hardware transcendental, reciprocal and
// square-root instructions (v_sin_f32 / v_cos_f32 / v_rcp_f32 / v_sqrt_f32,
// ~1e-6 relative or better); FP64
// uses the correctly rounded library functions the reference's results need.
template <class R>
__device__ __forceinline__ void sin_cos(R x, R& s, R& c) {
  if constexpr (sizeof(R) == 4) { s = __sinf(x); c = __cosf(x); } else sincos(x, &s, &c);
}
template <class R>
__device__ __forceinline__ R rsqrt_free_sqrt(R x) {
  if constexpr (sizeof(R) == 4) return __builtin_amdgcn_sqrtf(x); else return sqrt(x);
}
template <class R>
__device__ __forceinline__ R rdiv(R a, R b) {
  return quot<R>(a, b);
}
template <class R>
__device__ __forceinline__ R rpow(R x, R y) {
  if constexpr (sizeof(R) == 4) return powf(x, y); else return pow(x, y);
}

// Uniforms of one sample, in the reference's draw order: the two partner
// uniforms (pick_collision_partner), then (phi, theta) of each replenish
// location r = 1..NRepl (pick_random_velocity_on_sphere). q is the sample's
// index in its cell over the three integrals, (t, s) its integral and index
// in it.
// PhiloxSrc: counter-based Philox4x32-10, one block (s, t + 1 | r << 2, cell,
// ntime) per location r, the first one also carrying the partner uniforms
// (words 0-1) before (phi, theta) (words 2-3), each uniform word * 2^-32.
// pick_batch and vr_sample spell these steps out for it, with the seed of
// CollParams.
struct PhiloxSrc {
  uint64_t seed;
};

// StreamSrc: an explicit list of uniforms (the parity mode of
// dvm_set_uniform_stream): sample q of a cell reads the `per` = 2 + 2 NRepl
// values u[base[cell] + q * per ...] in the order above.
struct StreamSrc {
  const double* u;
  const long long* base;
  int per;
  __device__ __forceinline__ void pair(int cell, int ntime, int q, int t, int s, double& uA, double& uB,
                                       uint32_t& wa, uint32_t& wb) const {
    const double* p = u + base[cell] + (long long)q * per;
    uA = p[0];
    uB = p[1];
    wa = 0u;
    wb = 0u;
  }
  __device__ __forceinline__ void angles(int cell, int ntime, int q, int t, int s, int r, uint32_t wa, uint32_t wb,
                                         double& uphi, double& uth) const {
    const double* p = u + base[cell] + (long long)q * per + 2 * r;
    uphi = p[0];
    uth = p[1];
  }
};

// Partner picks of K samples (pick_collision_partner, collision_utilities.f90:
// 142-186, with referenced_search, sort_and_search.f90:40-86). The 2K
// inverse-CDF searches advance in lockstep so that their dependent loads from
// the cell's CDFs (L2/MALL resident, not LDS) overlap; each search follows the
// reference's bucket bracket and bisection exactly.
template <int K, class Src>
__device__ __forceinline__ void pick_batch(const CollParams& cp, const CellView& cv, int N, int cell, int ntime,
                                           const int* q, const int* t, const int* s, const bool* valid, int* lA,
                                           int* lB, uint32_t* wphi, uint32_t* wth, const Src& src) {
  constexpr int M = 2 * K;
  // the cell's two CDFs are adjacent: eq at cv.eqc, deviation at cv.eqc + N + 1
  int off[M], left[M], right[M], idx[M];
  double adj[M];
  bool active[M];
  int2 lr[M];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double uA, uB;
    if constexpr (std::is_same<Src, PhiloxSrc>::value) {
      uint32_t w4[4] = {0u, 0u, 0u, 0u};
      if (valid[k]) philox_block(cp.seed, (uint32_t)s[k], (uint32_t)(t[k] + 1), (uint32_t)cell, (uint32_t)ntime, w4);
      uA = u32_to_uniform(w4[0]);
      uB = u32_to_uniform(w4[1]);
      wphi[k] = w4[2];
      wth[k] = w4[3];
    } else {
      uA = 0.0;
      uB = 0.0;
      wphi[k] = 0u;
      wth[k] = 0u;
      if (valid[k]) src.pair(cell, ntime, q[k], t[k], s[k], uA, uB, wphi[k], wth[k]);
    }
    const bool eqA = t[k] == 0, eqB = t[k] == 1;
    off[2 * k] = eqA ? 0 : N + 1;
    off[2 * k + 1] = eqB ? 0 : N + 1;
    adj[2 * k] = uA;
    adj[2 * k + 1] = uB;
    int bA = (int)ceil(uA * 1000.0), bB = (int)ceil(uB * 1000.0);
    bA = bA < 1 ? 1 : (bA > kNBucket ? kNBucket : bA);
    bB = bB < 1 ? 1 : (bB > kNBucket ? kNBucket : bB);
    lr[2 * k] = cv.eqtab[(eqA ? 0 : kNBucket) + bA - 1];
    lr[2 * k + 1] = cv.eqtab[(eqB ? 0 : kNBucket) + bB - 1];
  }
  // st: the search's index while it runs, its bitwise complement once it has
  // stopped (the node found, the bracket closed, or nothing to search)
  int st[M];
#pragma unroll
  for (int j = 0; j < M; ++j) {
    const double u = adj[j];
    left[j] = lr[j].x;
    right[j] = lr[j].y;
    idx[j] = left[j] + (right[j] - left[j]) / 2;
    active[j] = valid[j >> 1] && !(u < kDoubleTol) && left[j] != right[j];
    if (valid[j >> 1] && u < kDoubleTol) idx[j] = 0;
    st[j] = active[j] ? idx[j] : ~idx[j];
    adj[j] = u * (off[j] == 0 ? cv.eq_total : cv.neq_total);
  }
  // Bisection without branches per search: a stopped search loads node 0 of
  // its CDF pair (one line for all such lanes). 0 <= left <= right here, so
  // (left + right) >> 1 is the reference's left + (right - left) / 2.
  for (int it = 0; it < kSearchCap; ++it) {
    double c0[M], c1[M];
#pragma unroll
    for (int j = 0; j < M; ++j) {
      const unsigned e = (unsigned)(off[j] + max(st[j], 0));
      const double* a = reinterpret_cast<const double*>(reinterpret_cast<const char*>(cv.eqc) + e * 8u);
      c0[j] = a[0];
      c1[j] = a[1];
    }
    bool any = false;
#pragma unroll
    for (int j = 0; j < M; ++j) {
      // each select is driven by compares of this step only (no boolean
      // carried between steps), which keeps the masks in scalar registers
      const int id = st[j];
      const bool act = id >= 0;
      const bool gt = adj[j] > c0[j];
      const bool le = adj[j] <= c1[j];
      left[j] = (act && gt && !le) ? id : left[j];
      right[j] = (act && !gt) ? id : right[j];
      const int nid = (int)((unsigned)(left[j] + right[j]) >> 1);
      st[j] = act ? ((gt && le) ? ~id : nid) : id;
      any = any || (act && !(gt && le));
    }
    if (!any) break;
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    lA[k] = st[2 * k] < 0 ? ~st[2 * k] : st[2 * k];
    lB[k] = st[2 * k + 1] < 0 ? ~st[2 * k + 1] : st[2 * k + 1];
  }
}

// One VR sample: variance_reduction.f90:686-816 / 838-961 / 983-1116 with
// vr_depletion_routine (:1127-1237) and the elastic branch of
// replenish_collision (replenishing_collisions.f90:134-354).
// R is the collision arithmetic (double, or float for the FP32-collision path);
// the remap runs in RM (double, or float with kRemap32 on the FP32 axes ax),
// with kUni its nearest nodes by rounding on equally spaced axes; the partner
// picks and the deviation signs are FP64 always, and every value scattered is
// FP64.
// scatter(node, value) receives the two depletions and every stencil point.
// The sign of the deviation at each deviation partner (+1 for an equilibrium
// partner) is read from the cell's f and f_eq at the start and first used
// after the velocity arithmetic, which covers the loads' latency.
template <bool kRecord, class R, bool kUni, bool kRemap32, class Src, class Scatter,
          class RM = typename std::conditional<kRemap32, float, double>::type>
__device__ __forceinline__ int vr_sample(const Grid& g, const Axes<RM>& ax, const CollParams& cp, const CellView& cv,
                                         int cell, int ntime, int q, int t, int s, double factor, int lA, int lB,
                                         uint32_t wphi, uint32_t wth, const Src& src, Scatter&& scatter,
                                         SampleRec* rec) {
  const int N = g.N, nx = g.n[0], nxy = g.nxy;
  if ((unsigned)lA >= (unsigned)N || (unsigned)lB >= (unsigned)N) return 1;  // unreachable for finite CDFs
  const double fa = cv.f[lA], ea = cv.eqv[lA], fb = cv.f[lB], eb = cv.eqv[lB];
  int iA, jA, kA, iB, jB, kB;
  node_ijk(g, lA, iA, jA, kA);
  node_ijk(g, lB, iB, jB, kB);
  const R x1 = (R)ax.a[0][iA], y1 = (R)ax.a[1][jA], z1 = (R)ax.a[2][kA];
  const R x2 = (R)ax.a[0][iB], y2 = (R)ax.a[1][jB], z2 = (R)ax.a[2][kB];
  if (kRecord) { rec->lA = lA; rec->lB = lB; rec->skipped = 0; }
  const R gx = x2 - x1, gy = y2 - y1, gz = z2 - z1;
  const R gg = rsqrt_free_sqrt<R>(gx * gx + gy * gy + gz * gz);
  if (gg < (R)kDoubleTol) {
    if (kRecord) rec->skipped = 1;
    return 0;
  }
  const R g_sigma = cp.hard_sphere ? gg : rpow<R>(gg, (R)cp.vhs_exponent);
  const R mf = (R)(cp.mass / (cp.mass + cp.mass));
  const R vc0 = mf * x1 + mf * x2, vc1 = mf * y1 + mf * y2, vc2 = mf * z1 + mf * z2;
  const R gx1 = fabs(x1 - vc0), gy1 = fabs(y1 - vc1), gz1 = fabs(z1 - vc2);
  const R gx2 = fabs(x2 - vc0), gy2 = fabs(y2 - vc1), gz2 = fabs(z2 - vc2);
  const R ur = gx1 + gx2, vr = gy1 + gy2, wr = gz1 + gz2;
  const R gp1 = rsqrt_free_sqrt<R>(gx1 * gx1 + gy1 * gy1 + gz1 * gz1);
  const R gp2 = rsqrt_free_sqrt<R>(gx2 * gx2 + gy2 * gy2 + gz2 * gz2);
  const R gr1 = rdiv<R>(gp1, gg), gr2 = rdiv<R>(gp2, gg);
  const R signA = (R)((t == 0) ? 1.0 : dsgn(fa - ea));
  const R signB = (R)((t == 1) ? 1.0 : dsgn(fb - eb));
  const R sumA = signA, sumB = signB;
  const R fac = (R)factor;
  const R dep1 = sumB * (R)1 * fac * g_sigma;
  const R dep2 = sumA * (R)1 * fac * g_sigma;
  const R kd1 = sumA * dep1, kd2 = sumB * dep2;
  if (kRecord) { rec->kd1 = (double)kd1; rec->kd2 = (double)kd2; }
  if (!kRecord) {
    scatter(iA, jA, kA, lA, -(double)kd1);
    scatter(iB, jB, kB, lB, -(double)kd2);
  }
  // kd / NRepl; NRepl = 1 in the reference's decks, where the quotient is kd
  // exactly (FP32: times the FP32 reciprocal of NRepl)
  R kr1 = kd1, kr2 = kd2;
  if (cp.num_repl != 1) {
    if constexpr (sizeof(R) == 4) {
      kr1 = kd1 * (R)cp.inv_repl;
      kr2 = kd2 * (R)cp.inv_repl;
    } else {
      kr1 = kd1 / (R)cp.num_repl;
      kr2 = kd2 / (R)cp.num_repl;
    }
  }
  int bad = 0;
  for (int r = 1; r <= cp.num_repl; ++r) {
    double uphi, uth;
    if constexpr (std::is_same<Src, PhiloxSrc>::value) {
      uphi = u32_to_uniform(wphi);
      uth = u32_to_uniform(wth);
      if (r > 1) {
        uint32_t w4[4];
        philox_block(cp.seed, (uint32_t)s, (uint32_t)((t + 1) | (r << 2)), (uint32_t)cell, (uint32_t)ntime, w4);
        uphi = u32_to_uniform(w4[0]);
        uth = u32_to_uniform(w4[1]);
      }
    } else {
      src.angles(cell, ntime, q, t, s, r, wphi, wth, uphi, uth);
    }
    // pick_random_velocity_on_sphere (collision_utilities.f90:284-352)
    const R ph = (R)2 * (R)kPiRef * (R)uphi;
    R sph, cph;
    sin_cos<R>(ph, sph, cph);
    R rx, ry, rz;
    if (cp.vss) {
      const R ct = (R)2 * rpow<R>((R)uth, (R)cp.ralpha) - (R)1;
      const R stt = rsqrt_free_sqrt<R>((R)1 - ct * ct);
      const R den = rsqrt_free_sqrt<R>(vr * vr + wr * wr);
      if (den > (R)kDoubleTol * gg) {
        rx = ct * ur + stt * sph * den;
        ry = ct * vr + rdiv<R>(stt * (gg * wr * cph - ur * vr * sph), den);
        rz = ct * wr - rdiv<R>(stt * (gg * vr * cph + ur * wr * sph), den);
      } else {
        rx = ct * ur;
        ry = stt * cph * ur;
        rz = stt * sph * ur;
      }
    } else {
      const R ct = (R)2 * (R)uth - (R)1;
      const R stt = rsqrt_free_sqrt<R>((R)1 - ct * ct);
      rx = gg * stt * cph;
      ry = gg * stt * sph;
      rz = gg * ct;
    }
    // post-collision velocities, remapped in RM
    // (remapping.f90:55-394 via apply_phi_replenishment :356-434)
    for (int side = 0; side < 2; ++side) {
      const R pu = (side == 0) ? vc0 + gr1 * rx : vc0 + -(gr2 * rx);
      const R pv = (side == 0) ? vc1 + gr1 * ry : vc1 + -(gr2 * ry);
      const R pw = (side == 0) ? vc2 + gr1 * rz : vc2 + -(gr2 * rz);
      if constexpr (kUni && !kRecord && sizeof(R) == 4) {
        // equally spaced axes, FP32 path: the closed-form weight solve away
        // from the axis ends (the general remap below otherwise)
        if (remap_uniform_interior<RM>(ax, pu, pv, pw, (RM)(side == 0 ? kr1 : kr2), nx, nxy, scatter)) continue;
      }
      StencilT<RM> st;
      remap_t<RM, kUni>(ax, (RM)pu, (RM)pv, (RM)pw, (RM)1, st);
      if (kRecord && r == 1) {
        int* o = side == 0 ? rec->st1 : rec->st2;
        double* m = side == 0 ? rec->m1 : rec->m2;
        o[0] = st.oi; o[1] = st.oj; o[2] = st.ok; o[3] = st.a; o[4] = st.b; o[5] = st.c; o[6] = st.flag;
#pragma unroll
        for (int q = 0; q < 7; ++q) m[q] = (double)st.m[q];
      }
      if (st.flag == 0) { bad = 1; continue; }
      if (!kRecord) {
        const RM kr = (RM)(side == 0 ? kr1 : kr2);
#pragma unroll
        for (int q = 0; q < 7; ++q) {
          if (q < 4 || (st.flag & (1 << (q - 4)))) {
            int si, sj, sk;
            stencil_ijk(st, q, si, sj, sk);
            scatter(si, sj, sk, si + nx * sj + nxy * sk, (double)(kr * st.m[q]));
          }
        }
      }
    }
  }
  return bad;
}

// Stage the three axes into LDS and point the grid at them.
__device__ __forceinline__ void stage_axes(Grid& g, double* sh) {
  double* p = sh;
  for (int d = 0; d < 3; ++d) {
    for (int q = threadIdx.x; q < g.n[d]; q += blockDim.x) p[q] = g.ax[d][q];
    g.ax[d] = p;
    p += g.n[d];
  }
}

// The axes of a collision workgroup in its arithmetic R, staged into LDS at
// sh (the caller's barrier publishes them): FP64 -- g then points at the
// staged copy; FP32 -- a float copy, and g keeps the global FP64 axes, which
// the FP64 conservation correction reads.
template <class R>
__device__ __forceinline__ Axes<R> stage_collision_axes(Grid& g, double* sh) {
  if constexpr (sizeof(R) == 8) {
    stage_axes(g, sh);
    return grid_axes(g);
  } else {
    Axes<float> x;
    float* p = reinterpret_cast<float*>(sh);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      for (int q = threadIdx.x; q < g.n[d]; q += blockDim.x) p[q] = (float)g.ax[d][q];
      x.a[d] = p;
      x.n[d] = g.n[d];
      x.vmin[d] = (float)g.vmin[d];
      x.vmax[d] = (float)g.vmax[d];
      x.factor[d] = (float)g.factor[d];
      p += g.n[d];
    }
    set_uniform_consts(x);
    return x;
  }
}

// Symmetric positive-definite 5x5 solve (Cholesky) for the conservation
// correction; returns false when the matrix is not numerically SPD.
__device__ bool solve5(double* M /* 5x5 row-major, overwritten */, double* b /* 5, overwritten */) {
  for (int j = 0; j < 5; ++j) {
    double d = M[j * 5 + j];
    for (int k = 0; k < j; ++k) d -= M[j * 5 + k] * M[j * 5 + k];
    if (!(d > 0.0)) return false;
    d = sqrt(d);
    M[j * 5 + j] = d;
    for (int i = j + 1; i < 5; ++i) {
      double v = M[i * 5 + j];
      for (int k = 0; k < j; ++k) v -= M[i * 5 + k] * M[j * 5 + k];
      M[i * 5 + j] = v / d;
    }
  }
  for (int i = 0; i < 5; ++i) {
    double v = b[i];
    for (int k = 0; k < i; ++k) v -= M[i * 5 + k] * b[k];
    b[i] = v / M[i * 5 + i];
  }
  for (int i = 4; i >= 0; --i) {
    double v = b[i];
    for (int k = i + 1; k < 5; ++k) v -= M[k * 5 + i] * b[k];
    b[i] = v / M[i * 5 + i];
  }
  return true;
}

// Collision phase of one cell per workgroup of BS lanes. Returns device error
// bits (kErrRemap for a remap without an exterior point).
//   kLds: the cell's increments accumulate in LDS (dynamic shared memory after
//         the axes); otherwise fd (a copy of f) takes global FP64 atomics.
//   kF32: FP32 collision arithmetic; the FP64-accumulated change is written to
//         dacc and k_conserve adds it to f with the FP64 conservation
//         correction. FP64: f + change is written to fd here. (LDS atomics
//         accumulate in FP64 in both paths: ds_add_f32 stalls far longer on
//         the repeated addresses of a cell's distribution peak.)
// Samples q in [q0, q1) of one cell (q indexes integral 1, then 2, then 3);
// every depletion and stencil point goes to scatter(node, value).
template <bool kF32, int BS, int K, bool kUni, bool kRemap32, class Src, class Scatter>
__device__ __forceinline__ int collide_samples(const Grid& g,
                                               const Axes<typename std::conditional<kRemap32, float, double>::type>& ax,
                                               const CollParams& cp, const CellView& cv,
                                               int rng_cell, int ntime, int n0, int n1, int q0, int q1,
                                               const double* __restrict__ fac3, const Src& src, Scatter&& scatter) {
  using R = typename std::conditional<kF32, float, double>::type;
  const int N = g.N;
  int bad = 0;
  for (int base = q0; base < q1; base += BS * K) {
    int qq[K], tt[K], ss[K], lA[K], lB[K];
    bool valid[K];
    uint32_t wphi[K], wth[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int q = base + k * BS + (int)threadIdx.x;
      qq[k] = q;
      valid[k] = q < q1;
      tt[k] = (q < n0) ? 0 : (q < n0 + n1 ? 1 : 2);
      ss[k] = (tt[k] == 0) ? q : (tt[k] == 1 ? q - n0 : q - n0 - n1);
    }
    pick_batch<K>(cp, cv, N, rng_cell, ntime, qq, tt, ss, valid, lA, lB, wphi, wth, src);
#pragma unroll 1
    for (int k = 0; k < K; ++k) {
      if (!valid[k]) continue;
      bad |= vr_sample<false, R, kF32 && kUni, kRemap32>(g, ax, cp, cv, rng_cell, ntime, qq[k], tt[k], ss[k], fac3[tt[k]], lA[k], lB[k],
                                 wphi[k], wth[k], src, scatter, nullptr);
    }
  }
  return bad;
}

__device__ __forceinline__ CellView cell_view(const double* f, const double* eqv, const double* cdf2,
                                              const int2* tab2, int N) {
  CellView cv;
  cv.f = f;
  cv.eqv = eqv;
  cv.eqc = cdf2;
  cv.neqc = cdf2 + (N + 1);
  cv.eqtab = tab2;
  cv.neqtab = tab2 + kNBucket;
  cv.eq_total = cv.eqc[N];
  cv.neq_total = cv.neqc[N];
  return cv;
}

// FP64 conservation correction of the collision change d of each cell computed
// with FP32 arithmetic: d is replaced by d - |f| (lambda . phi), phi = (1, v, |v|^2), with
// lambda solving sum |f| phi phi^T lambda = sum d phi. In exact arithmetic
// the collision step conserves mass, momentum and energy (each sample's
// depletions equal its replenishments and the remap keeps moments 0-2); the
// correction removes what FP32 rounding leaves, so the cell's moments change by
// FP64 round-off only. One workgroup per cell.
template <int BS>
struct ConserveShared {
  double red[BS / 64][20];
  double tot[20];
  double M[25];
  double lam[5];
  int flag;
};

template <int BS>
__device__ __forceinline__ int conserve_cell(const Grid& g, const double* __restrict__ fc,
                                              const double* __restrict__ da, double* __restrict__ out,
                                              ConserveShared<BS>& sh) {
  const int N = g.N;
  auto& red = sh.red;
  double* lam = sh.lam;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  // The 20 sums (sum d phi_a, then sum |f| phi_a phi_b for a <= b) in one
  // pass over the nodes.
  {
    double sums[20];
#pragma unroll
    for (int z = 0; z < 20; ++z) sums[z] = 0.0;
    constexpr int U = 4;  // loads of f in flight per lane
    for (int lb = threadIdx.x; lb < N; lb += U * BS) {
      double fv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) fv[u] = (lb + u * BS < N) ? fc[lb + u * BS] : 0.0;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int l = lb + u * BS;
        if (l >= N) break;
        int i, j, k;
        node_ijk(g, l, i, j, k);
        const double x = g.ax[0][i], y = g.ax[1][j], z = g.ax[2][k];
        const double phi[5] = {1.0, x, y, z, x * x + y * y + z * z};
        const double d = da[l];
        const double w = fabs(fv[u]);
#pragma unroll
        for (int a = 0; a < 5; ++a) sums[a] += d * phi[a];
#pragma unroll
        for (int b = 0; b < 5; ++b) sums[5 + b] += w * phi[0] * phi[b];
        int o = 10;
#pragma unroll
        for (int a = 1; a < 5; ++a)
#pragma unroll
          for (int b = a; b < 5; ++b) sums[o++] += w * phi[a] * phi[b];
      }
    }
#pragma unroll
    for (int z = 0; z < 20; ++z) {
      const double v = wave_sum(sums[z]);
      if (lane == 0) red[wid][z] = v;
    }
  }
  __syncthreads();
  // totals, the 5x5 system and its solve in LDS (one lane solves; register
  // arrays here would set the register budget of a heavy caller)
  double* tot = sh.tot;
  double* M = sh.M;
  if (threadIdx.x < 20) {
    double v = 0.0;
    for (int w = 0; w < BS / 64; ++w) v += red[w][threadIdx.x];
    tot[threadIdx.x] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int o = 5;
    for (int a = 0; a < 5; ++a)
      for (int c = a; c < 5; ++c) { M[a * 5 + c] = tot[o]; M[c * 5 + a] = tot[o]; ++o; }
    for (int a = 0; a < 5; ++a) lam[a] = tot[a];
    sh.flag = 0;
    if (!solve5(M, lam)) {
      for (int a = 0; a < 5; ++a) lam[a] = 0.0;
      sh.flag = kErrConserve;
    }
  }
  __syncthreads();
  const int flag = sh.flag;
  const double l0 = lam[0], l1 = lam[1], l2 = lam[2], l3 = lam[3], l4 = lam[4];
  constexpr int U = 8;  // loads of f in flight per lane
  for (int lb = threadIdx.x; lb < N; lb += U * BS) {
    double fv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) fv[u] = (lb + u * BS < N) ? fc[lb + u * BS] : 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int l = lb + u * BS;
      if (l >= N) break;
      int i, j, k;
      node_ijk(g, l, i, j, k);
      const double x = g.ax[0][i], y = g.ax[1][j], z = g.ax[2][k];
      const double fl = fv[u];
      const double corr = fabs(fl) * (l0 + l1 * x + l2 * y + l3 * z + l4 * (x * x + y * y + z * z));
      out[l] = fl + (da[l] - corr);
    }
  }
  return flag;
}

// LDS accumulation window of a cell: a box of w[0] x w[1] x w[2] nodes, the
// whole grid when the cell's FP64 slice fits LDS, otherwise centred on the
// node nearest to the cell's bulk velocity (where a relaxing distribution and
// its post-collision velocities concentrate) and clamped to the grid.
struct Window {
  int o[3], w[3], wxy, wn;
};

__device__ __forceinline__ Window make_window(const Grid& g, const int* wd, const double* __restrict__ props) {
  Window win;
  const double c[3] = {props[DVM_P_U], props[DVM_P_V], props[DVM_P_W]};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int n = g.n[d];
    win.w[d] = wd[d];
    double x = (c[d] - g.vmin[d]) * g.factor[d] + 0.5;
    x = fmin(fmax(x, 0.0), (double)(n - 1));  // also maps NaN to 0
    int o = (int)x - wd[d] / 2;
    o = o < 0 ? 0 : o;
    o = o > n - wd[d] ? n - wd[d] : o;
    win.o[d] = __builtin_amdgcn_readfirstlane(o);  // the same in every lane
  }
  win.wxy = win.w[0] * win.w[1];
  win.wn = win.wxy * win.w[2];
  return win;
}

// Inputs of the FP32 fast collision of a cell (collide_fast), beside
// the cell view: its deviation sign bytes staged in LDS and the Cholesky
// factor of sum |f| phi phi^T from the moments pass.
struct FastCell {
  const unsigned char* sgn;  // bit l % 8 of byte l / 8 set when f - f_eq < 0 (or -0); staged into LDS
  const double* cm;          // kCmat doubles: factor, then the not-positive-definite flag
};

// Conservation correction of a cell's FP32 collision change d (LDS) with the
// Cholesky factor L of M = sum |f| phi phi^T precomputed by the moments pass:
// lambda solves L L^T lambda = sum d phi, and out = f + (d - |f| lambda.phi),
// as conserve_cell but with one pass of five sums over d instead of twenty
// over d and f.
// kClear: the change is cleared as it is read by the write (each lane clears
// what it reads), leaving the slice zeroed for the next cell. mid() runs on
// every lane between the sums and the solve (the persistent kernel's hook).
struct NoHook {
  __device__ void operator()() const {}
};
template <int BS, bool kClear = false, class Hook = NoHook>
__device__ __forceinline__ int conserve_cell_factored(const Grid& g, const double* __restrict__ fc,
                                                      double* __restrict__ da, const double* __restrict__ cm,
                                                      double* __restrict__ out, ConserveShared<BS>& sh,
                                                      Hook&& mid = Hook{}) {
  const int N = g.N, nx = g.n[0], nxy = g.nxy;
  const float inv_nxy = (float)g.inv_nxy, inv_nx = (float)g.inv_n0;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double b[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int l = threadIdx.x; l < N; l += BS) {
    const int k = (int)(((float)l + 0.5f) * inv_nxy);
    const int r = l - k * nxy;
    const int j = (int)(((float)r + 0.5f) * inv_nx);
    const int i = r - j * nx;
    const double x = g.ax[0][i], y = g.ax[1][j], z = g.ax[2][k];
    const double d = da[l];
    b[0] += d;
    b[1] += d * x;
    b[2] += d * y;
    b[3] += d * z;
    b[4] += d * (x * x + y * y + z * z);
  }
#pragma unroll
  for (int z = 0; z < 5; ++z) {
    const double v = wave_sum(b[z]);
    if (lane == 0) sh.red[wid][z] = v;
  }
  __syncthreads();
  mid();
  if (threadIdx.x == 0) {
    double y[5];
#pragma unroll
    for (int a = 0; a < 5; ++a) {
      double v = 0.0;
      for (int w = 0; w < BS / 64; ++w) v += sh.red[w][a];
      y[a] = v;
    }
    const double* L = cm;
    const bool ok = L[15] == 0.0;
    sh.flag = ok ? 0 : kErrConserve;
    // L y' = b, then L^T lambda = y' (L packed by rows: row i at i (i + 1) / 2)
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      double v = y[i];
#pragma unroll
      for (int k = 0; k < i; ++k) v -= L[i * (i + 1) / 2 + k] * y[k];
      y[i] = v / L[i * (i + 1) / 2 + i];
    }
#pragma unroll
    for (int i = 4; i >= 0; --i) {
      double v = y[i];
#pragma unroll
      for (int k = i + 1; k < 5; ++k) v -= L[k * (k + 1) / 2 + i] * y[k];
      y[i] = v / L[i * (i + 1) / 2 + i];
    }
    for (int a = 0; a < 5; ++a) sh.lam[a] = ok ? y[a] : 0.0;
  }
  __syncthreads();
  const int flag = sh.flag;
  const double l0 = sh.lam[0], l1 = sh.lam[1], l2 = sh.lam[2], l3 = sh.lam[3], l4 = sh.lam[4];
  // U loads of f in flight per lane (the cell's f comes from HBM)
  constexpr int U = 8;
  for (int lb = threadIdx.x; lb < N; lb += U * BS) {
    double fv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) fv[u] = (lb + u * BS < N) ? fc[lb + u * BS] : 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int l = lb + u * BS;
      if (l >= N) break;
      const int k = (int)(((float)l + 0.5f) * inv_nxy);
      const int r = l - k * nxy;
      const int j = (int)(((float)r + 0.5f) * inv_nx);
      const int i = r - j * nx;
      const double x = g.ax[0][i], y = g.ax[1][j], z = g.ax[2][k];
      const double fl = fv[u];
      const double corr = fabs(fl) * (l0 + l1 * x + l2 * y + l3 * z + l4 * (x * x + y * y + z * z));
      out[l] = fl + (da[l] - corr);
      if constexpr (kClear) da[l] = 0.0;
    }
  }
  return flag;
}

// The FP32-collision sampling of samples [q0, q1) of one cell on equally
// spaced axes with isotropic scattering, accumulating the FP64 change through
// scat (the cell's LDS slice, or its LDS window and global atomics beyond
// it): variance_reduction.f90:686-1116 with
// vr_depletion_routine and the elastic replenish_collision, as vr_sample, in a
// leaner form for this case:
//   * one Philox block per sample (plus one per further replenish location)
//     and the two partner searches of pick_batch;
//   * node indices decoded and node velocities formed in FP32 on the
//     equally spaced axes;
//   * the deviation signs read from the sign bytes in LDS;
//   * for equal masses the post-collision velocities v_com +- r / 2 (the
//     reference's g_1 / g = m_2 / (m_1 + m_2) = 1/2);
//   * the closed-form remap away from the axis ends (remap_uniform_interior),
//     remap_t at the ends.
// Returns device error bits (kErrRemap: a pick outside the grid or a remap
// without an exterior point).
template <int BS, class Scatter>
__device__ __forceinline__ int collide_fast(const Grid& g, const Axes<double>& ax, const CollParams& cp,
                                            const CellView& cv, const unsigned char* sgn, int cell, int ntime,
                                            int q0, int q1, const int* __restrict__ cnt,
                                            const double* __restrict__ fac3, Scatter&& scat) {
  const int N = g.N, nx = g.n[0], nxy = g.nxy;
  const int n0 = cnt[0], n1 = cnt[1];
  const float inv_nxy = (float)g.inv_nxy, inv_nx = (float)g.inv_n0;
  const float vx0 = (float)ax.vmin[0], vy0 = (float)ax.vmin[1], vz0 = (float)ax.vmin[2];
  const float hx = (float)ax.h[0], hy = (float)ax.h[1], hz = (float)ax.h[2];
  const float fac0 = (float)fac3[0], fac1 = (float)fac3[1], fac2 = (float)fac3[2];
  const float vexp = (float)cp.vhs_exponent;
  const float two_pi = (float)(2.0 * kPiRef);
  int bad = 0;
  for (int base = q0; base < q1; base += BS) {
    const int q = base + (int)threadIdx.x;
    const bool valid = q < q1;
    const int t = (q < n0) ? 0 : (q < n0 + n1 ? 1 : 2);
    const int s = (t == 0) ? q : (t == 1 ? q - n0 : q - n0 - n1);
    // partner picks (pick_collision_partner with referenced_search): the
    // searches of pick_batch, whose lanes leave the bisection as their nodes
    // are found (their loads then stop)
    int lA, lB;
    uint32_t wphi, wth;
    pick_batch<1>(cp, cv, N, cell, ntime, &q, &t, &s, &valid, &lA, &lB, &wphi, &wth, PhiloxSrc{cp.seed});
    if (!valid) continue;
    if ((unsigned)lA >= (unsigned)N || (unsigned)lB >= (unsigned)N) {
      bad |= kErrRemap;
      continue;
    }
    const unsigned bytA = sgn[lA >> 3], bytB = sgn[lB >> 3];
    int iA, jA, kA, iB, jB, kB;
    {
      kA = (int)(((float)lA + 0.5f) * inv_nxy);
      const int rA = lA - kA * nxy;
      jA = (int)(((float)rA + 0.5f) * inv_nx);
      iA = rA - jA * nx;
      kB = (int)(((float)lB + 0.5f) * inv_nxy);
      const int rB = lB - kB * nxy;
      jB = (int)(((float)rB + 0.5f) * inv_nx);
      iB = rB - jB * nx;
    }
    const float x1 = fmaf((float)iA, hx, vx0), y1 = fmaf((float)jA, hy, vy0), z1 = fmaf((float)kA, hz, vz0);
    const float x2 = fmaf((float)iB, hx, vx0), y2 = fmaf((float)jB, hy, vy0), z2 = fmaf((float)kB, hz, vz0);
    const float gx = x2 - x1, gy = y2 - y1, gz = z2 - z1;
    const float gg = __builtin_amdgcn_sqrtf(gx * gx + gy * gy + gz * gz);
    if (gg < (float)kDoubleTol) continue;  // self-collision: counted, nothing moves
    const float g_sigma = cp.hard_sphere ? gg : powf(gg, vexp);
    const float vc0 = 0.5f * (x1 + x2), vc1 = 0.5f * (y1 + y2), vc2 = 0.5f * (z1 + z2);
    // signs: +1 for an equilibrium partner, dsgn(f - f_eq) for a deviation partner
    const float sA = (t == 0 || !((bytA >> (lA & 7)) & 1u)) ? 1.0f : -1.0f;
    const float sB = (t == 1 || !((bytB >> (lB & 7)) & 1u)) ? 1.0f : -1.0f;
    const float fac = (t == 0) ? fac0 : (t == 1 ? fac1 : fac2);
    const float kd = sA * sB * fac * g_sigma;  // kin_depl of both partners
    scat(iA, jA, kA, lA, -(double)kd);
    scat(iB, jB, kB, lB, -(double)kd);
    const float kr = (cp.num_repl == 1) ? kd : kd * cp.inv_repl;
    for (int r = 1; r <= cp.num_repl; ++r) {
      if (r > 1) {
        uint32_t x4[4];
        philox_block(cp.seed, (uint32_t)s, (uint32_t)((t + 1) | (r << 2)), (uint32_t)cell, (uint32_t)ntime, x4);
        wphi = x4[0];
        wth = x4[1];
      }
      // pick_random_velocity_on_sphere, isotropic (collision_utilities.f90:284-352)
      const float uphi = (float)wphi * 0x1.0p-32f, uth = (float)wth * 0x1.0p-32f;
      const float ph = two_pi * uphi;
      const float sph = __sinf(ph), cph = __cosf(ph);
      const float ct = 2.0f * uth - 1.0f;
      const float stt = __builtin_amdgcn_sqrtf(1.0f - ct * ct);
      const float hrx = 0.5f * (gg * stt * cph), hry = 0.5f * (gg * stt * sph), hrz = 0.5f * (gg * ct);
#pragma unroll
      for (int side = 0; side < 2; ++side) {
        const float pu = side == 0 ? vc0 + hrx : vc0 - hrx;
        const float pv = side == 0 ? vc1 + hry : vc1 - hry;
        const float pw = side == 0 ? vc2 + hrz : vc2 - hrz;
        if (remap_uniform_interior<double>(ax, pu, pv, pw, (double)kr, nx, nxy, scat)) continue;
        // an axis end: the reference's stencil
        Stencil st;
        remap_t<double, true>(ax, (double)pu, (double)pv, (double)pw, 1.0, st);
        if (st.flag == 0) {
          bad |= kErrRemap;
          continue;
        }
#pragma unroll
        for (int m = 0; m < 7; ++m) {
          if (m < 4 || (st.flag & (1 << (m - 4)))) {
            int si, sj, sk;
            stencil_ijk(st, m, si, sj, sk);
            const int node = si + nx * sj + nxy * sk;
            if ((unsigned)node < (unsigned)N) scat(si, sj, sk, node, (double)kr * st.m[m]);
          }
        }
      }
    }
  }
  return bad;
}

// One work item of the collision phase: samples [q0, q1) of one cell
// (q1 clamped to the cell's count) on BS lanes, accumulated in FP64 in the LDS
// window acc[] (LDS atomics accumulate in FP64 in both precisions: ds_add_f32
// stalls far longer on the repeated addresses of a distribution's peak).
//   full: the window is the whole grid; otherwise scatters outside it go to
//         global FP64 atomics on tgt (fd prefilled with f in FP64, dacc
//         zeroed in FP32).
//   sole: the item is the cell's only one and writes its window directly
//         (FP64: fd = f + change; FP32: dacc = change, which the conservation
//         kernel adds to f); otherwise it adds its window to tgt atomically.
//   FP32, full and sole: the whole change is in LDS, so the FP64
//         conservation correction runs here (csh its reduction space) and
//         fd = f + corrected change is written directly.
template <bool kF32, int BS, int K, bool kUni, bool kRemap32, class Src, bool kFastOnly = false>
__device__ __forceinline__ int collide_item(const Grid& g,
                                            const Axes<typename std::conditional<kRemap32, float, double>::type>& ax,
                                            const CollParams& cp, const CellView& cv, int rng_cell,
                                            int ntime, int q0, int q1, const int* __restrict__ cnt,
                                            const double* __restrict__ fac, const Window& win, bool full, bool sole,
                                            double* acc, double* __restrict__ fd, double* __restrict__ da,
                                            ConserveShared<BS>* csh, const Src& src,
                                            const FastCell* fast = nullptr) {
  const int N = g.N;
  const int total = cnt[0] + cnt[1] + cnt[2];
  q1 = q1 < total ? q1 : total;
  const int wn = full ? N : win.wn;
  if constexpr (kF32 && kUni && !kRemap32 && std::is_same<Src, PhiloxSrc>::value) {
    if (fast) {
      // the cell's slice (or window) in LDS, its sign bytes after it, copied
      // as 32-bit words (a cell's sign bytes start 4-byte aligned, 2 (N+1)
      // bytes per cell, and the words cover at most 3 bytes past them, inside
      // the cell's slot); the words are loaded before the slice is cleared
      unsigned char* sl = reinterpret_cast<unsigned char*>(acc + wn);
      const int nw = (N + 31) >> 5;
      const unsigned* sw = reinterpret_cast<const unsigned*>(fast->sgn);
      unsigned* lw = reinterpret_cast<unsigned*>(sl);
      unsigned w0 = 0u;
      if ((int)threadIdx.x < nw) w0 = sw[threadIdx.x];
      for (int q = threadIdx.x; q < wn; q += BS) acc[q] = 0.0;
      if ((int)threadIdx.x < nw) lw[threadIdx.x] = w0;
      for (int q = threadIdx.x + BS; q < nw; q += BS) lw[q] = sw[q];
      __syncthreads();
      int bad;
      if (full) {
        bad = collide_fast<BS>(g, ax, cp, cv, sl, rng_cell, ntime, q0, q1, cnt, fac,
                               [&](int, int, int, int node, double v) { atomicAdd(&acc[node], v); });
      } else {
        // the window in LDS, the rest of the grid to global FP64 atomics on
        // dacc (zeroed for the cell)
        const int o0 = win.o[0], o1 = win.o[1], o2 = win.o[2], wx = win.w[0], wxy = win.wxy;
        const unsigned w0u = win.w[0], w1u = win.w[1], w2u = win.w[2];
        bad = collide_fast<BS>(g, ax, cp, cv, sl, rng_cell, ntime, q0, q1, cnt, fac,
                               [&](int i, int j, int k, int node, double v) {
                                 const unsigned a = i - o0, b = j - o1, c = k - o2;
                                 if (a < w0u && b < w1u && c < w2u) atomicAdd(&acc[a + wx * b + wxy * c], v);
                                 else atomicAdd(&da[node], v);
                               });
      }
      __syncthreads();
      // the cell's only item with the whole grid in LDS corrects conservation
      // here; otherwise the window (a cell's only item: stored, the nodes
      // outside it having taken atomics; a part of a split cell: added) goes
      // to dacc for the separate correction
      if (full && sole)
        return bad | (fast->cm ? conserve_cell_factored<BS>(g, cv.f, acc, fast->cm, fd, *csh)
                               : conserve_cell<BS>(g, cv.f, acc, fd, *csh));
      const int nx = g.n[0], nxy = g.nxy;
      for (int q = threadIdx.x; q < wn; q += BS) {
        int l = q;
        if (!full) {
          const int c = q / win.wxy, r = q - c * win.wxy;
          const int b = r / win.w[0], a = r - b * win.w[0];
          l = (win.o[0] + a) + nx * (win.o[1] + b) + nxy * (win.o[2] + c);
        }
        const double v = acc[q];
        if (sole) da[l] = v;
        else if (v != 0.0) atomicAdd(&da[l], v);
      }
      return bad;
    }
  }
  // kFastOnly: a kernel whose items all take the fast path (the host checks
  // that they qualify); the general path is not compiled into it, so its
  // registers do not weigh on the fast path's
  if constexpr (kFastOnly) return 0;
  for (int q = threadIdx.x; q < wn; q += BS) acc[q] = 0.0;
  __syncthreads();
  double* tgt = kF32 ? da : fd;
  int bad;
  if (full) {
    bad = collide_samples<kF32, BS, K, kUni, kRemap32>(g, ax, cp, cv, rng_cell, ntime, cnt[0], cnt[1], q0, q1, fac, src,
                                       [&](int, int, int, int node, double v) {
                                         if ((unsigned)node < (unsigned)N) atomicAdd(&acc[node], v);
                                       });
  } else {
    const int o0 = win.o[0], o1 = win.o[1], o2 = win.o[2], wx = win.w[0], wxy = win.wxy;
    const unsigned w0 = win.w[0], w1 = win.w[1], w2 = win.w[2];
    bad = collide_samples<kF32, BS, K, kUni, kRemap32>(g, ax, cp, cv, rng_cell, ntime, cnt[0], cnt[1], q0, q1, fac, src,
                                       [&](int i, int j, int k, int node, double v) {
                                         const unsigned a = i - o0, b = j - o1, c = k - o2;
                                         if (a < w0 && b < w1 && c < w2) atomicAdd(&acc[a + wx * b + wxy * c], v);
                                         else if ((unsigned)node < (unsigned)N) atomicAdd(&tgt[node], v);
                                       });
  }
  __syncthreads();
  if (kF32 && full && sole) return bad | conserve_cell<BS>(g, cv.f, acc, fd, *csh);
  const int nx = g.n[0], nxy = g.nxy;
  if (full && sole && !kF32) {
    // fd = f + change over the whole grid, U loads of f in flight per lane
    constexpr int U = 8;
    for (int qb = threadIdx.x; qb < N; qb += U * BS) {
      double fv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) fv[u] = (qb + u * BS < N) ? cv.f[qb + u * BS] : 0.0;
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (qb + u * BS < N) fd[qb + u * BS] = fv[u] + acc[qb + u * BS];
    }
    return bad;
  }
  for (int q = threadIdx.x; q < wn; q += BS) {
    int l = q;
    if (!full) {
      const int c = q / win.wxy, r = q - c * win.wxy;
      const int b = r / win.w[0], a = r - b * win.w[0];
      l = (win.o[0] + a) + nx * (win.o[1] + b) + nxy * (win.o[2] + c);
    }
    const double v = acc[q];
    if (sole) {
      if (kF32) da[l] = v;
      else fd[l] = cv.f[l] + v;
    } else if (v != 0.0) {
      atomicAdd(&tgt[l], v);
    }
  }
  return bad;
}

// Collision phase: `split` workgroups of BS lanes per cell (each takes one
// part, a contiguous range of the cell's samples; the parts of a cell share an
// XCD).
//   kLds: the cell's FP64 slice fits LDS (window = grid); otherwise the
//         window of extents wd around the cell's bulk velocity.
//   split == 1 and kLds: fd (FP64) is written directly, or (FP32) the
//         conservation correction is applied here and fd written. Otherwise
//         fd is prefilled with f (FP64) or dacc zeroed (FP32, then
//         k_conserve) and each part adds its window atomically.
struct WinDims {
  int w[3];
};

// kStream: the uniforms come from the explicit list su (per-cell offsets sbase,
// sper values per sample) instead of Philox.
// kUni / kRemap32 (FP32 path only): the remap's nearest nodes by rounding on
// equally spaced axes / the remap in FP32 on FP32 axes.
template <bool kLds, bool kF32, int BS, int K, bool kStream, bool kUni = false, bool kRemap32 = false,
          bool kFastOnly = false>
__global__ void __launch_bounds__(BS) k_collide(Grid g, CollParams cp, int ntime, int split, WinDims wd,
                                                const double* __restrict__ props,
                                                const double* __restrict__ f, double* __restrict__ fd,
                                                double* __restrict__ dacc,
                                                const double* __restrict__ eqv, const double* __restrict__ cdf,
                                                const int2* __restrict__ tabs, const int* __restrict__ counts,
                                                const double* __restrict__ factors, int* __restrict__ err,
                                                const double* __restrict__ su, const long long* __restrict__ sbase,
                                                int sper, const unsigned char* __restrict__ sgn,
                                                const double* __restrict__ cmat, int fast) {
  extern __shared__ double dyn[];
  __shared__ ConserveShared<BS> csh;
  // Workgroup b is dispatched to XCD b % 8. With split > 1 the parts of one
  // cell go to consecutive workgroups of the same XCD, so that they share
  // that XCD's L2 copy of the cell's CDFs: XCD x takes the contiguous range
  // [start_x, start_x + count_x) of the (cell, part) order.
  int v = blockIdx.x;
  if (split > 1) {
    const int G = gridDim.x, q = G >> 3, r = G & 7, x = v & 7;
    v = x * q + (x < r ? x : r) + (v >> 3);
  }
  const int cell = v / split, part = v - cell * split;
  const int N = g.N;
  Window win{};
  if (!kLds) win = make_window(g, wd.w, props + (size_t)cell * DVM_NPROP);
  using RM = typename std::conditional<kF32 && kRemap32, float, double>::type;
  const Axes<RM> ax = stage_collision_axes<RM>(g, dyn);
  const CellView cv = cell_view(f + (size_t)cell * N, eqv + (size_t)cell * N, cdf + (size_t)cell * 2 * (N + 1),
                                tabs + (size_t)cell * 2 * kNBucket, N);
  const int* cnt = counts + cell * 4;
  const int chunk = (cnt[0] + cnt[1] + cnt[2] + split - 1) / split;
  const int q0 = part * chunk;
  using Src = typename std::conditional<kStream, StreamSrc, PhiloxSrc>::type;
  Src src;
  if constexpr (kStream) src = StreamSrc{su, sbase, sper}; else src = PhiloxSrc{cp.seed};
  // FP32 on equally spaced axes: the fast sampling path with the sign bytes
  // (and, for a cell's only part with the whole grid in LDS, the
  // precomputed conservation factor)
  FastCell fcs{};
  const FastCell* fc = nullptr;
  if (fast) {
    fcs.sgn = sgn + (size_t)cell * 2 * (N + 1);
    fcs.cm = cmat ? cmat + (size_t)cell * kCmat : nullptr;  // used by a cell's only part
    fc = &fcs;
  }
  const int bad = collide_item<kF32, BS, K, kUni, kF32 && kRemap32, Src, kFastOnly>(g, ax, cp, cv, cell, ntime, q0, split == 1 ? 0x7fffffff : q0 + chunk, cnt,
                                            factors + cell * 3, win, kLds, split == 1, dyn + kAxisSlots,
                                            fd + (size_t)cell * N, dacc + (size_t)cell * N, &csh, src, fc);
  if (bad) atomicOr(err, bad);
}

// The FP32 fast collision of whole cells (each its one part, the grid in
// LDS, the conservation factor valid) on persistent workgroups, one per CU:
// a workgroup stages the axes once and takes cells from a ticket counter
// until they run out, and each cell's epilogue overlaps the next cell's
// prologue: the next ticket is taken and that cell's sign words are loaded
// while the correction runs, and the slice is cleared as the corrected
// values are written. tick[0] is the ticket counter and tick[1] counts the
// workgroups that have finished; the last one resets both to zero, so every
// launch starts from zero without a memset. With kTabLds the cell's two
// bucket tables (2 x 1000 brackets, 16 KB) are staged in LDS after the sign
// words, with them, so that each partner search starts from an LDS read.
template <int BS, bool kTabLds>
__global__ void __launch_bounds__(BS) k_collide_fast_cells(Grid g, CollParams cp, int ntime, int ncell,
                                                           const double* __restrict__ f, double* __restrict__ fd,
                                                           const double* __restrict__ eqv,
                                                           const double* __restrict__ cdf,
                                                           const int2* __restrict__ tabs,
                                                           const int* __restrict__ counts,
                                                           const double* __restrict__ factors, int* __restrict__ err,
                                                           const unsigned char* __restrict__ sgn,
                                                           const double* __restrict__ cmat,
                                                           unsigned* __restrict__ tick) {
  extern __shared__ double dyn[];
  __shared__ ConserveShared<BS> csh;
  __shared__ int s_cell;
  const Axes<double> ax = stage_collision_axes<double>(g, dyn);
  const int N = g.N;
  double* acc = dyn + kAxisSlots;
  unsigned* lw = reinterpret_cast<unsigned*>(acc + N);
  const int nw = (N + 31) >> 5;  // sign words of a cell (at most BS: the slice fits LDS)
  const auto words = [&](int c) {
    return reinterpret_cast<const unsigned*>(sgn + (size_t)c * 2 * (N + 1));
  };
  // the bucket tables after the sign words, 8-byte aligned
  int2* ltab = reinterpret_cast<int2*>(lw + ((nw + 1) & ~1));
  constexpr int kTabPer = (2 * kNBucket + BS - 1) / BS;  // brackets staged per lane
  if (threadIdx.x == 0) s_cell = (int)atomicAdd(tick, 1u);
  __syncthreads();
  int cell = s_cell;
  if (cell < ncell) {
    const unsigned w0 = ((int)threadIdx.x < nw) ? words(cell)[threadIdx.x] : 0u;
    for (int q = threadIdx.x; q < N; q += BS) acc[q] = 0.0;
    if ((int)threadIdx.x < nw) lw[threadIdx.x] = w0;
    if constexpr (kTabLds)
      for (int r = threadIdx.x; r < 2 * kNBucket; r += BS) ltab[r] = tabs[(size_t)cell * 2 * kNBucket + r];
  }
  __syncthreads();
  int bad = 0;
  while (cell < ncell) {
    CellView cv = cell_view(f + (size_t)cell * N, eqv + (size_t)cell * N, cdf + (size_t)cell * 2 * (N + 1),
                            tabs + (size_t)cell * 2 * kNBucket, N);
    if constexpr (kTabLds) {
      cv.eqtab = ltab;
      cv.neqtab = ltab + kNBucket;
    }
    const int* cnt = counts + cell * 4;
    const int total = cnt[0] + cnt[1] + cnt[2];
    bad |= collide_fast<BS>(g, ax, cp, cv, reinterpret_cast<const unsigned char*>(lw), cell, ntime, 0, total, cnt,
                            factors + cell * 3,
                            [&](int, int, int, int node, double v) { atomicAdd(&acc[node], v); });
    if (threadIdx.x == 0) s_cell = (int)atomicAdd(tick, 1u);
    __syncthreads();  // the change is complete; the sign words are no longer read
    int next = ncell;
    unsigned wn = 0u;
    int2 tn[kTabPer];
    bad |= conserve_cell_factored<BS, true>(g, cv.f, acc, cmat + (size_t)cell * kCmat, fd + (size_t)cell * N, csh,
                                            [&] {
                                              next = s_cell;
                                              if (next < ncell && (int)threadIdx.x < nw)
                                                wn = words(next)[threadIdx.x];
                                              if constexpr (kTabLds) {
#pragma unroll
                                                for (int k = 0; k < kTabPer; ++k) {
                                                  const int r = (int)threadIdx.x + k * BS;
                                                  tn[k] = (next < ncell && r < 2 * kNBucket)
                                                              ? tabs[(size_t)next * 2 * kNBucket + r]
                                                              : make_int2(0, 0);
                                                }
                                              }
                                            });
    if ((int)threadIdx.x < nw) lw[threadIdx.x] = wn;
    if constexpr (kTabLds) {
      // the current cell's searches ended at the barrier before the correction
#pragma unroll
      for (int k = 0; k < kTabPer; ++k) {
        const int r = (int)threadIdx.x + k * BS;
        if (r < 2 * kNBucket) ltab[r] = tn[k];
      }
    }
    __syncthreads();  // the slice is cleared and the next cell's words and brackets staged
    cell = next;
  }
  if (bad) atomicOr(err, bad);
  // the last workgroup to finish resets the counters for the next launch
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(tick + 1, 1u) == gridDim.x - 1) {
      atomicExch(tick, 0u);
      atomicExch(tick + 1, 0u);
    }
  }
}

template <int BS>
__global__ void __launch_bounds__(BS) k_conserve(Grid g, const double* __restrict__ f,
                                                 const double* __restrict__ dacc, double* __restrict__ fd,
                                                 int* __restrict__ err) {
  __shared__ ConserveShared<BS> sh;
  const size_t off = (size_t)blockIdx.x * g.N;
  const int flag = conserve_cell<BS>(g, f + off, dacc + off, fd + off, sh);
  if (flag && threadIdx.x == 0) atomicOr(err, flag);
}

// Offsets of each cell's samples in an explicit uniform list: cells in order,
// each taking per * (its sample count) values from `cursor` on; total[0] is
// the number of values the step reads. One lane (a parity mode, not a hot
// path).
__global__ void k_stream_bases(const int* __restrict__ counts, int ncell, long long cursor, int per,
                               long long* __restrict__ base, long long* __restrict__ total) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  long long acc = 0;
  for (int c = 0; c < ncell; ++c) {
    base[c] = cursor + acc * per;
    acc += counts[c * 4 + 3];
  }
  total[0] = acc * per;
}

// Parity trace of one cell: the same samples as the FP64 k_collide, recorded
// instead of scattered.
template <bool kStream>
__global__ void __launch_bounds__(kCollBS) k_trace(Grid g, CollParams cp, int ntime, int cell,
                                                   const double* __restrict__ f, const double* __restrict__ eqv,
                                                   const double* __restrict__ cdf, const int2* __restrict__ tabs,
                                                   const int* __restrict__ counts, const double* __restrict__ factors,
                                                   long long max_records, int* __restrict__ irec,
                                                   double* __restrict__ drec, const double* __restrict__ su,
                                                   const long long* __restrict__ sbase, int sper) {
  using Src = typename std::conditional<kStream, StreamSrc, PhiloxSrc>::type;
  Src src;
  if constexpr (kStream) src = StreamSrc{su, sbase, sper}; else src = PhiloxSrc{cp.seed};
  const int N = g.N;
  const CellView cv = cell_view(f + (size_t)cell * N, eqv + (size_t)cell * N, cdf + (size_t)cell * 2 * (N + 1),
                                tabs + (size_t)cell * 2 * kNBucket, N);
  const int n0 = counts[cell * 4 + 0], n1 = counts[cell * 4 + 1], n2 = counts[cell * 4 + 2];
  const int total = n0 + n1 + n2;
  for (int q = threadIdx.x + blockIdx.x * blockDim.x; q < total && q < max_records; q += blockDim.x * gridDim.x) {
    int t = (q < n0) ? 0 : (q < n0 + n1 ? 1 : 2);
    int s = (t == 0) ? q : (t == 1 ? q - n0 : q - n0 - n1);
    SampleRec rec;
    for (int z = 0; z < 7; ++z) { rec.st1[z] = 0; rec.st2[z] = 0; rec.m1[z] = 0.0; rec.m2[z] = 0.0; }
    rec.kd1 = 0.0; rec.kd2 = 0.0;
    const bool valid = true;
    int lA, lB;
    uint32_t wphi, wth;
    const int qi = q;
    pick_batch<1>(cp, cv, N, cell, ntime, &qi, &t, &s, &valid, &lA, &lB, &wphi, &wth, src);
    rec.lA = lA; rec.lB = lB; rec.skipped = 0;
    vr_sample<true, double, false, false>(g, grid_axes(g), cp, cv, cell, ntime, qi, t, s, factors[cell * 3 + t], lA, lB, wphi, wth, src,
                            [](int, int, int, int, double) {}, &rec);
    int* ir = irec + (size_t)q * 32;
    double* dr = drec + (size_t)q * 16;
    for (int z = 0; z < 32; ++z) ir[z] = 0;
    for (int z = 0; z < 16; ++z) dr[z] = 0.0;
    ir[0] = t + 1; ir[1] = s; ir[2] = rec.lA; ir[3] = rec.lB; ir[4] = rec.skipped;
    if (!rec.skipped) {
      for (int z = 0; z < 7; ++z) {
        ir[5 + z] = rec.st1[z]; ir[12 + z] = rec.st2[z];
        dr[2 + z] = rec.m1[z]; dr[9 + z] = rec.m2[z];
      }
      dr[0] = rec.kd1; dr[1] = rec.kd2;
    }
  }
}

// perform_remapping over a batch of velocities (dvm_remap).
__global__ void k_remap_batch(Grid g, long long n, const double* __restrict__ uvw, double f_off,
                              int* __restrict__ st, double* __restrict__ m) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  Stencil s;
  remap(g, uvw[3 * q], uvw[3 * q + 1], uvw[3 * q + 2], f_off, s);
  int* o = st + 7 * q;
  o[0] = s.oi; o[1] = s.oj; o[2] = s.ok; o[3] = s.a; o[4] = s.b; o[5] = s.c; o[6] = s.flag;
  for (int z = 0; z < 7; ++z) m[7 * q + z] = s.m[z];
}

// ------------------------------------------------------------ convection
// Diffuse-wall flux balance of each wall-adjacent cell
// (boundary_conditions.f90:753-829 and the right/bottom/top analogues):
// ratio = flux_in / flux_out with flux_in = sum over incoming nodes of
// |v_n| * acc * f. blockIdx.y = side (0 left, 1 right, 2 bottom, 3 top).
// The lanes run over the incoming nodes only: runs of consecutive i in each
// (j, k) row for the left and right walls, blocks of whole rows in each k
// plane for the bottom and top walls; four loads in flight per lane.
__global__ void __launch_bounds__(kWallBS) k_wall_ratio(Grid g, ConvParams p, const double* __restrict__ f,
                                                        double* __restrict__ ratio) {
  __shared__ double red[kWallBS / 64];
  const int side = blockIdx.y, pos = blockIdx.x;
  const int along = (side < 2) ? p.ny : p.nx;
  if (pos >= along || !p.wall_side[side]) return;
  const int ix = (side == 0) ? 0 : (side == 1 ? p.nx - 1 : pos);
  const int iy = (side == 2) ? 0 : (side == 3 ? p.ny - 1 : pos);
  const double* fc = f + (size_t)(ix + p.nx * iy) * g.N;
  const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
  // the incoming range of the wall-normal index: [lo, lo + cnt)
  const bool xwall = side < 2;
  const int nn = xwall ? n0 : n1;
  const int lo = (side == 0 || side == 2) ? 0 : (xwall ? g.pos[0] : g.pos[1]);
  const int hi = (side == 0 || side == 2) ? (xwall ? g.neg[0] : g.neg[1]) + 1 : nn;
  const int cnt = hi > lo ? hi - lo : 0;
  // x walls: t = ii + cnt * r over rows r of (j, k); y walls: t = rem + (n0 cnt) * k
  const int run = xwall ? cnt : n0 * cnt;
  const int total = run * (xwall ? n1 * n2 : n2);
  const double inv_run = run > 0 ? 1.0 / (double)run : 0.0;
  const double inv_n0 = g.inv_n0;
  const double* __restrict__ axn = xwall ? g.ax[0] : g.ax[1];
  constexpr int U = 4;
  double s = 0.0;
  for (int t0 = threadIdx.x; t0 < total; t0 += U * kWallBS) {
    double fv[U], vn[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int t = t0 + u * kWallBS;
      fv[u] = 0.0;
      vn[u] = 0.0;
      if (t < total) {
        const int r = (int)(((double)t + 0.5) * inv_run);
        const int q = t - r * run;
        int l, w;
        if (xwall) {
          w = lo + q;
          l = w + n0 * r;
        } else {
          const int jj = (int)(((double)q + 0.5) * inv_n0);
          w = lo + jj;
          l = r * g.nxy + lo * n0 + q;
        }
        fv[u] = fc[l];
        vn[u] = axn[w];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) s += fabs(vn[u]) * p.acc * fv[u];
  }
  s = block_sum<kWallBS>(s, red);
  if (threadIdx.x == 0) ratio[side * p.max_along + pos] = s / p.flux_out[side];
}

// Heat flux into each wall cell's wall (boundary_conditions.f90:872-923 and
// the right / bottom / top analogues at :1069-1115, :1262-1307, :1450-1495):
// q = m/2 (sum over the incoming half of v_n f |c - u_cell|^2 + sum over the
// outgoing half of v_n g |c - u_wall|^2), g the wall ghost the sweep uses
// (diffuse part ratio * wallM, specular part spec), u_cell the cell's bulk
// velocity from the moments the step started with. The zero-velocity plane
// belongs to neither half. blockIdx.y = side.
struct WallVel {
  double u[4][3];
};

__global__ void __launch_bounds__(kMomBS) k_wall_heat(Grid g, ConvParams p, WallVel wv, double mass,
                                                      const double* __restrict__ f, const double* __restrict__ props,
                                                      const double* __restrict__ wallM,
                                                      const double* __restrict__ ratio,
                                                      const double* __restrict__ spec, double* __restrict__ q) {
  __shared__ double red[kMomBS / 64];
  const int side = blockIdx.y, pos = blockIdx.x;
  const int along = (side < 2) ? p.ny : p.nx;
  if (pos >= along || !p.wall_side[side]) return;
  const int ix = (side == 0) ? 0 : (side == 1 ? p.nx - 1 : pos);
  const int iy = (side == 2) ? 0 : (side == 3 ? p.ny - 1 : pos);
  const size_t cell = (size_t)ix + (size_t)p.nx * iy;
  const double* fc = f + cell * g.N;
  const double* pc = props + cell * DVM_NPROP;
  const double xb = pc[DVM_P_U], yb = pc[DVM_P_V], zb = pc[DVM_P_W];
  const double uw = wv.u[side][0], vw = wv.u[side][1], ww = wv.u[side][2];
  const double rt = (p.acc > 0.0) ? ratio[side * p.max_along + pos] : 0.0;
  const double* wm = wallM + (size_t)side * g.N;
  const double* sp = spec + ((size_t)side * p.max_along + pos) * g.N;
  double hin = 0.0, hout = 0.0;
  for (int l = threadIdx.x; l < g.N; l += kMomBS) {
    int i, j, k;
    node_ijk(g, l, i, j, k);
    const double x = g.ax[0][i], y = g.ax[1][j], z = g.ax[2][k];
    bool in, out;
    if (side == 0) { in = i <= g.neg[0]; out = i >= g.pos[0]; }
    else if (side == 1) { in = i >= g.pos[0]; out = i <= g.neg[0]; }
    else if (side == 2) { in = j <= g.neg[1]; out = j >= g.pos[1]; }
    else { in = j >= g.pos[1]; out = j <= g.neg[1]; }
    const double vn = (side < 2) ? x : y;
    if (in) {
      const double cx = (x - xb) * (x - xb), cy = (y - yb) * (y - yb), cz = (z - zb) * (z - zb);
      hin += vn * fc[l] * (cx + cy + cz);
    } else if (out) {
      double gv = 0.0;
      if (p.acc > 0.0) gv = rt * wm[l];
      if (p.acc < 1.0) gv += sp[l];
      const double cx = (x - uw) * (x - uw), cy = (y - vw) * (y - vw), cz = (z - ww) * (z - ww);
      hout += vn * gv * (cx + cy + cz);
    }
  }
  hin = block_sum<kMomBS>(hin, red);
  hout = block_sum<kMomBS>(hout, red);
  if (threadIdx.x == 0) q[side * p.max_along + pos] = 0.5 * mass * (hin + hout);
}

// Specular part of the wall ghosts (boundary_conditions.f90:832-858 and
// analogues): (1-acc)*f at each incoming node, reflected and remapped onto the
// ghost with apply_bc_remapping (:1852-1929). One lane per (incoming node).
__global__ void k_spec_ghost(Grid g, ConvParams p, const double* __restrict__ f, double* __restrict__ spec,
                             int* __restrict__ err) {
  const int side = blockIdx.z, pos = blockIdx.y;
  const int along = (side < 2) ? p.ny : p.nx;
  if (pos >= along || !p.wall_side[side]) return;
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= g.N) return;
  int i, j, k;
  node_ijk(g, l, i, j, k);
  bool in;
  if (side == 0) in = i <= g.neg[0];
  else if (side == 1) in = i >= g.pos[0];
  else if (side == 2) in = j <= g.neg[1];
  else in = j >= g.pos[1];
  if (!in) return;
  const int ix = (side == 0) ? 0 : (side == 1 ? p.nx - 1 : pos);
  const int iy = (side == 2) ? 0 : (side == 3 ? p.ny - 1 : pos);
  const double dens = (1.0 - p.acc) * f[(size_t)(ix + p.nx * iy) * g.N + l];
  double x = g.ax[0][i], y = g.ax[1][j];
  const double z = g.ax[2][k];
  if (side < 2) x = -x; else y = -y;
  Stencil s;
  remap(g, x, y, z, 1.0, s);
  if (s.flag == 0) { atomicOr(err, kErrSpecRemap); return; }
  double* gh = spec + ((size_t)side * p.max_along + pos) * g.N;
  for (int q = 0; q < 7; ++q) {
    if (q < 4 || (s.flag & (1 << (q - 4)))) {
      const int node = stencil_node(s, q, g.n[0], g.nxy);
      if ((unsigned)node < (unsigned)g.N) atomicAdd(&gh[node], dens * s.m[q]); else atomicOr(err, kErrSpecNode);
    }
  }
}

__global__ void k_fill(double* __restrict__ a, long long n, double v) {
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
    a[q] = v;
}

// Ghost value seen by cell (ix,iy) through `side` at node l.
__device__ __forceinline__ double ghost(const ConvParams& p, const double* __restrict__ f, const double* __restrict__ wallM,
                                        const double* __restrict__ ratio, const double* __restrict__ spec, int N,
                                        int side, int ix, int iy, int l, size_t idx) {
  const bool at_wall = (side == 0) ? ix == 0 : (side == 1) ? ix == p.nx - 1 : (side == 2) ? iy == 0 : iy == p.ny - 1;
  if (!at_wall) {
    const long long off = (side == 0) ? -(long long)N : (side == 1) ? (long long)N
                        : (side == 2) ? -(long long)p.nx * N : (long long)p.nx * N;
    return f[(long long)idx + off];
  }
  const int pos = (side < 2) ? iy : ix;
  if (p.wall_side[side]) {
    double gv = 0.0;
    if (p.acc > 0.0) gv = ratio[side * p.max_along + pos] * wallM[(size_t)side * N + l];
    if (p.acc < 1.0) gv += spec[((size_t)side * p.max_along + pos) * N + l];
    return gv;
  }
  if (p.couette_side[side]) {
    if (side == 0) return f[idx];
    return f[(size_t)(0 + p.nx * iy) * N + l];
  }
  return 0.0;  // open side in 2-D trunk semantics: zero ghost
}

// 2-D upwind sweep: interior cells (convection.f90:119-276) and boundary cells
// (boundary_conditions.f90:1548-1643) in one pass. A workgroup takes one chunk
// of kConvBS velocity nodes over a strip of S adjacent columns and a segment of
// rows: each lane owns one node in every column of the strip and walks the
// segment in y, holding the rows iy-1, iy and iy+1 of its node in registers, so
// the y-upwind value is read once per segment. The x-upwind value of a column
// is the neighbouring column's register, except at the strip's upwind edge,
// where one halo value per row is loaded (the column left of the strip for
// x >= 0 nodes, right of it for x < 0). Loads run D rows ahead of the row being
// computed, so (S + 1) D loads per lane are in flight.
// Grid: x = node chunks padded to a multiple of 8, y = strips, z = row
// segments. Workgroups are dispatched to the 8 XCDs round-robin, so chunk j of
// every strip and segment runs on XCD j % 8, and a strip's halo column (the edge
// column of the neighbouring strip, read by the same chunk) comes through that
// XCD's L2.
template <int S, int D>
__global__ void __launch_bounds__(kConvBS) k_convect2d(Grid g, ConvParams p, const double* __restrict__ f,
                                                       double* __restrict__ fo, const double* __restrict__ wallM,
                                                       const double* __restrict__ ratio,
                                                       const double* __restrict__ spec, int seg_rows) {
  const int N = g.N;
  const int l = blockIdx.x * kConvBS + threadIdx.x;
  const int ix0 = blockIdx.y * S;
  const int y0 = blockIdx.z * seg_rows;
  if (l >= N || ix0 >= p.nx || y0 >= p.ny) return;
  const int y1 = (p.ny - y0 < seg_rows) ? p.ny : y0 + seg_rows;  // rows [y0, y1)
  const int ylim = (y1 < p.ny) ? y1 + 1 : p.ny;                    // rows [y0 - 1, ylim) around the segment
  int i, j, k;
  node_ijk(g, l, i, j, k);
  const double xv = g.ax[0][i], yv = g.ax[1][j];
  const double cfl_x = p.fx * xv, cfl_y = p.fy * yv;
  const double ax_ = fabs(cfl_x), ay_ = fabs(cfl_y);
  // x region: 0 upwind from the right ghost, 1 no x term (zero plane), 2 from the left ghost
  int xr = -1, yr = -1;
  if (i <= g.neg[0]) xr = 0; else if (p.izflag && i == g.zero[0]) xr = 1; else if (i >= g.pos[0]) xr = 2;
  if (j <= g.neg[1]) yr = 0; else if (p.jzflag && j == g.zero[1]) yr = 1; else if (j >= g.pos[1]) yr = 2;
  const int ns = (p.nx - ix0 < S) ? p.nx - ix0 : S;  // columns of this strip
  const size_t row = (size_t)p.nx * N;              // distance between rows iy and iy+1
  const bool xneg = xv < 0.0;                       // x-upwind neighbour is ix+1
  const int hx = xneg ? ix0 + S : ix0 - 1;          // halo column
  const bool hval = hx >= 0 && hx < p.nx;
  const long long hoff = (long long)(hx - ix0) * N;
  const size_t col0 = (size_t)ix0 * N + l;          // column ix0 of row 0
  // the row below the segment is the y-upwind row of y >= 0 nodes, the row
  // above it that of y < 0 nodes; each node loads only its own
  const bool yneg = yv < 0.0;
  const int ylo = yneg ? y0 : y0 - 1;
  const int yhi = yneg ? ylim : y1;                 // rows read: [ylo, yhi)
  // r[0] row iy-1, r[1] row iy, r[1 + d] row iy + d; h[d] the halo of row iy + d
  double r[D + 2][S], h[D + 1];
#pragma unroll
  for (int c = 0; c < S; ++c)
    r[0][c] = (ylo < y0 && y0 > 0 && c < ns) ? f[col0 + (size_t)(y0 - 1) * row + (size_t)c * N] : 0.0;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const int yl = y0 + d;
    const size_t rb = col0 + (size_t)yl * row;
    const bool ld = yl < yhi;
#pragma unroll
    for (int c = 0; c < S; ++c) r[1 + d][c] = (ld && c < ns) ? f[rb + (size_t)c * N] : 0.0;
    h[d] = (yl < y1 && hval) ? f[(long long)rb + hoff] : 0.0;
  }
  for (int iy = y0; iy < y1; ++iy) {
    {
      const int yl = iy + D;
      const size_t rb = col0 + (size_t)yl * row;
      const bool ld = yl < yhi;
#pragma unroll
      for (int c = 0; c < S; ++c) r[D + 1][c] = (ld && c < ns) ? f[rb + (size_t)c * N] : 0.0;
      h[D] = (yl < y1 && hval) ? f[(long long)rb + hoff] : 0.0;
    }
    const size_t rb = col0 + (size_t)iy * row;
    const double hc = h[0];
    const bool yin = iy > 0 && iy < p.ny - 1;
#pragma unroll
    for (int c = 0; c < S; ++c) {
      if (c < ns) {
        const int ix = ix0 + c;
        const size_t idx = rb + (size_t)c * N;
        const double fc = r[1][c];
        double out;
        if (yin && ix > 0 && ix < p.nx - 1) {
          const double px = xneg ? (c + 1 < S ? r[1][c + 1 < S ? c + 1 : c] : hc)
                                 : (c > 0 ? r[1][c > 0 ? c - 1 : c] : hc);
          const double py = (yv < 0.0) ? r[2][c] : r[0][c];
          out = fc + ax_ * (px - fc) + ay_ * (py - fc);
        } else if (xr < 0 || yr < 0) {
          out = 0.0;
        } else {
          out = fc;
          if (xr != 1) {
            const double gx = ghost(p, f, wallM, ratio, spec, N, xr == 0 ? 1 : 0, ix, iy, l, idx);
            out = out + ax_ * (gx - fc);
          }
          if (yr != 1) {
            const int side = yr == 0 ? 3 : 2;
            const bool wall = (side == 2) ? iy == 0 : iy == p.ny - 1;
            const double gy = wall ? ghost(p, f, wallM, ratio, spec, N, side, ix, iy, l, idx)
                                   : (side == 2 ? r[0][c] : r[2][c]);
            out = out + ay_ * (gy - fc);
          }
        }
        fo[idx] = out;
      }
    }
#pragma unroll
    for (int d = 0; d <= D; ++d) {
#pragma unroll
      for (int c = 0; c < S; ++c) r[d][c] = r[d + 1][c];
    }
#pragma unroll
    for (int d = 0; d < D; ++d) h[d] = h[d + 1];
  }
}

// Host launch of the 2-D sweep.
void launch_convect2d(const Grid& g, const ConvParams& p, const double* f, double* fo, const double* wallM,
                      const double* ratio, const double* spec, hipStream_t st) {
  constexpr int S = kConvStrip, D = kConvDepth;
  // node chunks padded to a multiple of 8 (one XCD per chunk index)
  const unsigned chunks = (unsigned)(((g.N + kConvBS - 1) / kConvBS + 7) / 8 * 8);
  const int strips = (p.nx + S - 1) / S;
  const int segs = (p.ny + kConvSegRows - 1) / kConvSegRows;
  hipLaunchKernelGGL((k_convect2d<S, D>), dim3(chunks, strips, segs), dim3(kConvBS), 0, st, g, p, f, fo, wallM,
                     ratio, spec, kConvSegRows);
}

// 1-D sweep of the zone-grid solver (grid_ref_code convection.f90:111-256)
// with its in/out-flow (boundary_conditions.f90:783-913), zero-gradient
// (:915-973) and diffuse / specular wall (:566-781) boundary cells. Order 2
// adds the limited Lax-Wendroff correction with minmod slopes (zero slope in
// the boundary cells), which the reference does not have:
//   x >= 0: - nu (1 - nu) / 2 (s_i - s_(i-1)),   nu = dt/dx x
//   x <  0: + mu (1 - mu) / 2 (s_i - s_(i+1)),   mu = -nu.
__device__ __forceinline__ double minmod(double a, double b) {
  if (a * b <= 0.0) return 0.0;
  return fabs(a) < fabs(b) ? a : b;
}

__device__ __forceinline__ double slope1d(const double* __restrict__ f, int ix, int nx, size_t N, int l) {
  if (ix <= 0 || ix >= nx - 1) return 0.0;
  const double p0 = f[(size_t)ix * N + l];
  return minmod(f[(size_t)(ix + 1) * N + l] - p0, p0 - f[(size_t)(ix - 1) * N + l]);
}

__device__ __forceinline__ double upwind1d(const double* __restrict__ f, int ix, int nx, int N, int l, double cfl,
                                           double xv, int order) {
  const double p0 = f[(size_t)ix * N + l];
  if (xv < 0.0) {
    double out = p0 - cfl * (f[(size_t)(ix + 1) * N + l] - p0);
    if (order == 2) {
      const double mu = -cfl;
      out = out + (mu * (1.0 - mu) * 0.5) * (slope1d(f, ix, nx, N, l) - slope1d(f, ix + 1, nx, N, l));
    }
    return out;
  }
  double out = p0 - cfl * (p0 - f[(size_t)(ix - 1) * N + l]);
  if (order == 2) out = out - (cfl * (1.0 - cfl) * 0.5) * (slope1d(f, ix, nx, N, l) - slope1d(f, ix - 1, nx, N, l));
  return out;
}

// x-index range of the incoming and outgoing halves of a 1-D wall (side 0
// left, 1 right), from the trunk's zero points (the zone-grid solver reads
// i_neg / i_pos without setting them).
__device__ __forceinline__ void wall1d_halves(const Grid& g, int side, int& in0, int& in1, int& out0, int& out1) {
  if (side == 0) { in0 = 0; in1 = g.neg[0]; out0 = g.pos[0]; out1 = g.n[0] - 1; }
  else { in0 = g.pos[0]; in1 = g.n[0] - 1; out0 = 0; out1 = g.neg[0]; }
}

// One node l of 1-D cell ix from the post-collision f into fo.
__device__ __forceinline__ void convect1d_point(const Grid& g, const ConvParams& p, int i, double xv,
                                                const double* __restrict__ f, double* __restrict__ fo,
                                                const double* __restrict__ inflow, const double* __restrict__ wallM,
                                                const double* __restrict__ ratio, const double* __restrict__ spec,
                                                int ix, int l) {
  const int N = g.N, nx = p.nx;
  const double cfl = p.fx * xv;
  const size_t idx = (size_t)ix * N + l;
  if (ix > 0 && ix < nx - 1) { fo[idx] = upwind1d(f, ix, nx, N, l, cfl, xv, p.order); return; }
  const int side = (ix == 0) ? 0 : 1;
  const int b = p.bc[side];
  if (b == DVM_BC_INOUT) {
    fo[idx] = inflow[(size_t)side * N + l];
  } else if (b == DVM_BC_ZEROGRAD) {
    fo[idx] = upwind1d(f, side == 0 ? 1 : nx - 2, nx, N, l, cfl, xv, p.order);
  } else if (b == DVM_BC_WALL) {
    int in0, in1, out0, out1;
    wall1d_halves(g, side, in0, in1, out0, out1);
    double v = 0.0;
    if (i >= out0 && i <= out1) {
      v = ratio[side] * wallM[(size_t)side * N + l];
    } else if (i >= in0 && i <= in1) {
      const double f0 = f[idx], fn = f[side == 0 ? idx + N : idx - N];
      v = f0 - fabs(p.fx * xv) * (f0 - fn);
    }
    if (p.acc < 1.0) v = v + spec[(size_t)side * N + l];
    fo[idx] = v;
  } else {
    fo[idx] = f[idx];
  }
}

__global__ void __launch_bounds__(kConvBS) k_convect1d(Grid g, ConvParams p, const double* __restrict__ f,
                                                       double* __restrict__ fo, const double* __restrict__ inflow,
                                                       const double* __restrict__ wallM,
                                                       const double* __restrict__ ratio,
                                                       const double* __restrict__ spec) {
  const int N = g.N;
  const int l = blockIdx.x * kConvBS + threadIdx.x;
  if (l >= N) return;
  int i, j, k;
  node_ijk(g, l, i, j, k);
  const double xv = g.ax[0][i];
  for (int ix = blockIdx.y; ix < p.nx; ix += gridDim.y)
    convect1d_point(g, p, i, xv, f, fo, inflow, wallM, ratio, spec, ix, l);
}

// 1-D wall scale factor per side (grid.x = side): acc * (sum of the wall
// cell's f over the incoming half) / (sum of M_wall beta^3 over the outgoing
// half), apply_diffuse_reflection (boundary_conditions.f90:566-781).
__global__ void __launch_bounds__(kMomBS) k_wall1d_ratio(Grid g, ConvParams p, const double* __restrict__ f,
                                                         double* __restrict__ ratio) {
  __shared__ double red[kMomBS / 64];
  const int side = blockIdx.x;
  if (p.bc[side] != DVM_BC_WALL) return;
  int in0, in1, out0, out1;
  wall1d_halves(g, side, in0, in1, out0, out1);
  const double* fc = f + (size_t)(side == 0 ? 0 : p.nx - 1) * g.N;
  double s = 0.0;
  for (int l = threadIdx.x; l < g.N; l += kMomBS) {
    int i, j, k;
    node_ijk(g, l, i, j, k);
    if (i >= in0 && i <= in1) s += fc[l];
  }
  s = block_sum<kMomBS>(s, red);
  if (threadIdx.x == 0) ratio[side] = (p.acc * s) / p.flux_out[side];
}

// Specular part of a 1-D wall: (1 - acc) f of every incoming node of the wall
// cell, remapped from (-x, y, z), added into spec[side] (zeroed before). The
// reference visits the incoming nodes in index order and assigns each node
// before adding its own remap, so what a node adds to a later incoming node
// is overwritten: those contributions are dropped here.
__global__ void k_spec1d(Grid g, ConvParams p, const double* __restrict__ f, double* __restrict__ spec,
                         int* __restrict__ err) {
  const int side = blockIdx.y;
  if (p.bc[side] != DVM_BC_WALL) return;
  const int N = g.N;
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= N) return;
  int i, j, k;
  node_ijk(g, l, i, j, k);
  int in0, in1, out0, out1;
  wall1d_halves(g, side, in0, in1, out0, out1);
  if (i < in0 || i > in1) return;
  const double f0 = f[(size_t)(side == 0 ? 0 : p.nx - 1) * N + l];
  const double ds = (1.0 - p.acc) * f0;
  Stencil st;
  remap(g, -g.ax[0][i], g.ax[1][j], g.ax[2][k], 1.0, st);
  if (st.flag == 0) { atomicOr(err, kErrSpecRemap); return; }
  double* sp = spec + (size_t)side * N;
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    if (q >= 4 && !(st.flag & (1 << (q - 4)))) continue;
    int ti, tj, tk;
    stencil_ijk(st, q, ti, tj, tk);
    const int t = ti + g.n[0] * tj + g.nxy * tk;
    if ((unsigned)t >= (unsigned)N) continue;
    if (ti >= in0 && ti <= in1 && t > l) continue;  // overwritten by the later node's assignment
    atomicAdd(&sp[t], ds * st.m[q]);
  }
}

}  // namespace dvm
