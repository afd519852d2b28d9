// dvm_batch.hip -- many independent solves (space-homogeneous, 1-D and 2-D)
// with uneven velocity grids and cell counts, resident together and advanced
// by one replayed HIP graph per step. Included at the end of dvm_capi.hip (one
// translation unit).
//
// Each solve ("slot") owns a velocity grid, a time step, a sample count, a
// seed and its own step counter. Its distribution is a contiguous range of
// two node pools (ping-pong); its CDFs a range of the CDF pool; its per-cell
// moments, bucket tables and counters a range of the cell pools. Slots are
// admitted and retired between steps while the others stay resident.
//
// One step of every active slot:
//   k_zero_int         the work-queue heads (a fill kernel, graph-safe)
//   k_cdf_cells        one workgroup per resident cell (1024 lanes when fewer
//                      than four cells per CU are resident, else 256)
//   k_prep_cells       fd = f (FP64) or dacc = 0 (FP32) for cells that take
//                      global atomics (window smaller than the grid, or
//                      samples split over several work items)
//   k_collide_slots    persistent workgroups pull work items (a cell's sample
//                      range) from per-XCD queues, costliest cells first, a
//                      cell's items in one queue; each accumulates in
//                      an LDS window (the whole grid when it fits, else a box
//                      around the cell's bulk velocity, spilling the rest to
//                      global FP64 atomics)
//   k_conserve_cells   FP32 path: FP64 conservation correction of the cells
//                      the collision kernel could not correct itself
//   k_convect1d_cells  1-D slots: the zone-grid sweep with in/out flow
//   per 2-D slot       the wall flux ratios, specular ghosts and 2-D sweep of
//                      the slot's grid (k_wall_ratio, k_spec_ghost,
//                      k_convect2d), from the post-collision pool back into
//                      the pool the step started from
//   k_tick_slots       step counter += 1; ping-pong index flips unless the
//                      sweep brought f back into the original pool
//   k_moments_cells    moments of the new f
// Kernels read each slot's step counter and ping-pong index from device
// memory, so the captured graph stays valid for every step until the set of
// resident slots changes.
// The Philox key is the slot's seed and the counter carries the slot's own
// cell index and step number, so a solve's samples do not depend on its slot,
// on the other residents or on the launch shape.

namespace dvm {

constexpr int kBatchAxis = 6 * kMaxAxis;  // doubles of one slot's packed axes + betas

struct SlotDesc {
  Grid g;                  // axes / betas point into the batch's axes pool
  long long off;           // first node of cell 0 in the node pools
  long long coff;          // first entry of cell 0 in the CDF pool ([2][N+1] per cell)
  long long samples;       // samples per cell per step (explicit sample mode)
  double dt, kn, fx, beta3;
  uint64_t seed;
  int cell0, ncell;        // first cell in the cell pools, cell count
  int full;                // the FP64 slice fits LDS: the window is the grid
  int wd[3];               // LDS window extents
  int split;               // work items per cell (sample ranges)
  int convect;             // 1-D sweep after the collision
  int bc[2];               // 1-D boundary types (left, right): in/out flow or zero gradient
  int order;               // 1-D sweep order
};

struct SlotState {
  int ntime;               // steps done
  int cur;                 // pool holding the current f
};

struct CellRef {
  int slot, cell;
};

struct WorkItem {
  int slot, cell, q0, q1;  // samples [q0, q1) of the cell (q1 clamped to its count)
};

// The collision work items form kXcds queues, one per XCD: queue x is
// items[qstart[x], qstart[x + 1]) and its head is queue[x].
constexpr int kXcds = 8;

struct BatchPools {
  double* f[2];
  double *eqv, *dacc, *cdf, *eqpar, *props, *factors, *inflow;
  int2* tabs;
  int *counts, *err, *queue, *qstart;
  unsigned char* sgn;  // deviation sign bytes: a cell's at the byte offset of its CDF slice
  double* cmat;        // FP32 path: per cell the conservation factor (kCmat doubles), else null
};

__device__ __forceinline__ size_t cell_node0(const SlotDesc& d, int cell) { return (size_t)d.off + (size_t)cell * d.g.N; }

__global__ void k_zero_int(int* __restrict__ p, int n) {
  for (int q = threadIdx.x; q < n; q += blockDim.x) p[q] = 0;
}

template <int BS, int V, bool kFused>
__global__ void __launch_bounds__(BS) k_cdf_cells(const SlotDesc* __restrict__ slots,
                                                  const SlotState* __restrict__ state,
                                                  const CellRef* __restrict__ cells, CdfParams base,
                                                  BatchPools P, int sum_cap, int eq_mode) {
  __shared__ CdfShared<BS, V> sh;
  extern __shared__ float gdyn[];  // sum_cap floats: the scan's group ends
  const CellRef r = cells[blockIdx.x];
  const SlotDesc& d = slots[r.slot];
  const int cur = state[r.slot].cur;
  const int N = d.g.N;
  const size_t gc = (size_t)d.cell0 + r.cell;
  CdfParams cp = base;
  cp.dt = d.dt;
  cp.kn = d.kn;
  cp.beta3 = d.beta3;
  cp.samples_per_cell = d.samples;
  const size_t n0 = cell_node0(d, r.cell);
  const size_t cs = (size_t)d.coff + (size_t)r.cell * 2 * (N + 1);
  cdf_cell<BS, V, kFused>(d.g, cp, P.f[cur] + n0, P.props + gc * DVM_NPROP, P.eqv + n0, P.cdf + cs,
           P.tabs + gc * 2 * kNBucket, P.eqpar + gc * 5, P.counts + gc * 4, P.factors + gc * 3, sh,
           (2 * ((N + V - 1) / V) <= sum_cap) ? gdyn : nullptr, P.sgn + cs, eq_mode);
}

// fd = f (FP64) or dacc = 0 (FP32) over the cells that take global atomics.
__global__ void k_prep_cells(const SlotDesc* __restrict__ slots, const SlotState* __restrict__ state,
                             const CellRef* __restrict__ cells, int n_cells, int f32, BatchPools P) {
  for (int e = blockIdx.y; e < n_cells; e += gridDim.y) {
    const CellRef r = cells[e];
    const SlotDesc& d = slots[r.slot];
    if (d.full && d.split == 1) continue;
    const int cur = state[r.slot].cur;
    const int N = d.g.N;
    const size_t n0 = cell_node0(d, r.cell);
    const double* src = P.f[cur] + n0;
    double* dst = P.f[cur ^ 1] + n0;
    double* da = P.dacc + n0;
    for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < N; l += gridDim.x * blockDim.x) {
      if (f32) da[l] = 0.0;
      else dst[l] = src[l];
    }
  }
}

// Persistent collision workgroups: each takes the next work item of its home
// queue (the queue of the XCD it runs on: workgroups are dispatched to the
// XCDs round-robin) until that queue is drained, then of the other queues in
// turn; every workgroup reaches the exit once it has found all kXcds queues
// drained. A cell's items sit in one queue, so the workgroups that share its
// CDFs share one L2.
template <bool kF32, int BS, int K, bool kUni = false, bool kRemap32 = false, bool kFastOnly = false>
__global__ void __launch_bounds__(BS) k_collide_slots(const SlotDesc* __restrict__ slots,
                                                      const SlotState* __restrict__ state,
                                                      const WorkItem* __restrict__ items, CollParams base,
                                                      BatchPools P, int fast_lds) {
  extern __shared__ double dyn[];
  __shared__ ConserveShared<BS> csh;
  __shared__ int s_item;
  const int home = (int)(blockIdx.x % kXcds);
  int visited = 0;
  for (;;) {
    const int qx = (home + visited) % kXcds;
    if (threadIdx.x == 0) s_item = atomicAdd(P.queue + qx, 1);
    __syncthreads();
    // uniform across the workgroup: the item, its slot descriptor, pointers
    // and window bounds then live in scalar registers
    const int it = P.qstart[qx] + __builtin_amdgcn_readfirstlane(s_item);
    if (it >= P.qstart[qx + 1]) {
      __syncthreads();  // every lane has read s_item before the next head is taken
      if (++visited == kXcds) return;
      continue;
    }
    const WorkItem w = items[it];
    const SlotDesc& d = slots[w.slot];
    Grid g = d.g;
    const int N = g.N;
    const SlotState st = state[w.slot];
    const size_t gc = (size_t)d.cell0 + w.cell;
    const size_t n0 = cell_node0(d, w.cell);
    using RM = typename std::conditional<kF32 && kRemap32, float, double>::type;
    const Axes<RM> ax = stage_collision_axes<RM>(g, dyn);
    CollParams cp = base;
    cp.seed = d.seed;
    const int ntime = st.ntime + 1;
    const CellView cv = cell_view(P.f[st.cur] + n0, P.eqv + n0, P.cdf + d.coff + (size_t)w.cell * 2 * (N + 1),
                                  P.tabs + gc * 2 * kNBucket, N);
    Window win{};
    if (!d.full) win = make_window(g, d.wd, P.props + gc * DVM_NPROP);
    // the fast FP32 sampling for a cell whose slice (or window) and sign
    // words fit the launch's LDS (fast_lds bytes; 0 when the batch has no
    // such path)
    FastCell fcs{};
    const FastCell* fc = nullptr;
    const int wn = d.full ? N : d.wd[0] * d.wd[1] * d.wd[2];
    if (fast_lds > 0 && (long long)(kAxisSlots + wn) * 8 + 4 * ((N + 31) >> 5) <= (long long)fast_lds) {
      const size_t cs = (size_t)d.coff + (size_t)w.cell * 2 * (N + 1);
      fcs.sgn = P.sgn + cs;
      fcs.cm = P.cmat ? P.cmat + gc * kCmat : nullptr;
      fc = &fcs;
    }
    const int bad = collide_item<kF32, BS, K, kUni, kF32 && kRemap32, PhiloxSrc, kFastOnly>(g, ax, cp, cv, w.cell, ntime, w.q0, w.q1, P.counts + gc * 4,
                                              P.factors + gc * 3, win, d.full != 0, d.split == 1, dyn + kAxisSlots,
                                              P.f[st.cur ^ 1] + n0, P.dacc + n0, &csh, PhiloxSrc{cp.seed}, fc);
    if (bad) atomicOr(P.err, bad);
    __syncthreads();  // LDS axes, accumulator and s_item are reused by the next item
  }
}

template <int BS>
__global__ void __launch_bounds__(BS) k_conserve_cells(const SlotDesc* __restrict__ slots,
                                                       const SlotState* __restrict__ state,
                                                       const CellRef* __restrict__ cells, BatchPools P) {
  __shared__ ConserveShared<BS> sh;
  const CellRef r = cells[blockIdx.x];
  const SlotDesc& d = slots[r.slot];
  if (d.full && d.split == 1) return;  // corrected inside the collision kernel
  const int cur = state[r.slot].cur;
  const size_t n0 = cell_node0(d, r.cell);
  const int flag = conserve_cell<BS>(d.g, P.f[cur] + n0, P.dacc + n0, P.f[cur ^ 1] + n0, sh);
  if (flag && threadIdx.x == 0) atomicOr(P.err, flag);
}

// 1-D sweep of the convecting slots: from the post-collision pool back into
// the pool the step started from. grid.x covers the largest grid's nodes.
__global__ void __launch_bounds__(kConvBS) k_convect1d_cells(const SlotDesc* __restrict__ slots,
                                                             const SlotState* __restrict__ state,
                                                             const CellRef* __restrict__ cells, int n_cells,
                                                             BatchPools P) {
  const int l = blockIdx.x * kConvBS + threadIdx.x;
  for (int e = blockIdx.y; e < n_cells; e += gridDim.y) {
    const CellRef r = cells[e];
    const SlotDesc& d = slots[r.slot];
    const int N = d.g.N;
    if (l >= N) continue;
    const int cur = state[r.slot].cur;
    int i, j, k;
    node_ijk(d.g, l, i, j, k);
    ConvParams p{};
    p.nx = d.ncell;
    p.fx = d.fx;
    p.acc = 1.0;
    p.bc[0] = d.bc[0];
    p.bc[1] = d.bc[1];
    p.order = d.order;
    convect1d_point(d.g, p, i, d.g.ax[0][i], P.f[cur ^ 1] + d.off, P.f[cur] + d.off, P.inflow + d.off, nullptr,
                    nullptr, nullptr, r.cell, l);
  }
}

__global__ void k_tick_slots(const SlotDesc* __restrict__ slots, SlotState* __restrict__ state,
                             const int* __restrict__ active, int n) {
  for (int q = threadIdx.x; q < n; q += blockDim.x) {
    const int s = active[q];
    state[s].ntime += 1;
    if (!slots[s].convect) state[s].cur ^= 1;
  }
}

__global__ void __launch_bounds__(kMomBS) k_moments_cells(const SlotDesc* __restrict__ slots,
                                                          const SlotState* __restrict__ state,
                                                          const CellRef* __restrict__ cells, double mass,
                                                          BatchPools P) {
  __shared__ double red[1][6][kMomBS / 64];
  __shared__ MomCmatShared<1, kMomBS> msh;
  const CellRef r = cells[blockIdx.x];
  const SlotDesc& d = slots[r.slot];
  const int cur = state[r.slot].cur;
  const size_t gc = (size_t)d.cell0 + r.cell;
  // the FP32 fast path's batches: also the next collision's conservation factor
  if (P.cmat)
    moments_cmat_cells<1, kMomBS>(d.g, P.f[cur] + cell_node0(d, r.cell), 0, 1, P.props + gc * DVM_NPROP,
                                  P.cmat + gc * kCmat, mass, msh);
  else
    moments_cell(d.g, P.f[cur] + cell_node0(d, r.cell), P.props + gc * DVM_NPROP, mass, red);
}

// Distance of each space-homogeneous slot's f from equilibrium (l1_cell).
// One workgroup per slot.
__global__ void __launch_bounds__(kMomBS) k_l1_slots(const SlotDesc* __restrict__ slots,
                                                     const SlotState* __restrict__ state,
                                                     const int* __restrict__ active, double mass, BatchPools P,
                                                     double* __restrict__ out) {
  __shared__ double red[kMomBS / 64];
  const int s = active[blockIdx.x];
  const SlotDesc& d = slots[s];
  const double e = l1_cell(d.g, P.f[state[s].cur] + d.off, P.props + (size_t)d.cell0 * DVM_NPROP, mass, red);
  if (threadIdx.x == 0) out[s] = e;
}

}  // namespace dvm

// ====================================================================== host

namespace {

// Best-fit allocator over [0, cap) with coalescing on release.
struct RangeAlloc {
  long long cap = 0;
  std::vector<std::pair<long long, long long>> free_;  // (start, length), sorted by start
  void init(long long c) {
    cap = c;
    free_.assign(1, {0, c});
  }
  long long alloc(long long n) {
    int best = -1;
    for (int q = 0; q < (int)free_.size(); ++q)
      if (free_[q].second >= n && (best < 0 || free_[q].second < free_[best].second)) best = q;
    if (best < 0) return -1;
    const long long s = free_[best].first;
    free_[best].first += n;
    free_[best].second -= n;
    if (free_[best].second == 0) free_.erase(free_.begin() + best);
    return s;
  }
  void release(long long s, long long n) {
    auto it = free_.begin();
    while (it != free_.end() && it->first < s) ++it;
    it = free_.insert(it, {s, n});
    auto nx = it + 1;
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = it - 1;
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  long long largest() const {
    long long m = 0;
    for (auto& r : free_) m = r.second > m ? r.second : m;
    return m;
  }
  long long total() const {
    long long t = 0;
    for (auto& r : free_) t += r.second;
    return t;
  }
};

struct HostSlot {
  bool active = false;
  bool uniform = false;          // equally spaced axes
  int N = 0, ncell = 0, dim = 0;
  long long nodes = 0, cdf = 0;  // pool lengths held
  long long samples = 0;
  int steps = 0;                 // steps launched since admission
  // 2-D slots that convect: the sweep's constants and the slot's own wall
  // buffers (wall Maxwellians [4][N], flux ratios [4][max(nx,ny)], specular
  // ghosts [4][max(nx,ny)][N] when acc < 1)
  bool conv2d = false;
  ConvParams cp{};
  double *wallM = nullptr, *ratio = nullptr, *spec = nullptr;
};

static void free_slot_walls(HostSlot& h) {
  for (double* p : {h.wallM, h.ratio, h.spec})
    if (p) (void)hipFree(p);
  h.wallM = h.ratio = h.spec = nullptr;
}

}  // namespace

struct dvm_batch {
  dvm_config species{};
  int max_slots = 0;
  long long node_cap = 0;
  int cell_cap = 0;
  hipStream_t stream = nullptr;
  int cus = 0;
  // device
  SlotDesc* d_slots = nullptr;
  SlotState* d_state = nullptr;
  double* d_axes = nullptr;
  BatchPools P{};
  CellRef *d_cells = nullptr, *d_cells1d = nullptr, *d_adm = nullptr;
  int* d_active = nullptr;
  int* d_active0 = nullptr;   // active space-homogeneous slots
  WorkItem* d_items = nullptr;
  long long items_cap = 0;
  double* d_l1 = nullptr;
  // host mirror
  std::vector<SlotDesc> slots;
  std::vector<HostSlot> hs;
  RangeAlloc nodes, cells, cdfs;
  CdfParams cdfp{};
  CollParams collp{};
  // launch state
  int n_cells = 0, n_cells1d = 0, n_active = 0, n_active0 = 0, n_items = 0, max_n = 0, max_n1d = 0;
  bool any_global = false;
  bool all_uniform = false;  // every resident slot's axes equally spaced
  size_t coll_lds = 0;
  bool fast_ok = false;      // the species and precision admit the fast FP32 sampling
  bool all_fast = false;     // every resident cell takes the fast FP32 sampling
  int fast_lds = 0;          // LDS bytes a cell's slice and sign bytes may take on the fast FP32 path (0: none)
  int persist = 0;
  int bs = 1024, k = 1, item_samples = 0;  // 0: sized from the resident load
  long long item_used = 0;
  hipGraphExec_t exec = nullptr;
  bool dirty = true;
  char msg[512] = {0};
};

static char g_batch_msg[512] = "";

static int bfail(dvm_batch* b, int code, const char* fmt, ...) {
  char* dst = b ? b->msg : g_batch_msg;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(dst, 512, fmt, ap);
  va_end(ap);
  return code;
}

#define BHIP(b, x)                                                                          \
  do {                                                                                      \
    hipError_t e_ = (x);                                                                    \
    if (e_ != hipSuccess) return bfail((b), DVM_E_HIP, "%s: %s", #x, hipGetErrorString(e_)); \
  } while (0)

template <class T>
static int balloc(dvm_batch* b, T** p, size_t n) {
  if (hipMalloc((void**)p, n * sizeof(T) + 16) != hipSuccess)
    return bfail(b, DVM_E_NOMEM, "hipMalloc of %zu bytes failed", n * sizeof(T));
  return DVM_OK;
}

// uni: FP32 with every resident slot's axes equally spaced (the remap then
// rounds to the nearest node, as a context with those axes does); r32: with
// uni, the remap in FP32 as well (DVM_PREC_FP32_REMAP32).
template <bool kF32, int BS>
static const void* batch_kernel_k(int k, bool uni, bool r32) {
  if (kF32 && uni && r32)
    return k == 2 ? (const void*)k_collide_slots<kF32, BS, 2, true, true>
                  : (const void*)k_collide_slots<kF32, BS, 1, true, true>;
  if (kF32 && uni)
    return k == 2 ? (const void*)k_collide_slots<kF32, BS, 2, true> : (const void*)k_collide_slots<kF32, BS, 1, true>;
  return k == 2 ? (const void*)k_collide_slots<kF32, BS, 2> : (const void*)k_collide_slots<kF32, BS, 1>;
}

static const void* batch_collide_kernel(bool f32, int bs, int k, bool uni, bool r32, bool all_fast = false) {
  // every item on the fast FP32 path: its own kernel (1024 lanes, one pick per lane)
  if (all_fast && f32 && uni && !r32 && bs == 1024 && k == 1)
    return (const void*)k_collide_slots<true, 1024, 1, true, false, true>;
  switch (bs) {
    case 256: return f32 ? batch_kernel_k<true, 256>(k, uni, r32) : batch_kernel_k<false, 256>(k, uni, r32);
    case 512: return f32 ? batch_kernel_k<true, 512>(k, uni, r32) : batch_kernel_k<false, 512>(k, uni, r32);
    case 768: return f32 ? batch_kernel_k<true, 768>(k, uni, r32) : batch_kernel_k<false, 768>(k, uni, r32);
    default: return f32 ? batch_kernel_k<true, 1024>(k, uni, r32) : batch_kernel_k<false, 1024>(k, uni, r32);
  }
}

static bool batch_r32(const dvm_batch* b) { return b->species.precision == DVM_PREC_FP32_REMAP32; }

static bool batch_f32(const dvm_batch* b) { return prec_f32(b->species.precision); }

extern "C" const char* dvm_batch_last_error(const dvm_batch* b) { return b ? b->msg : g_batch_msg; }

extern "C" void dvm_batch_destroy(dvm_batch* b) {
  if (!b) return;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  void* ptrs[] = {b->d_slots, b->d_state, b->d_axes, b->P.f[0], b->P.f[1], b->P.eqv, b->P.dacc, b->P.cdf,
                  b->P.eqpar, b->P.props, b->P.factors, b->P.inflow, b->P.tabs, b->P.counts, b->P.err, b->P.sgn,
                  b->P.cmat,
                  b->P.queue, b->P.qstart, b->d_cells, b->d_cells1d, b->d_adm, b->d_active, b->d_active0,
                  b->d_items, b->d_l1};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (HostSlot& h : b->hs) free_slot_walls(h);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

// Node-pool, cell-pool and CDF-pool lengths an instance occupies.
static void instance_extent(const dvm_config* c, long long& nodes, int& ncell, long long& cdf) {
  const long long N = (long long)c->nvx * c->nvy * c->nvz;
  ncell = c->nx * c->ny;
  nodes = N * ncell;
  cdf = 2 * (N + 1) * ncell;
}

static long long bytes_per_node(bool f32) { return (long long)sizeof(double) * (3 + 2 + (f32 ? 1 : 0) + 1); }

static int batch_create_impl(dvm_batch* b, const dvm_config* species, int max_slots, long long node_cap,
                             int cell_cap) {
  b->species = *species;
  b->max_slots = max_slots;
  b->node_cap = node_cap;
  b->cell_cap = cell_cap;
  BHIP(b, hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  int dev = 0;
  BHIP(b, hipGetDevice(&dev));
  BHIP(b, hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, dev));
  species_params(species, 1.0, b->cdfp, b->collp);
  const bool f32 = batch_f32(b);
  const size_t NC = (size_t)node_cap, CC = (size_t)cell_cap;
  int rc;
  if ((rc = balloc(b, &b->d_slots, max_slots)) || (rc = balloc(b, &b->d_state, max_slots)) ||
      (rc = balloc(b, &b->d_axes, (size_t)max_slots * kBatchAxis)) || (rc = balloc(b, &b->P.f[0], NC)) ||
      (rc = balloc(b, &b->P.f[1], NC)) || (rc = balloc(b, &b->P.eqv, NC)) || (rc = balloc(b, &b->P.inflow, NC)) ||
      (rc = balloc(b, &b->P.cdf, 2 * (NC + CC))) || (rc = balloc(b, &b->P.eqpar, CC * 5)) ||
      (rc = balloc(b, &b->P.props, CC * DVM_NPROP)) || (rc = balloc(b, &b->P.factors, CC * 3)) ||
      (rc = balloc(b, &b->P.tabs, CC * 2 * kNBucket)) || (rc = balloc(b, &b->P.counts, CC * 4)) ||
      (rc = balloc(b, &b->P.err, 4)) || (rc = balloc(b, &b->P.queue, kXcds)) ||
      (rc = balloc(b, &b->P.qstart, kXcds + 1)) || (rc = balloc(b, &b->d_cells, CC)) ||
      (rc = balloc(b, &b->d_cells1d, CC)) || (rc = balloc(b, &b->d_adm, CC)) ||
      (rc = balloc(b, &b->d_active, max_slots)) || (rc = balloc(b, &b->d_active0, max_slots)) ||
      (rc = balloc(b, &b->d_l1, max_slots)))
    return rc;
  if (f32 && (rc = balloc(b, &b->P.dacc, NC))) return rc;
  if ((rc = balloc(b, &b->P.sgn, 2 * (NC + CC)))) return rc;
  b->P.cmat = nullptr;
  // the fast FP32 sampling: equally spaced axes, the FP64 remap, isotropic
  // scattering, whole cells in LDS; the moments kernel computes its
  // conservation factors (after admission and after every step, so they are
  // those of the f each collision starts from)
  b->fast_ok = f32 && !batch_r32(b) && !b->collp.vss;
  if (b->fast_ok && (rc = balloc(b, &b->P.cmat, CC * kCmat))) return rc;
  BHIP(b, hipMemset(b->P.err, 0, 4 * sizeof(int)));
  BHIP(b, hipMemset(b->P.counts, 0, CC * 4 * sizeof(int)));
  BHIP(b, hipMemset(b->d_state, 0, (size_t)max_slots * sizeof(SlotState)));
  // the full Monte Carlo solver never writes eqv: its deviation signs are dsgn(f - 0)
  BHIP(b, hipMemset(b->P.eqv, 0, NC * sizeof(double)));
  b->slots.assign(max_slots, SlotDesc{});
  b->hs.assign(max_slots, HostSlot{});
  b->nodes.init(node_cap);
  b->cells.init(cell_cap);
  b->cdfs.init(2 * (node_cap + cell_cap));
  const int lds_max = (int)(kLdsBytes - kStaticLds);
  for (int bs : {256, 512, 768, 1024})
    for (int k : {1, 2})
      for (bool uni : {false, true})
        for (bool af : {false, true})
          BHIP(b, hipFuncSetAttribute(batch_collide_kernel(f32, bs, k, uni, batch_r32(b), af),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
  return DVM_OK;
}

extern "C" int dvm_batch_create(const dvm_config* species, int32_t max_slots, int64_t node_capacity,
                                int32_t cell_capacity, dvm_batch** out) {
  if (!out) return bfail(nullptr, DVM_E_ARG, "null output pointer");
  *out = nullptr;
  if (!species) return bfail(nullptr, DVM_E_ARG, "null species config");
  if (max_slots < 1 || node_capacity < 1 || cell_capacity < 1)
    return bfail(nullptr, DVM_E_ARG, "max_slots, node_capacity and cell_capacity must be >= 1");
  if (!(species->mass > 0.0)) return bfail(nullptr, DVM_E_ARG, "mass must be > 0");
  if (species->num_repl < 1) return bfail(nullptr, DVM_E_ARG, "num_repl must be >= 1");
  if (species->coll_solver != DVM_COLL_VR && species->coll_solver != DVM_COLL_FMC)
    return bfail(nullptr, DVM_E_ARG, "batched solves use the variance-reduction or full Monte Carlo collision solver");
  if (species->precision != DVM_PREC_FP64 && !prec_f32(species->precision))
    return bfail(nullptr, DVM_E_ARG, "precision must be DVM_PREC_FP64, DVM_PREC_FP32 or DVM_PREC_FP32_REMAP32");
  if (species->sample_mode != DVM_SAMPLES_REFERENCE && species->sample_mode != DVM_SAMPLES_EXPLICIT)
    return bfail(nullptr, DVM_E_ARG, "invalid sample_mode");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return bfail(nullptr, DVM_E_HIP, "no HIP device is visible");
  dvm_batch* b = new dvm_batch();
  const int rc = batch_create_impl(b, species, max_slots, node_capacity, cell_capacity);
  if (rc) {
    snprintf(g_batch_msg, sizeof g_batch_msg, "%s", b->msg);
    dvm_batch_destroy(b);
    return rc;
  }
  *out = b;
  return DVM_OK;
}

extern "C" int dvm_batch_footprint(const dvm_config* inst, int32_t precision, int64_t* nodes, int64_t* bytes) {
  if (!inst) return bfail(nullptr, DVM_E_ARG, "null config");
  long long nn, cdf;
  int nc;
  instance_extent(inst, nn, nc, cdf);
  if (nodes) *nodes = nn;
  if (bytes) {
    *bytes = nn * bytes_per_node(prec_f32(precision)) + (long long)nc * (16 * kNBucket + 8 * 20 + 16);
    if (inst->dim == DVM_DIM_2D && inst->enable_convection) {
      // the slot's wall buffers
      const long long N = (long long)inst->nvx * inst->nvy * inst->nvz;
      const long long along = inst->nx > inst->ny ? inst->nx : inst->ny;
      *bytes += 8 * (4 * N + 4 * along + (inst->acc_coeff < 1.0 ? 4 * along * N : 0));
    }
  }
  return DVM_OK;
}

static int check_instance(dvm_batch* b, const dvm_config* inst) {
  if (int rc = validate(inst)) return bfail(b, rc, "%s", g_create_msg);
  if (inst->dim != DVM_DIM_0D && inst->dim != DVM_DIM_1D && inst->dim != DVM_DIM_2D)
    return bfail(b, DVM_E_ARG, "batched solves are space-homogeneous, 1-D or 2-D");
  const dvm_config& s = b->species;
  if (inst->mass != s.mass || inst->diam != s.diam || inst->omega != s.omega || inst->alpha != s.alpha ||
      inst->csection_coeff != s.csection_coeff || inst->viscosity != s.viscosity ||
      inst->t_viscosity != s.t_viscosity || inst->coll_solver != s.coll_solver ||
      inst->cdf_tolerance != s.cdf_tolerance || inst->num_repl != s.num_repl || inst->coln_rms != s.coln_rms ||
      inst->sample_mode != s.sample_mode || inst->precision != s.precision)
    return bfail(b, DVM_E_ARG, "instance species / collision settings differ from the batch's");
  if (inst->dim == DVM_DIM_1D && inst->enable_convection &&
      (inst->bc_left == DVM_BC_WALL || inst->bc_right == DVM_BC_WALL))
    return bfail(b, DVM_E_ARG, "batched 1-D solves take in/out flow or zero-gradient ends (1-D walls: dvm_ctx)");
  if (inst->sample_mode == DVM_SAMPLES_EXPLICIT && inst->samples_per_cell > kMaxCellSamples)
    return bfail(b, DVM_E_ARG, "samples_per_cell must be <= 2^30");
  return DVM_OK;
}

extern "C" int dvm_batch_admit(dvm_batch* b, const dvm_config* inst, const double* f0, int32_t* slot_out) {
  if (!b || !inst || !f0 || !slot_out) return bfail(b, DVM_E_ARG, "null argument");
  *slot_out = -1;
  if (int rc = check_instance(b, inst)) return rc;
  int slot = -1;
  for (int q = 0; q < b->max_slots; ++q)
    if (!b->hs[q].active) { slot = q; break; }
  if (slot < 0) return bfail(b, DVM_E_NOMEM, "all %d slots are in use", b->max_slots);
  long long nn, ncdf;
  int nc;
  instance_extent(inst, nn, nc, ncdf);
  const long long off = b->nodes.alloc(nn);
  if (off < 0) return bfail(b, DVM_E_NOMEM, "node pool: %lld nodes requested, largest free range %lld", nn, b->nodes.largest());
  const long long c0 = b->cells.alloc(nc);
  if (c0 < 0) {
    b->nodes.release(off, nn);
    return bfail(b, DVM_E_NOMEM, "cell pool: %d cells requested, largest free range %lld", nc, b->cells.largest());
  }
  const long long coff = b->cdfs.alloc(ncdf);
  if (coff < 0) {
    b->nodes.release(off, nn);
    b->cells.release(c0, nc);
    return bfail(b, DVM_E_NOMEM, "CDF pool exhausted");
  }
  HostGrid hg;
  build_host_grid(inst, hg);
  double* ax = b->d_axes + (size_t)slot * kBatchAxis;
  BHIP(b, hipMemcpyAsync(ax, hg.packed.data(), hg.packed.size() * sizeof(double), hipMemcpyHostToDevice, b->stream));
  CdfParams cp;
  CollParams co;
  species_params(inst, hg.b3min, cp, co);
  SlotDesc d{};
  d.g = hg.g;
  place_grid(d.g, ax);
  d.off = off;
  d.coff = coff;
  d.samples = inst->samples_per_cell;
  d.dt = inst->dt;
  d.kn = inst->kn;
  d.beta3 = cp.beta3;
  d.seed = inst->seed;
  d.cell0 = (int)c0;
  d.ncell = nc;
  // a slot on the fast FP32 sampling keeps room for its sign words beside
  // its window
  const int N = hg.g.N;
  d.full = window_dims(hg.g.n, d.wd, (b->fast_ok && hg.uniform) ? fast_window_cap(N) : kWindowCap) ? 1 : 0;
  d.split = 1;
  // a convecting slot's sweep brings f back into the pool the step started
  // from, so its ping-pong index stays 0 (pool 0 holds f, pool 1 the
  // post-collision f) for its whole residence
  d.convect = (inst->dim != DVM_DIM_0D && inst->enable_convection) ? 1 : 0;
  d.fx = inst->dt / (inst->dx > 0.0 ? inst->dx : 1.0);
  d.bc[0] = inst->bc_left;
  d.bc[1] = inst->bc_right;
  d.order = inst->conv_order == 2 ? 2 : 1;
  HostSlot fresh{};
  if (inst->dim == DVM_DIM_2D && d.convect) {
    fresh.conv2d = true;
    conv_params(inst, hg, fresh.cp);
    std::vector<double> wm;
    conv2d_walls(inst, hg, fresh.cp, wm);
    const size_t along = (size_t)fresh.cp.max_along;
    int rc = balloc(b, &fresh.wallM, (size_t)4 * N);
    if (!rc) rc = balloc(b, &fresh.ratio, 4 * along);
    if (!rc && inst->acc_coeff < 1.0) rc = balloc(b, &fresh.spec, 4 * along * N);
    if (rc) {
      free_slot_walls(fresh);
      b->nodes.release(off, nn);
      b->cells.release(c0, nc);
      b->cdfs.release(coff, ncdf);
      return rc;
    }
    BHIP(b, hipMemcpy(fresh.wallM, wm.data(), wm.size() * sizeof(double), hipMemcpyHostToDevice));
    BHIP(b, hipMemset(fresh.ratio, 0, 4 * along * sizeof(double)));
  }
  if (d.convect && inst->dim == DVM_DIM_1D) {
    std::vector<double> inf;
    inflow_maxwellians(inst, hg, inf);  // 2N <= nx*N: the slot's own range of the inflow pool
    BHIP(b, hipMemcpyAsync(b->P.inflow + off, inf.data(), inf.size() * sizeof(double), hipMemcpyHostToDevice,
                           b->stream));
  }
  b->slots[slot] = d;
  BHIP(b, hipMemcpyAsync(b->d_slots + slot, &b->slots[slot], sizeof(SlotDesc), hipMemcpyHostToDevice, b->stream));
  const SlotState st{0, 0};
  BHIP(b, hipMemcpyAsync(b->d_state + slot, &st, sizeof st, hipMemcpyHostToDevice, b->stream));
  BHIP(b, hipMemcpyAsync(b->P.f[0] + off, f0, (size_t)nn * sizeof(double), hipMemcpyHostToDevice, b->stream));
  BHIP(b, hipMemsetAsync(b->P.eqpar + c0 * 5, 0, (size_t)nc * 5 * sizeof(double), b->stream));
  BHIP(b, hipMemsetAsync(b->P.counts + c0 * 4, 0, (size_t)nc * 4 * sizeof(int), b->stream));
  // initial moments of the admitted cells
  std::vector<CellRef> refs(nc);
  for (int c = 0; c < nc; ++c) refs[c] = CellRef{slot, c};
  BHIP(b, hipMemcpyAsync(b->d_adm, refs.data(), refs.size() * sizeof(CellRef), hipMemcpyHostToDevice, b->stream));
  hipLaunchKernelGGL(k_moments_cells, dim3(nc), dim3(kMomBS), 0, b->stream, b->d_slots, b->d_state, b->d_adm,
                     b->species.mass, b->P);
  BHIP(b, hipGetLastError());
  BHIP(b, hipStreamSynchronize(b->stream));  // host vectors above are freed on return
  HostSlot& h = b->hs[slot];
  h = fresh;
  h.active = true;
  h.uniform = hg.uniform;
  h.N = N;
  h.ncell = nc;
  h.dim = inst->dim;
  h.nodes = nn;
  h.cdf = ncdf;
  h.samples = inst->samples_per_cell;
  h.steps = 0;
  b->dirty = true;
  *slot_out = slot;
  return DVM_OK;
}

extern "C" int dvm_batch_retire(dvm_batch* b, int32_t slot) {
  if (!b) return bfail(b, DVM_E_ARG, "null batch");
  if (slot < 0 || slot >= b->max_slots || !b->hs[slot].active) return bfail(b, DVM_E_ARG, "slot %d is not resident", slot);
  // queued steps may still read the slot's ranges
  BHIP(b, hipStreamSynchronize(b->stream));
  HostSlot& h = b->hs[slot];
  const SlotDesc& d = b->slots[slot];
  b->nodes.release(d.off, h.nodes);
  b->cells.release(d.cell0, h.ncell);
  b->cdfs.release(d.coff, h.cdf);
  free_slot_walls(h);
  h = HostSlot{};
  b->dirty = true;
  return DVM_OK;
}

// Work lists, work items and the persistent grid for the resident slots.
static int batch_lists(dvm_batch* b) {
  std::vector<CellRef> cl, cl1;
  std::vector<int> act, act0;
  // work items of each cell, contiguous in `items`
  struct CellWork {
    long long cost;
    int first, count;
  };
  std::vector<WorkItem> items;
  std::vector<CellWork> cw;
  const bool explicit_mode = b->species.sample_mode == DVM_SAMPLES_EXPLICIT;
  b->max_n = 0;
  b->max_n1d = 0;
  b->any_global = false;
  int lds_n = 0;
  // Samples per work item: the tile when set; else about four items per CU
  // over the resident samples (never below 2^15: an item that splits a cell
  // flushes its window with one atomic per node; never above 2^18, below),
  // so that a few large solves still spread over the whole chip.
  long long step = b->item_samples;
  if (step <= 0) {
    long long total = 0;
    for (int s = 0; s < b->max_slots; ++s)
      if (b->hs[s].active) total += (explicit_mode ? b->hs[s].samples : 0) * b->hs[s].ncell;
    step = total / (4LL * (b->cus > 0 ? b->cus : 256));
    if (step < (1LL << 15)) step = 1LL << 15;
    // and at most 2^18, so that a large cell's items are taken by several
    // workgroups of its XCD, which share its CDFs in L2 (the autotuner's
    // choice on the 4096-instance batch)
    if (step > (1LL << 18)) step = 1LL << 18;
  }
  b->item_used = step;
  b->all_uniform = true;
  for (int s = 0; s < b->max_slots; ++s) {
    const HostSlot& h = b->hs[s];
    if (!h.active) continue;
    b->all_uniform = b->all_uniform && h.uniform;
    SlotDesc& d = b->slots[s];
    act.push_back(s);
    if (h.dim == DVM_DIM_0D) act0.push_back(s);
    b->max_n = h.N > b->max_n ? h.N : b->max_n;
    const int wn = d.wd[0] * d.wd[1] * d.wd[2];
    lds_n = wn > lds_n ? wn : lds_n;
    // sample ranges of item_samples per work item (the count is known up
    // front only in the explicit sample mode); the last range is open-ended
    const long long est = explicit_mode ? h.samples : (long long)h.N;
    d.split = (explicit_mode && est > step) ? (int)((est + step - 1) / step) : 1;
    if (!d.full || d.split > 1) b->any_global = true;
    for (int c = 0; c < h.ncell; ++c) {
      cl.push_back(CellRef{s, c});
      if (d.convect && h.dim == DVM_DIM_1D) cl1.push_back(CellRef{s, c});
      cw.push_back(CellWork{est, (int)items.size(), d.split});
      for (int q = 0; q < d.split; ++q) {
        const long long q0 = (long long)q * step;
        const bool last = q == d.split - 1;
        items.push_back(WorkItem{s, c, (int)q0, last ? 0x7fffffff : (int)(q0 + step)});
      }
    }
    if (d.convect && h.dim == DVM_DIM_1D) b->max_n1d = h.N > b->max_n1d ? h.N : b->max_n1d;
  }
  // cells to the XCD queues, costliest first, each to the least loaded queue
  // (longest-processing-time balancing); a queue lists its cells' items in
  // that order, so the costliest items run first
  std::stable_sort(cw.begin(), cw.end(), [](const CellWork& x, const CellWork& y) { return x.cost > y.cost; });
  std::vector<int> home(cw.size());
  {
    long long load[kXcds] = {0};
    for (size_t q = 0; q < cw.size(); ++q) {
      int x = 0;
      for (int y = 1; y < kXcds; ++y)
        if (load[y] < load[x]) x = y;
      home[q] = x;
      load[x] += cw[q].cost;
    }
  }
  std::vector<WorkItem> wi;
  wi.reserve(items.size());
  int qstart[kXcds + 1];
  for (int x = 0; x < kXcds; ++x) {
    qstart[x] = (int)wi.size();
    for (size_t q = 0; q < cw.size(); ++q)
      if (home[q] == x) wi.insert(wi.end(), items.begin() + cw[q].first, items.begin() + cw[q].first + cw[q].count);
  }
  qstart[kXcds] = (int)wi.size();
  if ((long long)wi.size() > b->items_cap) {
    if (b->d_items) BHIP(b, hipFree(b->d_items));
    b->d_items = nullptr;
    b->items_cap = 0;
    const long long cap = (long long)wi.size() * 2 + 64;
    if (int rc = balloc(b, &b->d_items, (size_t)cap)) return rc;
    b->items_cap = cap;
  }
  if (!cl.empty()) BHIP(b, hipMemcpyAsync(b->d_cells, cl.data(), cl.size() * sizeof(CellRef), hipMemcpyHostToDevice, b->stream));
  if (!cl1.empty())
    BHIP(b, hipMemcpyAsync(b->d_cells1d, cl1.data(), cl1.size() * sizeof(CellRef), hipMemcpyHostToDevice, b->stream));
  if (!act.empty()) BHIP(b, hipMemcpyAsync(b->d_active, act.data(), act.size() * sizeof(int), hipMemcpyHostToDevice, b->stream));
  if (!act0.empty())
    BHIP(b, hipMemcpyAsync(b->d_active0, act0.data(), act0.size() * sizeof(int), hipMemcpyHostToDevice, b->stream));
  if (!wi.empty()) BHIP(b, hipMemcpyAsync(b->d_items, wi.data(), wi.size() * sizeof(WorkItem), hipMemcpyHostToDevice, b->stream));
  BHIP(b, hipMemcpyAsync(b->P.qstart, qstart, sizeof(qstart), hipMemcpyHostToDevice, b->stream));
  // descriptors carry the split chosen above
  BHIP(b, hipMemcpyAsync(b->d_slots, b->slots.data(), b->slots.size() * sizeof(SlotDesc), hipMemcpyHostToDevice,
                         b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  b->n_cells = (int)cl.size();
  b->n_cells1d = (int)cl1.size();
  b->n_active = (int)act.size();
  b->n_active0 = (int)act0.size();
  b->n_items = (int)wi.size();
  b->coll_lds = (size_t)(kAxisSlots + lds_n) * sizeof(double);
  // the fast FP32 path also stages a cell's sign words after its slice or
  // window; it applies to the slots for which the two fit
  b->fast_lds = 0;
  b->all_fast = false;
  if (b->fast_ok && b->all_uniform) {
    size_t need = b->coll_lds;
    bool all = true;
    for (int s = 0; s < b->max_slots; ++s) {
      const HostSlot& h = b->hs[s];
      if (!h.active) continue;
      const int* w = b->slots[s].wd;
      const size_t n = (size_t)(kAxisSlots + w[0] * w[1] * w[2]) * sizeof(double) + (size_t)sign_lds_bytes(h.N);
      if (n > kLdsBytes - kStaticLds) {
        all = false;
        continue;
      }
      if (n > need) need = n;
    }
    b->coll_lds = need;
    b->fast_lds = (int)need;
    b->all_fast = all;
  }
  const void* kern = batch_collide_kernel(batch_f32(b), b->bs, b->k, b->all_uniform, batch_r32(b), b->all_fast);
  int per_cu = 0;
  BHIP(b, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, b->bs, b->coll_lds));
  if (per_cu < 1) per_cu = 1;
  long long grid = (long long)per_cu * b->cus;
  if (grid > b->n_items) grid = b->n_items;
  b->persist = (int)(grid < 1 ? 1 : grid);
  return DVM_OK;
}

// Enqueue one step of every resident slot on the batch stream.
static int batch_enqueue_step(dvm_batch* b) {
  const bool f32 = batch_f32(b);
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, b->stream, b->P.queue, kXcds);
  // f_eq: stored and read back on a cache hit (1) unless the batch may take
  // the fast FP32 sampling, whose collision never reads it: then computed
  // again on a hit, and stored (2) only while some item takes the general
  // sampling, which reads it; neither CDF kernel then reads a stored f_eq,
  // whichever of them earlier steps used
  const int eq_mode = !b->fast_ok ? 1 : (b->all_fast ? 0 : 2);
  if (b->n_cells >= 4 * b->cus) {
    const int cap = cdf_sum_cap((int)b->max_n, kCdfV);
    hipLaunchKernelGGL((k_cdf_cells<kCdfBS, kCdfV, true>), dim3(b->n_cells), dim3(kCdfBS), cap * sizeof(float),
                       b->stream, b->d_slots, b->d_state, b->d_cells, b->cdfp, b->P, cap, eq_mode);
  } else {
    const int cap = cdf_sum_cap((int)b->max_n, kCdfWideV);
    hipLaunchKernelGGL((k_cdf_cells<kCdfWideBS, kCdfWideV, false>), dim3(b->n_cells), dim3(kCdfWideBS), cap * sizeof(float),
                       b->stream, b->d_slots, b->d_state, b->d_cells, b->cdfp, b->P, cap, eq_mode);
  }
  if (b->any_global) {
    const unsigned chunks = (unsigned)((b->max_n + 255) / 256);
    const unsigned ys = (unsigned)(b->n_cells < 65535 ? b->n_cells : 65535);
    hipLaunchKernelGGL(k_prep_cells, dim3(chunks, ys), dim3(256), 0, b->stream, b->d_slots, b->d_state, b->d_cells,
                       b->n_cells, f32 ? 1 : 0, b->P);
  }
  {
    const SlotDesc* sl = b->d_slots;
    const SlotState* st = b->d_state;
    const WorkItem* it = b->d_items;
    CollParams cp = b->collp;
    BatchPools P = b->P;
    int fl = b->fast_lds;
    void* args[] = {&sl, &st, &it, &cp, &P, &fl};
    BHIP(b, hipLaunchKernel(batch_collide_kernel(f32, b->bs, b->k, b->all_uniform, batch_r32(b), b->all_fast), dim3(b->persist), dim3(b->bs), args, b->coll_lds,
                            b->stream));
  }
  if (f32)
    hipLaunchKernelGGL(k_conserve_cells<256>, dim3(b->n_cells), dim3(256), 0, b->stream, b->d_slots, b->d_state,
                       b->d_cells, b->P);
  if (b->n_cells1d > 0) {
    const unsigned chunks = (unsigned)((b->max_n1d + kConvBS - 1) / kConvBS);
    const unsigned ys = (unsigned)(b->n_cells1d < 65535 ? b->n_cells1d : 65535);
    hipLaunchKernelGGL(k_convect1d_cells, dim3(chunks, ys), dim3(kConvBS), 0, b->stream, b->d_slots, b->d_state,
                       b->d_cells1d, b->n_cells1d, b->P);
  }
  // 2-D slots: walls and sweep of each slot's own grid, from the
  // post-collision pool 1 into pool 0 (a convecting slot's index stays 0)
  for (int s = 0; s < b->max_slots; ++s) {
    const HostSlot& h = b->hs[s];
    if (!h.active || !h.conv2d) continue;
    const SlotDesc& d = b->slots[s];
    const ConvParams& p = h.cp;
    const double* fc = b->P.f[1] + d.off;
    double* fo = b->P.f[0] + d.off;
    if (p.acc > 0.0)
      hipLaunchKernelGGL(k_wall_ratio, dim3(p.max_along, 4), dim3(kWallBS), 0, b->stream, d.g, p, fc, h.ratio);
    if (p.acc < 1.0) {
      hipLaunchKernelGGL(k_fill, dim3(1024), dim3(256), 0, b->stream, h.spec, (long long)4 * p.max_along * h.N, 0.0);
      hipLaunchKernelGGL(k_spec_ghost, dim3((h.N + 255) / 256, p.max_along, 4), dim3(256), 0, b->stream, d.g, p, fc,
                         h.spec, b->P.err);
    }
    launch_convect2d(d.g, p, fc, fo, h.wallM, h.ratio, h.spec, b->stream);
  }
  hipLaunchKernelGGL(k_tick_slots, dim3(1), dim3(256), 0, b->stream, b->d_slots, b->d_state, b->d_active, b->n_active);
  hipLaunchKernelGGL(k_moments_cells, dim3(b->n_cells), dim3(kMomBS), 0, b->stream, b->d_slots, b->d_state,
                     b->d_cells, b->species.mass, b->P);
  BHIP(b, hipGetLastError());
  return DVM_OK;
}

static int batch_capture(dvm_batch* b) {
  if (b->exec) {
    BHIP(b, hipGraphExecDestroy(b->exec));
    b->exec = nullptr;
  }
  if (int rc = batch_lists(b)) return rc;
  if (b->n_active == 0) {
    b->dirty = false;
    return DVM_OK;
  }
  hipGraph_t graph = nullptr;
  BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  const int rc = batch_enqueue_step(b);
  const hipError_t e = hipStreamEndCapture(b->stream, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return bfail(b, DVM_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
  const hipError_t ei = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) return bfail(b, DVM_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(ei));
  b->dirty = false;
  return DVM_OK;
}

static int batch_check_err(dvm_batch* b) {
  int h[4] = {0, 0, 0, 0};
  BHIP(b, hipMemcpyAsync(h, b->P.err, sizeof h, hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  if (h[0]) {
    BHIP(b, hipMemset(b->P.err, 0, sizeof h));
    char text[400];
    const int rc = device_error(h[0], text, sizeof text);
    return bfail(b, rc, "%s", text);
  }
  return DVM_OK;
}

extern "C" int dvm_batch_step(dvm_batch* b, int32_t nsteps) {
  if (!b) return bfail(b, DVM_E_ARG, "null batch");
  if (nsteps < 0) return bfail(b, DVM_E_ARG, "nsteps must be >= 0");
  if (b->dirty)
    if (int rc = batch_capture(b)) return rc;
  if (b->n_active == 0 || nsteps == 0) return DVM_OK;
  for (int q = 0; q < nsteps; ++q) BHIP(b, hipGraphLaunch(b->exec, b->stream));
  for (int s = 0; s < b->max_slots; ++s)
    if (b->hs[s].active) b->hs[s].steps += nsteps;
  return DVM_OK;
}

extern "C" int dvm_batch_synchronize(dvm_batch* b) {
  if (!b) return bfail(b, DVM_E_ARG, "null batch");
  return batch_check_err(b);
}

static int slot_state(dvm_batch* b, int slot, SlotState& st) {
  BHIP(b, hipMemcpyAsync(&st, b->d_state + slot, sizeof st, hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  return DVM_OK;
}

extern "C" int dvm_batch_download_f(dvm_batch* b, int32_t slot, double* f) {
  if (!b || !f) return bfail(b, DVM_E_ARG, "null argument");
  if (slot < 0 || slot >= b->max_slots || !b->hs[slot].active) return bfail(b, DVM_E_ARG, "slot %d is not resident", slot);
  SlotState st;
  if (int rc = slot_state(b, slot, st)) return rc;
  const SlotDesc& d = b->slots[slot];
  BHIP(b, hipMemcpyAsync(f, b->P.f[st.cur] + d.off, (size_t)b->hs[slot].nodes * sizeof(double), hipMemcpyDeviceToHost,
                         b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  return batch_check_err(b);
}

extern "C" int dvm_batch_moments(dvm_batch* b, int32_t slot, double* props) {
  if (!b || !props) return bfail(b, DVM_E_ARG, "null argument");
  if (slot < 0 || slot >= b->max_slots || !b->hs[slot].active) return bfail(b, DVM_E_ARG, "slot %d is not resident", slot);
  const SlotDesc& d = b->slots[slot];
  BHIP(b, hipMemcpyAsync(props, b->P.props + (size_t)d.cell0 * DVM_NPROP,
                         (size_t)b->hs[slot].ncell * DVM_NPROP * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  return batch_check_err(b);
}

extern "C" int dvm_batch_l1(dvm_batch* b, double* l1) {
  if (!b || !l1) return bfail(b, DVM_E_ARG, "null argument");
  if (b->dirty)
    if (int rc = batch_capture(b)) return rc;
  std::vector<double> h(b->max_slots, -1.0);
  if (b->n_active0 > 0) {
    BHIP(b, hipMemcpyAsync(b->d_l1, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, b->stream));
    hipLaunchKernelGGL(k_l1_slots, dim3(b->n_active0), dim3(kMomBS), 0, b->stream, b->d_slots, b->d_state,
                       b->d_active0, b->species.mass, b->P, b->d_l1);
    BHIP(b, hipGetLastError());
    BHIP(b, hipMemcpyAsync(h.data(), b->d_l1, h.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
  }
  for (int s = 0; s < b->max_slots; ++s) l1[s] = h[s];
  return batch_check_err(b);
}

extern "C" int dvm_batch_slot_info(dvm_batch* b, int32_t slot, int32_t* steps, int64_t* nodes, int32_t* ncell,
                                   int32_t* in_lds) {
  if (!b) return bfail(b, DVM_E_ARG, "null batch");
  if (slot < 0 || slot >= b->max_slots || !b->hs[slot].active) return bfail(b, DVM_E_ARG, "slot %d is not resident", slot);
  const HostSlot& h = b->hs[slot];
  if (steps) *steps = h.steps;
  if (nodes) *nodes = h.nodes;
  if (ncell) *ncell = h.ncell;
  if (in_lds) *in_lds = b->slots[slot].full;
  return DVM_OK;
}

extern "C" int dvm_batch_capacity(dvm_batch* b, int64_t* free_nodes, int64_t* largest_free, int32_t* free_cells,
                                  int32_t* free_slots) {
  if (!b) return bfail(b, DVM_E_ARG, "null batch");
  int fs = 0;
  for (auto& h : b->hs) fs += h.active ? 0 : 1;
  if (free_nodes) *free_nodes = b->nodes.total();
  if (largest_free) *largest_free = b->nodes.largest();
  if (free_cells) *free_cells = (int32_t)b->cells.total();
  if (free_slots) *free_slots = fs;
  return DVM_OK;
}

extern "C" int dvm_batch_set_tiles(dvm_batch* b, int32_t lanes, int32_t pick_k, int32_t item_samples) {
  if (!b) return bfail(b, DVM_E_ARG, "null batch");
  if (lanes != 256 && lanes != 512 && lanes != 768 && lanes != 1024)
    return bfail(b, DVM_E_ARG, "lanes must be 256, 512, 768 or 1024");
  if (pick_k != 1 && pick_k != 2) return bfail(b, DVM_E_ARG, "pick_k must be 1 or 2");
  if (item_samples != 0 && item_samples < 1024)
    return bfail(b, DVM_E_ARG, "item_samples must be 0 (sized from the load) or >= 1024");
  BHIP(b, hipStreamSynchronize(b->stream));
  b->bs = lanes;
  b->k = pick_k;
  b->item_samples = item_samples;
  b->dirty = true;
  return DVM_OK;
}

extern "C" int dvm_batch_get_tiles(dvm_batch* b, int32_t* tiles) {
  if (!b || !tiles) return bfail(b, DVM_E_ARG, "null argument");
  tiles[0] = b->bs;
  tiles[1] = b->k;
  tiles[2] = (int32_t)(b->item_used < 0x7fffffff ? b->item_used : 0x7fffffff);
  tiles[3] = b->persist;
  return DVM_OK;
}

// Autotuner: times `steps` graph replays of the resident set for each
// candidate tile (lanes per workgroup x samples per lane in lockstep, then
// samples per global-atomic work item), restoring the resident state before
// every candidate and at the end, and keeps the fastest. The state a tile
// produces differs only by the order of floating-point atomic additions.
extern "C" int dvm_batch_tune(dvm_batch* b, int32_t steps, int32_t* tiles_out, double* ms_out) {
  if (!b) return bfail(b, DVM_E_ARG, "null batch");
  if (steps < 1) return bfail(b, DVM_E_ARG, "steps must be >= 1");
  if (b->dirty)
    if (int rc = batch_capture(b)) return rc;
  if (b->n_active == 0) return bfail(b, DVM_E_STATE, "no resident slots to tune on");
  BHIP(b, hipStreamSynchronize(b->stream));
  const size_t NC = (size_t)b->node_cap, CC = (size_t)b->cell_cap;
  double *s0 = nullptr, *s1 = nullptr, *se = nullptr, *sq = nullptr, *sp = nullptr;
  SlotState* ss = nullptr;
  int rc;
  auto release = [&]() {
    for (void* p : {(void*)s0, (void*)s1, (void*)se, (void*)sq, (void*)sp, (void*)ss})
      if (p) (void)hipFree(p);
  };
  if ((rc = balloc(b, &s0, NC)) || (rc = balloc(b, &s1, NC)) || (rc = balloc(b, &se, NC)) ||
      (rc = balloc(b, &sq, CC * 5)) ||
      (rc = balloc(b, &sp, CC * DVM_NPROP)) || (rc = balloc(b, &ss, (size_t)b->max_slots))) {
    release();
    return rc;
  }
  auto copy_all = [&](bool save) -> hipError_t {
    const hipMemcpyKind kd = hipMemcpyDeviceToDevice;
    hipError_t e = hipSuccess;
    auto cp = [&](void* dst, void* src, size_t n) {
      if (e == hipSuccess) e = hipMemcpyAsync(save ? dst : src, save ? src : dst, n, kd, b->stream);
    };
    cp(s0, b->P.f[0], NC * sizeof(double));
    cp(s1, b->P.f[1], NC * sizeof(double));
    cp(se, b->P.eqv, NC * sizeof(double));
    cp(sq, b->P.eqpar, CC * 5 * sizeof(double));
    cp(sp, b->P.props, CC * DVM_NPROP * sizeof(double));
    cp(ss, b->d_state, (size_t)b->max_slots * sizeof(SlotState));
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    return e;
  };
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = copy_all(true);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  const int old_bs = b->bs, old_k = b->k, old_item = b->item_samples;
  int best_bs = old_bs, best_k = old_k, best_item = old_item;
  double best_ms = 1e300;
  auto trial = [&](int bs, int k, int item) -> int {
    b->bs = bs;
    b->k = k;
    b->item_samples = item;
    if (int r = batch_capture(b)) return r;
    if ((e = copy_all(false)) != hipSuccess) return bfail(b, DVM_E_HIP, "tune restore: %s", hipGetErrorString(e));
    BHIP(b, hipGraphLaunch(b->exec, b->stream));  // warm-up replay
    BHIP(b, hipEventRecord(e0, b->stream));
    for (int q = 0; q < steps; ++q) BHIP(b, hipGraphLaunch(b->exec, b->stream));
    BHIP(b, hipEventRecord(e1, b->stream));
    BHIP(b, hipEventSynchronize(e1));
    float ms = 0.0f;
    BHIP(b, hipEventElapsedTime(&ms, e0, e1));
    const double per = (double)ms / steps;
    if (per < best_ms) {
      best_ms = per;
      best_bs = bs;
      best_k = k;
      best_item = item;
    }
    return DVM_OK;
  };
  rc = (e == hipSuccess) ? DVM_OK : bfail(b, DVM_E_HIP, "tune setup: %s", hipGetErrorString(e));
  for (int bs : {256, 512, 768, 1024})
    for (int k : {1, 2})
      if (rc == DVM_OK) rc = trial(bs, k, old_item);
  if (rc == DVM_OK && b->species.sample_mode == DVM_SAMPLES_EXPLICIT) {
    const int bb = best_bs, bk = best_k;
    for (int item : {0, 1 << 18, 1 << 20, 1 << 30})
      if (rc == DVM_OK && item != old_item) rc = trial(bb, bk, item);
  }
  b->bs = (rc == DVM_OK) ? best_bs : old_bs;
  b->k = (rc == DVM_OK) ? best_k : old_k;
  b->item_samples = (rc == DVM_OK) ? best_item : old_item;
  b->dirty = true;
  if (rc == DVM_OK) {
    rc = batch_capture(b);
    if (rc == DVM_OK && (e = copy_all(false)) != hipSuccess) rc = bfail(b, DVM_E_HIP, "tune restore: %s", hipGetErrorString(e));
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  release();
  if (rc) return rc;
  if (tiles_out) {
    tiles_out[0] = b->bs;
    tiles_out[1] = b->k;
    tiles_out[2] = b->item_samples;
  }
  if (ms_out) *ms_out = best_ms;
  return DVM_OK;
}
