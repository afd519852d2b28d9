"""Run configurations and initial distributions.

`SolverConfig` holds what the reference reads from its input deck and module
setters (read_input_2D.f90:75-777) for the hot path; `to_c()` packs it into the
`dvm_config` struct of include/dvm_hip.h. The workload builders below set up
the synthetic cases the time step is measured and checked on, in the
reference's nondimensional units (mass = diameter = 1 after scaling, Maxwellian
proportional to exp(-m|c-u|^2/T)).
"""
from dataclasses import dataclass, field, replace
import ctypes as C
import math

import numpy as np

from . import capi

PI_REF = 3.1459265358979323  # constants.f90:30


def uniform_axis(n, vmin, vmax):
    """velocity_grid.f90:636-648 generate_uniform_grid."""
    i = np.arange(n, dtype=np.float64)
    return vmin + i * (vmax - vmin) / float(n - 1)


def beta_axis(a):
    """velocity_grid.f90:486-540 compute_beta: half gaps, full gap at the ends."""
    a = np.asarray(a, dtype=np.float64)
    b = np.empty_like(a)
    b[0] = a[1] - a[0]
    b[-1] = a[-1] - a[-2]
    b[1:-1] = (a[2:] - a[1:-1]) / 2.0 + (a[1:-1] - a[:-2]) / 2.0
    return b


@dataclass
class SolverConfig:
    dim: int = capi.DVM_DIM_0D
    nx: int = 1
    ny: int = 1
    dx: float = 1.0
    dy: float = 1.0
    vx: np.ndarray = None
    vy: np.ndarray = None
    vz: np.ndarray = None
    dt: float = 0.05
    kn: float = 1.0
    mass: float = 1.0
    diam: float = 1.0
    omega: float = 0.5
    alpha: float = 1.0
    viscosity: float = 1.0
    t_viscosity: float = 1.0
    csection_coeff: int = capi.DVM_CS_CLARKE
    coll_solver: int = capi.DVM_COLL_VR
    cdf_tolerance: float = 0.01
    num_repl: int = 1
    coln_rms: float = 1.0
    sample_mode: int = capi.DVM_SAMPLES_EXPLICIT
    samples_per_cell: int = 1000
    seed: int = 12345
    precision: int = capi.DVM_PREC_FP64
    enable_convection: int = 0
    bc_left: int = capi.DVM_BC_WALL
    bc_right: int = capi.DVM_BC_WALL
    acc_coeff: float = 1.0
    wall: np.ndarray = field(default_factory=lambda: np.tile(np.array([1.0, 0.0, 0.0, 0.0, 1.0]), (4, 1)))
    conv_order: int = capi.DVM_CONV_UPWIND1
    name: str = "custom"

    @property
    def nvx(self):
        return len(self.vx)

    @property
    def nvy(self):
        return len(self.vy)

    @property
    def nvz(self):
        return len(self.vz)

    @property
    def num_points(self):
        return self.nvx * self.nvy * self.nvz

    def with_(self, **kw):
        return replace(self, **kw)

    def to_c(self):
        keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (self.vx, self.vy, self.vz)]
        c = capi.DvmConfig()
        c.dim, c.nx, c.ny, c.dx, c.dy = self.dim, self.nx, self.ny, self.dx, self.dy
        c.nvx, c.nvy, c.nvz = len(keep[0]), len(keep[1]), len(keep[2])
        c.vx, c.vy, c.vz = (a.ctypes.data_as(C.POINTER(C.c_double)) for a in keep)
        c.dt, c.kn = self.dt, self.kn
        c.mass, c.diam, c.omega, c.alpha = self.mass, self.diam, self.omega, self.alpha
        c.viscosity, c.t_viscosity = self.viscosity, self.t_viscosity
        c.csection_coeff, c.coll_solver = self.csection_coeff, self.coll_solver
        c.cdf_tolerance, c.num_repl, c.coln_rms = self.cdf_tolerance, self.num_repl, self.coln_rms
        c.sample_mode, c.samples_per_cell = self.sample_mode, int(self.samples_per_cell)
        c.seed, c.precision = int(self.seed) & ((1 << 64) - 1), self.precision
        c.enable_convection, c.bc_left, c.bc_right = self.enable_convection, self.bc_left, self.bc_right
        c.acc_coeff = self.acc_coeff
        w = np.asarray(self.wall, dtype=np.float64).reshape(4, 5)
        for s in range(4):
            for q in range(5):
                c.wall[s][q] = w[s, q]
        c.conv_order = self.conv_order
        return c, keep


# ----------------------------------------------------------------- initial data
def maxwellian(cfg, dens, u, v, w, temp, renormalise=True, pi=math.pi):
    """Discrete Maxwellian times beta^3 on the node-fastest layout; renormalised
    to `dens` as the reference's Maxwellian initial condition is."""
    x, y, z = cfg.vx, cfg.vy, cfg.vz
    bx, by, bz = beta_axis(x), beta_axis(y), beta_axis(z)
    m = cfg.mass
    coeff = dens * math.sqrt(m ** 3 / (pi ** 3 * temp ** 3))
    cx = (x - u) ** 2
    cy = (y - v) ** 2
    cz = (z - w) ** 2
    csq = cz[:, None, None] + cy[None, :, None] + cx[None, None, :]
    f = coeff * np.exp(-csq * m / temp) * (bz[:, None, None] * by[None, :, None] * bx[None, None, :])
    f = f.reshape(-1)
    if renormalise and f.sum() > 0:
        f *= dens / f.sum()
    return f


def node_velocities(cfg):
    """(N, 3) velocity of every node, i fastest."""
    z, y, x = np.meshgrid(cfg.vz, cfg.vy, cfg.vx, indexing="ij")
    return np.stack([x.reshape(-1), y.reshape(-1), z.reshape(-1)], axis=1)


# -------------------------------------------------------------------- workloads
def homogeneous_relaxation(n_v=21, vmax=5.0, samples=100_000, dt=0.05, seed=1, temp=1.0):
    """Space-homogeneous relaxation of two Maxwellians (equal density, T and 2T,
    mean velocities +-1 along x); hard spheres, Kn = 1."""
    ax = uniform_axis(n_v, -vmax, vmax)
    cfg = SolverConfig(dim=capi.DVM_DIM_0D, nx=1, ny=1, vx=ax, vy=ax.copy(), vz=ax.copy(), dt=dt, kn=1.0,
                       omega=0.5, alpha=1.0, sample_mode=capi.DVM_SAMPLES_EXPLICIT, samples_per_cell=samples,
                       seed=seed, enable_convection=0, name=f"homogeneous {n_v}^3, {samples} samples/step")
    f = maxwellian(cfg, 0.5, 1.0, 0.0, 0.0, temp) + maxwellian(cfg, 0.5, -1.0, 0.0, 0.0, 2.0 * temp)
    return cfg, f.reshape(1, -1)


def lid_driven_cavity(nx=64, ny=64, n_v=25, vmax=5.0, samples=20_000, kn=0.1, lid=0.3, t_wall=1.0,
                      cfl=0.5, seed=7, precision=capi.DVM_PREC_FP64):
    """2-D lid-driven cavity microflow: four diffuse walls at T_w, the top wall
    moving at `lid` along x, hard spheres, uniform equilibrium start. Unit box,
    dt from the 2-D CFL condition |c_x| + |c_y| <= cfl."""
    ax = uniform_axis(n_v, -vmax, vmax)
    dx, dy = 1.0 / nx, 1.0 / ny
    dt = cfl / (vmax / dx + vmax / dy)
    wall = np.array([[1.0, 0.0, 0.0, 0.0, t_wall]] * 4)
    wall[3, capi.W_U] = lid
    cfg = SolverConfig(dim=capi.DVM_DIM_2D, nx=nx, ny=ny, dx=dx, dy=dy, vx=ax, vy=ax.copy(), vz=ax.copy(),
                       dt=dt, kn=kn, omega=0.5, alpha=1.0, sample_mode=capi.DVM_SAMPLES_EXPLICIT,
                       samples_per_cell=samples, seed=seed, precision=precision, enable_convection=1,
                       bc_left=capi.DVM_BC_WALL, bc_right=capi.DVM_BC_WALL, acc_coeff=1.0, wall=wall,
                       name=f"cavity {nx}x{ny}, {n_v}^3, {samples} samples/cell")
    f0 = maxwellian(cfg, 1.0, 0.0, 0.0, 0.0, t_wall)
    return cfg, np.tile(f0, (nx * ny, 1))


def rankine_hugoniot(mach, gamma=5.0 / 3.0):
    """Density, velocity, temperature and pressure ratios across a normal shock
    (shock_conditions.f90:140-160)."""
    m2 = mach * mach
    dens = (gamma + 1.0) * m2 / ((gamma - 1.0) * m2 + 2.0)
    press = (2.0 * gamma * m2 - (gamma - 1.0)) / (gamma + 1.0)
    temp = press / dens
    return dens, 1.0 / dens, temp, press


def sound_speed(gamma, temp, mass=1.0):
    """sqrt(gamma T / (2 m)): the speed of sound in the reference's units, where
    the Maxwellian is exp(-m |c - u|^2 / T) (shock_conditions.f90:54-171; the
    shock calculator's u_up = M sqrt(gamma T / (2 m)) with the mixture's mass,
    software/shock_calculator/normal_shock_parameters.txt)."""
    return math.sqrt(gamma * temp / (2.0 * mass))


def normal_shock_1d(nx=200, n_v=31, vmax=6.0, mach=2.0, samples=50_000, kn=1.0, dx=0.1, cfl=0.5, seed=3,
                    omega=0.81, width=10.0, gamma=5.0 / 3.0):
    """1-D normal shock: upstream Maxwellian at Mach `mach` (T = 1), downstream
    state from Rankine-Hugoniot, tanh profile between them, fixed in/out flow
    at both ends, VHS omega with isotropic scattering; gamma = 5/3 for the
    monatomic gas."""
    ax = uniform_axis(n_v, -vmax, vmax)
    dt = cfl * dx / vmax
    u1 = mach * sound_speed(gamma, 1.0, 1.0)
    rd, ru, rt, _ = rankine_hugoniot(mach, gamma)
    wall = np.array([[1.0, u1, 0.0, 0.0, 1.0], [rd, u1 * ru, 0.0, 0.0, rt],
                     [1.0, 0.0, 0.0, 0.0, 1.0], [1.0, 0.0, 0.0, 0.0, 1.0]])
    cfg = SolverConfig(dim=capi.DVM_DIM_1D, nx=nx, ny=1, dx=dx, dy=1.0, vx=ax, vy=ax.copy(), vz=ax.copy(),
                       dt=dt, kn=kn, omega=omega, alpha=1.0, sample_mode=capi.DVM_SAMPLES_EXPLICIT,
                       samples_per_cell=samples, seed=seed, enable_convection=1, bc_left=capi.DVM_BC_INOUT,
                       bc_right=capi.DVM_BC_INOUT, acc_coeff=1.0, wall=wall,
                       name=f"1-D shock M{mach}, {nx} cells, {n_v}^3")
    xc = (np.arange(nx) + 0.5) - nx / 2.0
    s = 0.5 * (1.0 + np.tanh(xc / (width / 2.0)))
    f = np.empty((nx, cfg.num_points))
    for c in range(nx):
        d = 1.0 + (rd - 1.0) * s[c]
        u = u1 + (u1 * ru - u1) * s[c]
        t = 1.0 + (rt - 1.0) * s[c]
        f[c] = maxwellian(cfg, d, u, 0.0, 0.0, t)
    return cfg, f


def mixture_instance(rng, edge, samples_per_node=50, dt=0.05, kn=1.0, seed=0, precision=capi.DVM_PREC_FP64,
                     vpad=4.0):
    """One space-homogeneous instance of the batched workload: a grid of
    edge^3 nodes and a sum of 1-3 Maxwellians with T ~ U[0.5, 3],
    u ~ U[-1.5, 1.5]^3 and positive weights (total density 1). The grid spans
    +-(1.5 + vpad sqrt(T_max)) so every component's tails are resolved; hard
    spheres, explicit sample count samples_per_node * edge^3."""
    k = int(rng.integers(1, 4))
    temps = rng.uniform(0.5, 3.0, size=k)
    vels = rng.uniform(-1.5, 1.5, size=(k, 3))
    w = rng.uniform(0.2, 1.0, size=k)
    w = w / w.sum()
    vmax = 1.5 + vpad * math.sqrt(float(temps.max()))
    ax = uniform_axis(edge, -vmax, vmax)
    cfg = SolverConfig(dim=capi.DVM_DIM_0D, nx=1, ny=1, vx=ax, vy=ax.copy(), vz=ax.copy(), dt=dt, kn=kn,
                       omega=0.5, alpha=1.0, sample_mode=capi.DVM_SAMPLES_EXPLICIT,
                       samples_per_cell=int(samples_per_node * edge ** 3), seed=seed, precision=precision,
                       enable_convection=0, name=f"mixture {edge}^3 x{k}")
    f = np.zeros(cfg.num_points)
    for q in range(k):
        f += maxwellian(cfg, w[q], *vels[q], temps[q])
    return cfg, f.reshape(1, -1)


def batched_homogeneous(count=4096, seed=11, edges=tuple(range(15, 42, 2)), samples_per_node=50, dt=0.05,
                        precision=capi.DVM_PREC_FP64):
    """The batched workload: `count` space-homogeneous instances, edge drawn
    uniformly from `edges` (odd 15..41: mean nodes 27 412), each relaxing a
    mixture of Maxwellians with 50 samples per node per step. Instance i has
    seed seed * 1_000_003 + i."""
    rng = np.random.default_rng(seed)
    es = rng.choice(np.asarray(edges), size=count)
    out = []
    for i, e in enumerate(es):
        out.append(mixture_instance(rng, int(e), samples_per_node=samples_per_node, dt=dt,
                                    seed=seed * 1_000_003 + i, precision=precision))
    return out
