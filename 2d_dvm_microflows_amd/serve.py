"""Continuous-batching request loop over a resident `Batch`.

Requests (independent space-homogeneous relaxations and 1-D shocks, each with
its own velocity grid and cell count) arrive over time. The loop admits them
into the batch's device pools while their footprint fits (first come first
served, later requests backfill around one that does not fit yet), advances
every resident solve together `check_every` steps at a time (one replayed HIP
graph per step), checks each solve's stop rule, retires finished solves and
records each request's latency (completion - arrival).

Stop rules (a convergence test of this library; the reference runs a fixed
number of steps, solver/dvm_solver.f90:299):
  homogeneous  L1 = sum |f - M| / n <= tol (M the discrete Maxwellian at the
               solve's own moments), or an L1 plateau: no 5 % improvement on
               the best L1 for `plateau_steps` steps (the discrete collision
               operator's fixed point sits slightly off the discrete
               Maxwellian on coarse grids), or max_steps
  shock        steady state of a stationary shock, tested on windows of
               `steady_window` checks: the window-mean density profile
               differs from the previous window's by at most tol of its peak
               (the Monte Carlo floor of 100-step means is 2-7e-4 on the
               served shocks at 20 000 samples per cell), and the window-mean
               mass flux n u of every interior cell lies within `flux_tol`
               (relative) of the inflow flux (a stationary shock carries one
               flux through every cell), in two consecutive windows; or
               max_steps
An instance's trajectory depends only on its own config (seeded counter-based
sampling), and its checks fall every check_every steps from its own
admission, so a request ends in the same state and step whether it is served
alone or in a mix.
"""
import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import capi
from .batch import Batch
from .config import SolverConfig, maxwellian, mixture_instance, rankine_hugoniot, sound_speed, uniform_axis


@dataclass
class Request:
    rid: int
    arrival: float               # seconds after the loop starts
    cfg: SolverConfig
    f0: np.ndarray
    kind: str = "homogeneous"    # or "shock"
    tol: float = 1e-4            # L1 (homogeneous) or window density change (shock) tolerance
    max_steps: int = 400


@dataclass
class Outcome:
    rid: int
    kind: str
    arrival: float
    admitted: float = math.nan
    done: float = math.nan
    steps: int = 0
    reason: str = ""             # tol | plateau | steady | max_steps | too_large
    metric: float = math.nan     # final L1 or steady-state change
    nodes: int = 0
    f: np.ndarray = None         # final distribution when kept
    props: np.ndarray = None     # final moments [nx, DVM_NPROP] when kept

    @property
    def latency(self):
        return self.done - self.arrival


@dataclass
class ServeStats:
    outcomes: list
    wall_s: float
    steps_launched: int
    max_resident: int
    samples: float               # collision samples drawn over the run
    node_updates: float          # cell-node updates over the run
    latency_p50: float = math.nan
    latency_p99: float = math.nan
    extra: dict = field(default_factory=dict)


class Server:
    """species       SolverConfig carrying the shared species / precision
    node_capacity   pool nodes (the HBM budget: dvm_batch_footprint gives an
                    instance's bytes)
    batch           an existing Batch (or an object with its admit / step /
                    l1 / moments / retire interface) instead of a new one
    steady_window   checks per window of the shock steady-state rule
    flux_tol        its mass-flux uniformity tolerance
    """

    def __init__(self, species, node_capacity, cell_capacity, max_slots, check_every=10, plateau_steps=50,
                 keep_f=False, keep_moments=False, clock=time.perf_counter, sleep=time.sleep, batch=None,
                 steady_window=10, flux_tol=0.02):
        self.batch = batch if batch is not None else Batch(species, max_slots=max_slots,
                                                           node_capacity=node_capacity,
                                                           cell_capacity=cell_capacity)
        self.node_capacity = int(node_capacity)
        self.cell_capacity = int(cell_capacity)
        self.check_every = int(check_every)
        self.plateau_steps = int(plateau_steps)
        self.steady_window = max(1, int(steady_window))
        self.flux_tol = float(flux_tol)
        self.keep_f = keep_f
        self.keep_moments = keep_moments
        self.clock = clock
        self.sleep = sleep

    def close(self):
        self.batch.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _shock_check(self, st, props, req):
        """One check of the shock steady-state rule (see the module docstring);
        returns (reason, metric), metric the window's largest density change
        relative to the peak density."""
        dens = props[:, capi.P_DENS]
        st["win"].append(dens.copy())
        nu = dens * props[:, capi.P_U]
        st["flux"] = nu if st["flux"] is None else st["flux"] + nu
        if len(st["win"]) < self.steady_window:
            return "", math.nan
        mean = np.mean(np.array(st["win"]), axis=0)
        flux = st["flux"] / len(st["win"])
        st["win"], st["flux"] = [], None
        c = req.cfg
        inflow = c.wall[0][capi.W_DENS] * c.wall[0][capi.W_U]
        inner = flux[2:-2] if len(flux) > 4 else flux
        flux_dev = float(np.max(np.abs(inner - inflow)) / abs(inflow)) if inflow != 0 else math.inf
        prev, st["prev"] = st["prev"], mean
        if prev is None:
            return "", math.nan
        metric = float(np.max(np.abs(mean - prev)) / max(float(np.max(np.abs(mean))), 1e-300))
        st["ok"] = st["ok"] + 1 if (metric <= req.tol and flux_dev <= self.flux_tol) else 0
        return ("steady" if st["ok"] >= 2 else ""), metric

    def _fits_at_all(self, req):
        nc = req.cfg.nx * req.cfg.ny
        return nc * req.cfg.num_points <= self.node_capacity and nc <= self.cell_capacity

    def run(self, requests):
        reqs = sorted(requests, key=lambda r: r.arrival)
        out = {r.rid: Outcome(r.rid, r.kind, r.arrival, nodes=r.cfg.nx * r.cfg.num_points) for r in reqs}
        waiting = deque()
        resident = {}                # slot -> dict(req, steps, best, since, prev)
        b = self.batch
        i, launched, max_res = 0, 0, 0
        samples = updates = 0.0
        t0 = self.clock()
        while i < len(reqs) or waiting or resident:
            now = self.clock() - t0
            while i < len(reqs) and reqs[i].arrival <= now:
                waiting.append(reqs[i])
                i += 1
            for req in list(waiting):
                if not self._fits_at_all(req):
                    waiting.remove(req)
                    o = out[req.rid]
                    o.reason, o.done = "too_large", now
                    continue
                slot = b.try_admit(req.cfg, req.f0)
                if slot is None:
                    continue
                waiting.remove(req)
                out[req.rid].admitted = self.clock() - t0
                resident[slot] = dict(req=req, steps=0, best=math.inf, since=0, prev=None, win=[], flux=None,
                                      ok=0)
            max_res = max(max_res, len(resident))
            if not resident:
                if i < len(reqs):
                    self.sleep(max(0.0, reqs[i].arrival - (self.clock() - t0)))
                    continue
                for req in waiting:       # cannot be admitted even into empty pools
                    out[req.rid].reason, out[req.rid].done = "too_large", self.clock() - t0
                break
            n = self.check_every
            b.step(n)
            launched += n
            for st in resident.values():
                c = st["req"].cfg
                samples += n * c.samples_per_cell * c.nx
                updates += n * c.num_points * c.nx
            l1 = b.l1()                   # waits for the steps
            now = self.clock() - t0
            finished = []
            for slot, st in resident.items():
                req = st["req"]
                st["steps"] += n
                reason, metric = "", math.nan
                if req.kind == "homogeneous":
                    metric = float(l1[slot])
                    if metric < 0.95 * st["best"]:
                        st["best"], st["since"] = metric, 0
                    else:
                        st["since"] += n
                    if metric <= req.tol:
                        reason = "tol"
                    elif st["since"] >= self.plateau_steps:
                        reason = "plateau"
                else:
                    reason, metric = self._shock_check(st, b.moments(slot), req)
                if not reason and st["steps"] >= req.max_steps:
                    reason = "max_steps"
                if reason:
                    o = out[req.rid]
                    o.done, o.steps, o.reason, o.metric = now, st["steps"], reason, metric
                    if self.keep_f:
                        o.f = b.download_f(slot)
                    if self.keep_moments:
                        o.props = b.moments(slot)
                    finished.append(slot)
            for slot in finished:
                b.retire(slot)
                del resident[slot]
        wall = self.clock() - t0
        outs = [out[r.rid] for r in reqs]
        lat = np.array([o.latency for o in outs if o.reason != "too_large"])
        stats = ServeStats(outs, wall, launched, max_res, samples, updates)
        if lat.size:
            stats.latency_p50 = float(np.percentile(lat, 50))
            stats.latency_p99 = float(np.percentile(lat, 99))
        return stats


# ------------------------------------------------------------------- workload
def shock_instance(rng, nx, edge, mach, samples_per_cell, seed, precision, dx=0.1, cfl=0.5, kn=1.0,
                   gamma=5.0 / 3.0):
    """A 1-D normal shock request: Mach `mach` upstream (T = 1), the
    Rankine-Hugoniot state downstream, a tanh profile between them, in/out
    flow at both ends; hard spheres so that it shares the homogeneous
    requests' species (monatomic: gamma = 5/3)."""
    u1 = mach * sound_speed(gamma, 1.0, 1.0)
    rd, ru, rt, _ = rankine_hugoniot(mach, gamma)
    vmax = max(6.0, u1 + 4.0 * math.sqrt(max(1.0, rt)))
    ax = uniform_axis(edge, -vmax, vmax)
    wall = np.array([[1.0, u1, 0.0, 0.0, 1.0], [rd, u1 * ru, 0.0, 0.0, rt],
                     [1.0, 0.0, 0.0, 0.0, 1.0], [1.0, 0.0, 0.0, 0.0, 1.0]])
    cfg = SolverConfig(dim=capi.DVM_DIM_1D, nx=nx, ny=1, dx=dx, vx=ax, vy=ax.copy(), vz=ax.copy(),
                       dt=cfl * dx / vmax, kn=kn, omega=0.5, alpha=1.0, sample_mode=capi.DVM_SAMPLES_EXPLICIT,
                       samples_per_cell=int(samples_per_cell), seed=seed, precision=precision, enable_convection=1,
                       bc_left=capi.DVM_BC_INOUT, bc_right=capi.DVM_BC_INOUT, wall=wall,
                       name=f"shock M{mach:.2f} {nx} cells {edge}^3")
    xc = (np.arange(nx) + 0.5) - nx / 2.0
    s = 0.5 * (1.0 + np.tanh(xc / 5.0))
    f = np.empty((nx, cfg.num_points))
    for c in range(nx):
        f[c] = maxwellian(cfg, 1.0 + (rd - 1.0) * s[c], u1 + (u1 * ru - u1) * s[c], 0.0, 0.0,
                          1.0 + (rt - 1.0) * s[c])
    return cfg, f


def synthetic_requests(count=120, rate=20.0, seed=5, homogeneous_share=0.7, precision=capi.DVM_PREC_FP32,
                       edges=tuple(range(15, 42, 2)), samples_per_node=50, shock_cells=(50, 400),
                       shock_edges=tuple(range(21, 32, 2)), shock_samples=20000, l1_tol=1e-4, steady_tol=1e-3,
                       max_steps=300, shock_max_steps=6000):
    """The serving workload: Poisson arrivals at `rate` per second; 70 %
    space-homogeneous mixtures (edge 15..41 odd, 50 samples per node), 30 %
    1-D shocks (Mach U[1.2, 3], 50-400 cells, edge 21..31 odd) that run to
    steady state (100-step mean density profile steady to steady_tol and mass
    flux uniform; at most shock_max_steps)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for rid in range(count):
        t += rng.exponential(1.0 / rate)
        if rng.random() < homogeneous_share:
            cfg, f = mixture_instance(rng, int(rng.choice(edges)), samples_per_node=samples_per_node,
                                      seed=seed * 7919 + rid, precision=precision)
            out.append(Request(rid, t, cfg, f, "homogeneous", l1_tol, max_steps))
        else:
            nx = int(rng.integers(shock_cells[0], shock_cells[1] + 1))
            cfg, f = shock_instance(rng, nx, int(rng.choice(shock_edges)), float(rng.uniform(1.2, 3.0)),
                                    shock_samples, seed * 7919 + rid, precision)
            out.append(Request(rid, t, cfg, f, "shock", steady_tol, shock_max_steps))
    return out


def moment_errors(p, ref):
    """Max over cells of the relative error of density, of velocity (in units
    of the thermal speed sqrt(T)) and of temperature, p against ref."""
    th = np.sqrt(np.abs(ref[:, capi.P_TRTEMP]))
    return {"density": float(np.max(np.abs(p[:, 0] - ref[:, 0]) / np.abs(ref[:, 0]))),
            "velocity": float(np.max(np.max(np.abs(p[:, 1:4] - ref[:, 1:4]), axis=1) / th)),
            "temperature": float(np.max(np.abs(p[:, 4] - ref[:, 4]) / np.abs(ref[:, 4])))}


def replay_fp64(requests, outcomes, node_capacity, cell_capacity, max_slots, check_every=10, seed_offset=0):
    """Runs every served request again with plain FP64 collision arithmetic for
    exactly the steps it was served (one batch, admitted in arrival order as
    pool space allows) and returns {rid: errors}: moment_errors of the served
    moments against the FP64 ones and, when the served f was kept, "f_l1" =
    sum |f - f64| / sum |f64|. seed_offset != 0 replays with another seed: the
    spread between two FP64 realisations is the Monte Carlo noise that the
    FP32 error is to be read against."""
    done = {o.rid: o for o in outcomes if o.props is not None and o.steps > 0}
    todo = deque(r for r in sorted(requests, key=lambda r: r.arrival) if r.rid in done)
    if not todo:
        return {}
    species = todo[0].cfg.with_(precision=capi.DVM_PREC_FP64)
    errs = {}
    with Batch(species, max_slots=max_slots, node_capacity=node_capacity, cell_capacity=cell_capacity) as b:
        resident = {}
        while todo or resident:
            while todo:
                r = todo[0]
                cfg = r.cfg.with_(precision=capi.DVM_PREC_FP64, seed=r.cfg.seed + seed_offset)
                slot = b.try_admit(cfg, r.f0)
                if slot is None:
                    break
                todo.popleft()
                resident[slot] = [r, 0]
            if not resident:
                raise RuntimeError("a request does not fit the FP64 replay pools")
            b.step(check_every)
            for slot in list(resident):
                r, n = resident[slot]
                resident[slot][1] = n = n + check_every
                if n >= done[r.rid].steps:
                    o = done[r.rid]
                    e = moment_errors(o.props, b.moments(slot))
                    if o.f is not None:
                        f64 = b.download_f(slot)
                        e["f_l1"] = float(np.abs(o.f - f64).sum() / np.abs(f64).sum())
                    errs[r.rid] = e
                    b.retire(slot)
                    del resident[slot]
    return errs
