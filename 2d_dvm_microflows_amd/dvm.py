"""Host-side mirror of the reference's time-step interface over the C-ABI.

The reference drives its hot path through Fortran module procedures
(software/trunk/src):

  compute_collision_integral   collisions/collision_integral.f90:163
  compute_convection           convection/convection.f90:69
  calculate_physical_properties dist_funcs/physical_props.f90:285
  solve (the time loop)        solver/dvm_solver.f90:231-369

`Solver` keeps those names and their order of calls, and forwards each to the
HIP library through include/dvm_hip.h. Device memory (the resident
distribution, CDFs and moments) belongs to the library context; numpy arrays
cross the boundary only at upload, download and moment read-back.
"""
import ctypes as C
import os

import numpy as np

from . import capi
from .config import SolverConfig

_LIB = None


def load_library(path=None):
    """Load libdvmhip.so; raise if it is missing (there is no CPU fallback)."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or capi.LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"HIP library {p} is not built; run __graft_entry__.build() (hipcc --offload-arch=gfx950)")
    lib = capi.declare_batch(capi.declare(C.CDLL(p), "dvm"))
    lib.dvm_timing_reset.restype = C.c_int
    lib.dvm_timing_reset.argtypes = [C.c_void_p]
    lib.dvm_timing_read.restype = C.c_int
    lib.dvm_timing_read.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
    if path is None:
        _LIB = lib
    return lib


class DvmError(RuntimeError):
    """A non-zero status of the C-ABI, with the library's message."""

    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Solver:
    """One simulation context (one HIP stream) behind the reference's API.

    `lib`/`prefix` select the implementation of the C-ABI; the product uses the
    HIP library (`dvm_*`). Tests pass the CPU oracle (`orc_*`) through the same
    wrapper so that both run the same calls on the same inputs.
    """

    def __init__(self, cfg: SolverConfig, lib=None, prefix="dvm"):
        self.cfg = cfg
        self.lib = lib if lib is not None else load_library()
        self.prefix = prefix
        self._ccfg, self._keep = cfg.to_c()
        h = C.c_void_p()
        rc = self._fn("ctx_create")(C.byref(self._ccfg), C.byref(h))
        if rc != 0:
            raise DvmError(rc, self._fn("last_error")(None).decode())
        self.h = h
        self.ncell = cfg.nx * cfg.ny
        self.N = cfg.nvx * cfg.nvy * cfg.nvz
        self.ntime = 0

    def _fn(self, name):
        return getattr(self.lib, f"{self.prefix}_{name}")

    def _check(self, rc):
        if rc != 0:
            raise DvmError(rc, self._fn("last_error")(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self._fn("ctx_destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # --- distribution transfer ---------------------------------------------
    def upload_f(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(self.ncell, self.N)
        self._check(self._fn("upload_f")(self.h, 0, _dptr(f)))

    def download_f(self):
        out = np.empty((self.ncell, self.N), dtype=np.float64)
        self._check(self._fn("download_f")(self.h, 0, _dptr(out)))
        return out

    # --- the reference's operators -----------------------------------------
    def calculate_physical_properties(self):
        """Moments of every cell, [ncell, DVM_NPROP] (physical_props.f90:285)."""
        out = np.empty((self.ncell, capi.DVM_NPROP), dtype=np.float64)
        self._check(self._fn("moments")(self.h, _dptr(out)))
        return out

    def output_moments(self):
        """Heat flux, viscous stress and entropy of every cell, [ncell, DVM_NEXT]
        (compute_heat_flux / compute_viscous_stress / compute_entropy,
        physical_props.f90:1871-2065)."""
        out = np.empty((self.ncell, capi.DVM_NEXT), dtype=np.float64)
        self._check(self._fn("moments_ext")(self.h, _dptr(out)))
        return out

    def set_wall_heat_flux_freq(self, every):
        """Record the wall heat fluxes during the convection of every step whose
        ntime is a multiple of `every` (set_wall_heat_flux_dump_freq,
        boundary_conditions.f90:108-116; 0 stops)."""
        self._check(self._fn("set_wall_heat_flux_freq")(self.h, int(every)))

    def wall_heat_flux(self):
        """The last recorded wall heat fluxes, [4, max(nx, ny)] (left, right,
        bottom, top; index along the wall), and their step (-1: none);
        boundary_conditions.f90:872-923 and analogues."""
        out = np.zeros((4, max(self.cfg.nx, self.cfg.ny)), dtype=np.float64)
        nt = C.c_int32(-1)
        self._check(self._fn("wall_heat_flux")(self.h, _dptr(out), C.byref(nt)))
        return out, nt.value

    def compute_collision_integral(self, ntime):
        """Collision phase of every cell (collision_integral.f90:163); returns
        the samples drawn per cell."""
        out = np.zeros(self.ncell, dtype=np.int64)
        self._check(self._fn("collide")(self.h, int(ntime), out.ctypes.data_as(C.POINTER(C.c_int64))))
        return out

    def compute_convection(self, ntime):
        """Boundary conditions and upwind sweep (convection.f90:69)."""
        self._check(self._fn("convect")(self.h, 0, int(ntime)))

    def step(self, ntime=None):
        """collide -> convect -> moments, without a host synchronisation."""
        if ntime is None:
            self.ntime += 1
            ntime = self.ntime
        self._check(self._fn("step")(self.h, int(ntime)))

    def initialize(self, f):
        """Upload f and compute its moments (solve() does this before the loop,
        solver/dvm_solver.f90:275-278)."""
        self.upload_f(f)
        return self.calculate_physical_properties()

    def solve(self, num_steps, props_every=0):
        """The time loop of solver/dvm_solver.f90:299-357; returns the moment
        history when props_every > 0."""
        hist = []
        for _ in range(num_steps):
            self.step()
            if props_every and self.ntime % props_every == 0:
                hist.append(self.calculate_physical_properties())
        return hist

    # --- parity hooks ------------------------------------------------------
    def remap(self, uvw, f_off=1.0):
        uvw = np.ascontiguousarray(uvw, dtype=np.float64).reshape(-1, 3)
        n = uvw.shape[0]
        st = np.zeros((n, 7), dtype=np.int32)
        m = np.zeros((n, 7), dtype=np.float64)
        self._check(self._fn("remap")(self.h, n, _dptr(uvw), float(f_off),
                                      st.ctypes.data_as(C.POINTER(C.c_int32)), _dptr(m)))
        return st, m

    def collide_trace(self, ntime, cell, max_records):
        irec = np.zeros((max_records, capi.TRACE_INTS), dtype=np.int32)
        drec = np.zeros((max_records, capi.TRACE_DOUBLES), dtype=np.float64)
        n = C.c_int64(0)
        self._check(self._fn("collide_trace")(self.h, int(ntime), int(cell), int(max_records),
                                              irec.ctypes.data_as(C.POINTER(C.c_int32)), _dptr(drec),
                                              C.byref(n)))
        k = min(n.value, max_records)
        return irec[:k], drec[:k], n.value

    def l1_to_equilibrium(self):
        """Per-cell sum |f - M| / n against the discrete Maxwellian at the cell's
        own moments (dvm_l1)."""
        out = np.zeros(self.ncell, dtype=np.float64)
        self._check(self._fn("l1")(self.h, _dptr(out)))
        return out

    def set_uniform_stream(self, u):
        """Parity mode: later collisions read their uniforms from the list u
        (cells in order, samples in order, 2 + 2 NRepl values per sample: the
        two partner uniforms, then phi and theta per replenish location)
        instead of the counter-based generator; an empty list returns to it.
        The library keeps its own copy of the list."""
        u = np.ascontiguousarray(u, dtype=np.float64).ravel()
        self._check(self._fn("set_uniform_stream")(self.h, _dptr(u) if u.size else None, int(u.size)))

    def uniform_stream_position(self):
        """Values of the uniform list read so far."""
        n = C.c_int64(0)
        self._check(self._fn("uniform_stream_position")(self.h, C.byref(n)))
        return n.value

    PHASES = ("cdf", "prep", "collide", "conserve", "walls", "sweep", "moments")

    def timing_reset(self):
        """Wait for queued work and zero the per-phase device-time totals."""
        self._check(self.lib.dvm_timing_reset(self.h))

    def timing_read(self):
        """Per-phase device ms since the last reset (HIP events on the
        context's stream) and the number of calls they cover."""
        ms = (C.c_double * len(self.PHASES))()
        n = C.c_int64(0)
        self._check(self.lib.dvm_timing_read(self.h, ms, C.byref(n)))
        return dict(zip(self.PHASES, list(ms))), n.value

    def synchronize(self):
        """Wait for the context's queued work (drains the timing ring)."""
        self.timing_read()
