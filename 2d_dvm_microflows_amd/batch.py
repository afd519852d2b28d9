"""Batched solves over the dvm_batch_* entry points of include/dvm_hip.h.

A `Batch` keeps many independent space-homogeneous, 1-D and 2-D solves resident on
the device and advances all of them with one replayed HIP graph per step.
Every instance keeps the reference's per-solve semantics (the loop body of
solve(), software/trunk/src/solver/dvm_solver.f90:299-357): its trajectory is
the one `Solver` produces for the same config, whatever else is resident.
"""
import ctypes as C

import numpy as np

from . import capi
from .config import SolverConfig
from .dvm import DvmError, load_library

DVM_E_NOMEM = -3  # include/dvm_hip.h


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Batch:
    """Resident pool of instances sharing one species / collision model.

    species      SolverConfig whose species and collision fields every
                 admitted instance must share (precision, sample mode, ...)
    max_slots    instances resident at once
    node_capacity, cell_capacity
                 pool sizes: sum of nx*ny * nvx*nvy*nvz and of nx*ny over residents
    """

    def __init__(self, species: SolverConfig, max_slots, node_capacity, cell_capacity, lib=None):
        self.lib = lib if lib is not None else load_library()
        self.species = species
        self._ccfg, self._keep = species.to_c()
        h = C.c_void_p()
        rc = self.lib.dvm_batch_create(C.byref(self._ccfg), int(max_slots), int(node_capacity),
                                       int(cell_capacity), C.byref(h))
        if rc != 0:
            raise DvmError(rc, self.lib.dvm_batch_last_error(None).decode())
        self.h = h
        self.max_slots = int(max_slots)
        self.shape = {}  # slot -> (ncell, N)

    def _check(self, rc):
        if rc != 0:
            raise DvmError(rc, self.lib.dvm_batch_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.dvm_batch_destroy(self.h)
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

    @staticmethod
    def footprint(cfg: SolverConfig, precision=None, lib=None):
        """(pool nodes, HBM bytes) an instance occupies."""
        lib = lib if lib is not None else load_library()
        c, keep = cfg.to_c()
        n = C.c_int64(0)
        b = C.c_int64(0)
        prec = cfg.precision if precision is None else precision
        rc = lib.dvm_batch_footprint(C.byref(c), int(prec), C.byref(n), C.byref(b))
        if rc != 0:
            raise DvmError(rc, lib.dvm_batch_last_error(None).decode())
        return n.value, b.value

    def admit(self, cfg: SolverConfig, f0):
        """Admit one instance with initial f0 [nx*ny, N]; returns its slot.
        Raises DvmError with code DVM_E_NOMEM (-3) when it does not fit now."""
        c, keep = cfg.to_c()
        ncell, N = cfg.nx * cfg.ny, cfg.num_points
        f0 = np.ascontiguousarray(f0, dtype=np.float64).reshape(ncell, N)
        slot = C.c_int32(-1)
        self._check(self.lib.dvm_batch_admit(self.h, C.byref(c), _dptr(f0), C.byref(slot)))
        self.shape[slot.value] = (ncell, N)
        return slot.value

    def try_admit(self, cfg, f0):
        """admit(), or None when the pools or slots are full."""
        try:
            return self.admit(cfg, f0)
        except DvmError as e:
            if e.code == DVM_E_NOMEM:
                return None
            raise

    def retire(self, slot):
        self._check(self.lib.dvm_batch_retire(self.h, int(slot)))
        self.shape.pop(slot, None)

    def step(self, nsteps=1):
        """Enqueue nsteps steps of every resident instance (no host wait)."""
        self._check(self.lib.dvm_batch_step(self.h, int(nsteps)))

    def synchronize(self):
        self._check(self.lib.dvm_batch_synchronize(self.h))

    def download_f(self, slot):
        ncell, N = self.shape[slot]
        out = np.empty((ncell, N), dtype=np.float64)
        self._check(self.lib.dvm_batch_download_f(self.h, int(slot), _dptr(out)))
        return out

    def moments(self, slot):
        ncell, _ = self.shape[slot]
        out = np.empty((ncell, capi.DVM_NPROP), dtype=np.float64)
        self._check(self.lib.dvm_batch_moments(self.h, int(slot), _dptr(out)))
        return out

    def l1(self):
        """[max_slots] distance from equilibrium of each resident space-homogeneous
        instance (-1 elsewhere)."""
        out = np.empty(self.max_slots, dtype=np.float64)
        self._check(self.lib.dvm_batch_l1(self.h, _dptr(out)))
        return out

    def slot_info(self, slot):
        steps, ncell, lds = C.c_int32(0), C.c_int32(0), C.c_int32(0)
        nodes = C.c_int64(0)
        self._check(self.lib.dvm_batch_slot_info(self.h, int(slot), C.byref(steps), C.byref(nodes),
                                                 C.byref(ncell), C.byref(lds)))
        return dict(steps=steps.value, nodes=nodes.value, ncell=ncell.value, in_lds=bool(lds.value))

    def capacity(self):
        fn, lf = C.c_int64(0), C.c_int64(0)
        fc, fs = C.c_int32(0), C.c_int32(0)
        self._check(self.lib.dvm_batch_capacity(self.h, C.byref(fn), C.byref(lf), C.byref(fc), C.byref(fs)))
        return dict(free_nodes=fn.value, largest_free=lf.value, free_cells=fc.value, free_slots=fs.value)

    def set_tiles(self, lanes, pick_k, item_samples):
        self._check(self.lib.dvm_batch_set_tiles(self.h, int(lanes), int(pick_k), int(item_samples)))

    def tiles(self):
        t = (C.c_int32 * 4)()
        self._check(self.lib.dvm_batch_get_tiles(self.h, t))
        return dict(lanes=t[0], pick_k=t[1], item_samples=t[2], workgroups=t[3])

    def tune(self, steps=1):
        """Autotune the collision tiles on the resident set; the state is left
        as it was. Returns (tiles, ms per step of the best)."""
        t = (C.c_int32 * 3)()
        ms = C.c_double(0.0)
        self._check(self.lib.dvm_batch_tune(self.h, int(steps), t, C.byref(ms)))
        return dict(lanes=t[0], pick_k=t[1], item_samples=t[2]), ms.value
