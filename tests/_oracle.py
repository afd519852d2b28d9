"""Loads the CPU oracle (oracle/liboracle.so) for the tests. The oracle is the
checker only; product code never imports this module."""
import ctypes as C
import importlib
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
GOLDEN = os.path.join(ROOT, "tests", "golden")

_lib = None


def load_oracle():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "liboracle.so"], check=True,
                       capture_output=True)
    capi = importlib.import_module("2d_dvm_microflows_amd.capi")
    lib = capi.declare(C.CDLL(ORACLE_SO), "orc")
    lib.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    lib.orc_philox4x32_10.restype = None
    lib.orc_referenced_search.argtypes = [C.POINTER(C.c_double), C.c_int32, C.POINTER(C.c_int32),
                                          C.POINTER(C.c_int32), C.c_double]
    lib.orc_referenced_search.restype = C.c_int32
    lib.orc_build_buckets.argtypes = [C.POINTER(C.c_double), C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    lib.orc_build_buckets.restype = None
    lib.orc_uniform_axis.argtypes = [C.c_int32, C.c_double, C.c_double, C.POINTER(C.c_double)]
    lib.orc_uniform_axis.restype = None
    lib.orc_collide_cells.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]
    lib.orc_collide_cells.restype = C.c_int
    pi = C.POINTER(C.c_int32)
    lib.orc_cell_cdf.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double), pi, pi, pi,
                                 pi, pi]
    lib.orc_cell_cdf.restype = C.c_int
    _lib = lib
    return lib


def oracle_solver(cfg):
    dvm = importlib.import_module("2d_dvm_microflows_amd.dvm")
    return dvm.Solver(cfg, lib=load_oracle(), prefix="orc")


def golden(name):
    import numpy as np
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def golden_cases(name):
    g = golden(name)
    cases = {}
    for k, v in g.items():
        case, field = k.split("__", 1)
        cases.setdefault(case, {})[field] = v
    return cases
