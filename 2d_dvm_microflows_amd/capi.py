"""ctypes view of the C-ABI declared in include/dvm_hip.h.

The struct layout and constants here mirror the header one for one; both the
HIP library (libdvmhip.so) and, in tests, the CPU oracle take a pointer to
this struct.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libdvmhip.so")

DVM_OK = 0
DVM_COLL_NONE, DVM_COLL_FMC, DVM_COLL_VR = 0, 2, 3
DVM_BC_WALL, DVM_BC_INOUT, DVM_BC_ZEROGRAD, DVM_BC_COUETTE = 1, 2, 3, 4
DVM_CS_CLARKE, DVM_CS_BIRD, DVM_CS_VKI = 1, 2, 3
DVM_DIM_0D, DVM_DIM_1D, DVM_DIM_2D = 0, 1, 2
DVM_PREC_FP64, DVM_PREC_FP32, DVM_PREC_FP32_REMAP32 = 0, 1, 2
DVM_SAMPLES_REFERENCE, DVM_SAMPLES_EXPLICIT = 0, 1
DVM_CONV_UPWIND1, DVM_CONV_MINMOD2 = 1, 2
DVM_NPROP = 8
DVM_NEXT = 10  # qx qy qz txx tyy tzz txy txz tyz entropy
P_DENS, P_U, P_V, P_W, P_TRTEMP, P_TEMP, P_PRESSURE, P_AVGSPEED = range(8)
W_DENS, W_U, W_V, W_W, W_TEMP = range(5)
TRACE_INTS, TRACE_DOUBLES = 32, 16


class DvmConfig(C.Structure):
    _fields_ = [
        ("dim", C.c_int32),
        ("nx", C.c_int32), ("ny", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double),
        ("nvx", C.c_int32), ("nvy", C.c_int32), ("nvz", C.c_int32),
        ("vx", C.POINTER(C.c_double)), ("vy", C.POINTER(C.c_double)), ("vz", C.POINTER(C.c_double)),
        ("dt", C.c_double), ("kn", C.c_double),
        ("mass", C.c_double), ("diam", C.c_double), ("omega", C.c_double), ("alpha", C.c_double),
        ("viscosity", C.c_double), ("t_viscosity", C.c_double),
        ("csection_coeff", C.c_int32),
        ("coll_solver", C.c_int32),
        ("cdf_tolerance", C.c_double),
        ("num_repl", C.c_int32),
        ("coln_rms", C.c_double),
        ("sample_mode", C.c_int32),
        ("samples_per_cell", C.c_int64),
        ("seed", C.c_uint64),
        ("precision", C.c_int32),
        ("enable_convection", C.c_int32),
        ("bc_left", C.c_int32), ("bc_right", C.c_int32),
        ("acc_coeff", C.c_double),
        ("wall", (C.c_double * 5) * 4),
        ("conv_order", C.c_int32),
    ]


def declare(lib, prefix):
    """Attach argtypes/restype for the entry points named `prefix`_* (dvm or orc)."""
    P, I32, I64, D = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    pd, pi32, pi64 = C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_int64)
    sig = {
        "ctx_create": (C.c_int, [C.POINTER(DvmConfig), C.POINTER(C.c_void_p)]),
        "ctx_destroy": (None, [P]),
        "last_error": (C.c_char_p, [P]),
        "upload_f": (C.c_int, [P, I32, pd]),
        "download_f": (C.c_int, [P, I32, pd]),
        "moments": (C.c_int, [P, pd]),
        "moments_ext": (C.c_int, [P, pd]),
        "collide": (C.c_int, [P, I32, pi64]),
        "convect": (C.c_int, [P, I32, I32]),
        "step": (C.c_int, [P, I32]),
        "remap": (C.c_int, [P, I64, pd, D, pi32, pd]),
        "collide_trace": (C.c_int, [P, I32, I32, I64, pi32, pd, pi64]),
        "l1": (C.c_int, [P, pd]),
        "set_uniform_stream": (C.c_int, [P, pd, I64]),
        "uniform_stream_position": (C.c_int, [P, pi64]),
        "set_wall_heat_flux_freq": (C.c_int, [P, I32]),
        "wall_heat_flux": (C.c_int, [P, pd, pi32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, f"{prefix}_{name}")
        fn.restype = res
        fn.argtypes = args
    return lib


def declare_batch(lib):
    """Attach argtypes/restype for the dvm_batch_* entry points."""
    P, I32, I64 = C.c_void_p, C.c_int32, C.c_int64
    pd, pi32, pi64 = C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_int64)
    pcfg = C.POINTER(DvmConfig)
    sig = {
        "create": (C.c_int, [pcfg, I32, I64, I32, C.POINTER(C.c_void_p)]),
        "destroy": (None, [P]),
        "last_error": (C.c_char_p, [P]),
        "footprint": (C.c_int, [pcfg, I32, pi64, pi64]),
        "admit": (C.c_int, [P, pcfg, pd, pi32]),
        "retire": (C.c_int, [P, I32]),
        "step": (C.c_int, [P, I32]),
        "synchronize": (C.c_int, [P]),
        "download_f": (C.c_int, [P, I32, pd]),
        "moments": (C.c_int, [P, I32, pd]),
        "l1": (C.c_int, [P, pd]),
        "slot_info": (C.c_int, [P, I32, pi32, pi64, pi32, pi32]),
        "capacity": (C.c_int, [P, pi64, pi64, pi32, pi32]),
        "set_tiles": (C.c_int, [P, I32, I32, I32]),
        "get_tiles": (C.c_int, [P, pi32]),
        "tune": (C.c_int, [P, I32, pi32, pd]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, f"dvm_batch_{name}")
        fn.restype = res
        fn.argtypes = args
    return lib
