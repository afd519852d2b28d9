"""The HIP library loads and exports every entry point include/dvm_hip.h
declares (no GPU needed: nothing is called)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "dvm_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dvm_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_boundary():
    syms = declared_symbols()
    for s in ("dvm_ctx_create", "dvm_ctx_destroy", "dvm_upload_f", "dvm_download_f", "dvm_moments",
              "dvm_collide", "dvm_convect", "dvm_step", "dvm_remap", "dvm_collide_trace", "dvm_last_error"):
        assert s in syms


def test_library_exports_every_declared_symbol(pkg):
    path = pkg.capi.LIB_PATH
    if not os.path.exists(path):
        pytest.fail(f"{path} is not built (run __graft_entry__.build())")
    lib = C.CDLL(path)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_create_without_device_reports_error(pkg):
    """On a machine without a GPU, context creation fails with a message rather
    than falling back to the CPU."""
    import importlib
    torch_free = importlib.import_module("2d_dvm_microflows_amd.dvm")
    lib = torch_free.load_library()
    n = C.c_int(0)
    hip = C.CDLL("libamdhip64.so")
    if hip.hipGetDeviceCount(C.byref(n)) == 0 and n.value > 0:
        pytest.skip("a GPU is visible")
    cfg, _ = pkg.homogeneous_relaxation(n_v=5, samples=10)
    with pytest.raises(torch_free.DvmError):
        pkg.Solver(cfg)
