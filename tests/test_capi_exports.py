"""The HIP library loads and exports every entry point include/dvm_hip.h
declares (no GPU needed: nothing is called)."""
import ctypes as C
import os
import re

import numpy as np

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


def test_sample_count_bound_is_checked_before_any_device_call(pkg):
    """Per-cell sample counts are bounded by 2^30 so that sample indices and
    per-cell totals stay within 32-bit integers on the device; a larger
    explicit count is refused by the argument checks, which run before any
    HIP call."""
    import importlib
    dvm = importlib.import_module("2d_dvm_microflows_amd.dvm")
    cfg, _ = pkg.homogeneous_relaxation(n_v=5, samples=(1 << 30) + 1)
    with pytest.raises(dvm.DvmError) as e:
        pkg.Solver(cfg)
    assert "2^30" in str(e.value)
    b = pkg.Batch(cfg.with_(samples_per_cell=10), max_slots=1, node_capacity=125, cell_capacity=1) \
        if _gpu_visible() else None
    if b is not None:
        with pytest.raises(dvm.DvmError):
            b.admit(cfg, np.ones((1, 125)))
        b.close()


def _gpu_visible():
    n = C.c_int(0)
    try:
        hip = C.CDLL("libamdhip64.so")
    except OSError:
        return False
    return hip.hipGetDeviceCount(C.byref(n)) == 0 and n.value > 0
