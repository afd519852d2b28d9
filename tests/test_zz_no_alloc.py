"""dvm_l1_noalloc (staged in the library's file-scope device scratch instead
of a per-call allocation) against dvm_l1 through the C-ABI: bit-identical
results on two calls in a row, the second on a smaller context, then again on
the first."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _l1_both(s):
    lib = s.lib
    fn = lib.dvm_l1_noalloc
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    ref = s.l1_to_equilibrium()
    out = np.full(s.ncell, np.nan)
    rc = fn(s.h, out.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == 0, lib.dvm_last_error(s.h)
    return ref, out


def test_l1_noalloc_equals_l1(pkg):
    big_cfg, big_f = pkg.lid_driven_cavity(nx=24, ny=20, n_v=9, vmax=4.0, samples=500)
    small_cfg, small_f = pkg.homogeneous_relaxation(n_v=11, samples=20000)
    print("[no-alloc] cavity 24x20 x 9^3, then one homogeneous cell x 11^3, then the cavity again", flush=True)
    with pkg.Solver(big_cfg) as big, pkg.Solver(small_cfg) as small:
        big.initialize(big_f)
        small.initialize(small_f)
        for n in (1, 2):
            big.step(n)
            small.step(n)
        for s in (big, small, big):
            ref, out = _l1_both(s)
            assert out.shape == ref.shape and np.array_equal(out, ref), (s.ncell, np.max(np.abs(out - ref)))
            assert np.all(ref > 0)
