"""xnb_simulation_energies_noalloc (the entry point that allocates nothing) against
xnb_simulation_energies on the same boxes: bitwise equal, on two calls in a row, the second
on a smaller box (fewer reduction partials than the first left in the shared array).  The
output buffer is device memory the test allocates through the HIP runtime the library uses."""
import ctypes as C

import numpy as np
import pytest

from exanbody_amd import PRECISION_FP64, PRECISION_MIXED
from tests.helpers import RCUT, SKIN, fcc_inputs, make_pair

pytestmark = pytest.mark.gpu


def _hip():
    import exanbody_amd
    exanbody_amd.load()  # loads the HIP runtime libxnb_amd.so is linked against
    h = C.CDLL("libamdhip64.so.7")  # that same, already loaded runtime
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return h


def _energies_noalloc(box, hip, dev):
    rc = box.L.xnb_simulation_energies_noalloc(box._h, dev)
    assert rc == 0, box.L.xnb_last_error(box._h)
    box.synchronize()
    out = np.zeros(2, np.float64)
    assert hip.hipMemcpy(out.ctypes.data, dev, 16, 2) == 0  # hipMemcpyDeviceToHost
    return out


def test_noalloc_energies_equal_allocating_entry_point():
    hip = _hip()
    dev = C.c_void_p()
    assert hip.hipMalloc(C.byref(dev), 16) == 0
    try:
        # first call on a box with many partials, second on a smaller one
        for ncell, prec, steps in ((16, PRECISION_FP64, 12), (5, PRECISION_MIXED, 7)):
            r, v, ids, L = fcc_inputs(ncell, temperature=0.9, jitter=0.02, seed=ncell)
            _, b, _ = make_pair(r, v, ids, L, gpu=True)
            b.init(1.0, 1.0, RCUT, SKIN, prec)
            b.run(steps, 0.005, 1.0, 1.0, RCUT, SKIN, precision=prec, rebuild_every=5)
            assert hip.hipMemset(dev, 0xFF, 16) == 0  # stale bytes in the output must be overwritten
            got = _energies_noalloc(b, hip, dev)
            ke, pe = b.energies()
            np.testing.assert_array_equal(got.view(np.uint64), np.array([ke, pe]).view(np.uint64),
                                          err_msg=f"{len(r)} atoms")
            assert ke > 0 and pe < 0
            b.close()
    finally:
        hip.hipFree(dev)


def test_noalloc_energies_rejects_null_output():
    r, v, ids, L = fcc_inputs(4, seed=3)
    _, b, _ = make_pair(r, v, ids, L, gpu=True)
    b.init(1.0, 1.0, RCUT, SKIN, PRECISION_FP64)
    assert b.L.xnb_simulation_energies_noalloc(b._h, None) != 0
    b.close()
