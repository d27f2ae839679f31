"""nb_direct_sum_f64_noalloc: the allocation-free variant of the FP64 direct
sum (device arrays in, partial sums in the library's fixed scratch array)
against nb_direct_sum_f64, bit for bit, on two calls in a row, the second
with a smaller input. Device buffers come from the HIP runtime through
ctypes; the entry point itself allocates nothing."""
import ctypes as C

import numpy as np
import pytest

from nbody2d import abi
from nbody2d import ics

pytestmark = pytest.mark.gpu


class Hip:
    """The few HIP runtime calls the test needs for device buffers."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def check(self, rc, what):
        assert rc == 0, f"{what} failed: {rc}"

    def to_device(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        self.check(self.L.hipMalloc(C.byref(p), max(a.nbytes, 16)), "hipMalloc")
        self.bufs.append(p)
        self.check(self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1), "H2D")
        return p

    def empty(self, nbytes):
        p = C.c_void_p()
        self.check(self.L.hipMalloc(C.byref(p), max(nbytes, 16)), "hipMalloc")
        self.bufs.append(p)
        return p

    def to_host(self, p, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        self.check(self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2), "D2H")
        return out

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def run_noalloc(hip, pos, mass, G, eps, targets):
    n, m = len(mass), len(targets)
    dp, dm = hip.to_device(pos.reshape(-1)), hip.to_device(mass)
    dt, da = hip.to_device(targets.astype(np.int64)), hip.empty(16 * m)
    rc = abi.lib().nb_direct_sum_f64_noalloc(n, dp, dm, G, eps, m, dt, da, None)
    assert rc == abi.NB_OK, abi.lib().nb_last_error(None)
    hip.check(hip.L.hipDeviceSynchronize(), "sync")
    return hip.to_host(da, (m, 2), np.float64)


def test_noalloc_direct_sum_equals_allocating_entry_point():
    hip = Hip()
    try:
        cases = [(ics.exp_disk(20000, seed=31), 3000, 0.01),
                 (ics.uniform_disk(5000, seed=32), 700, 0.0)]   # second call: smaller
        for (m_, p, v), k, eps in cases:
            t = ics.sample_targets(len(m_), k, seed=33)
            print(f"[gpu] noalloc n={len(m_)} m={k} eps={eps}", flush=True)
            got = run_noalloc(hip, p, m_, 2.0, eps, t)
            want = abi.direct_sum_f64(p, m_, 2.0, eps, t)
            assert np.array_equal(got, want)
    finally:
        hip.free()


def test_noalloc_direct_sum_rejects_oversized_and_marks_bad_targets():
    hip = Hip()
    try:
        m_, p, v = ics.exp_disk(1000, seed=34)
        L = abi.lib()
        too_many = abi.NOALLOC_SCRATCH_BYTES // 16 + 1
        rc = L.nb_direct_sum_f64_noalloc(1000, C.c_void_p(1), C.c_void_p(1), 1.0, 0.0, too_many,
                                         C.c_void_p(1), C.c_void_p(1), None)
        assert rc == abi.NB_ERR_INVALID_ARG
        t = np.array([0, 5, 1000, -1, 999], dtype=np.int64)  # 1000 and -1 are out of range
        got = run_noalloc(hip, p, m_, 1.0, 0.01, t)
        assert np.isnan(got[2]).all() and np.isnan(got[3]).all()
        want = abi.direct_sum_f64(p, m_, 1.0, 0.01, t[[0, 1, 4]])
        assert np.array_equal(got[[0, 1, 4]], want)
    finally:
        hip.free()
