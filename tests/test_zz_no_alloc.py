"""annie_hip_square_mul_noalloc (work registers in the library's static
device array) against annie_hip_square_mul (the context's own work
registers): word-exact results on two calls in a row, the second on a
smaller transform; in place (y == x); the context's registers untouched;
sizes above the limit and misaligned buffers rejected.  The caller's device
buffers come from the HIP runtime the engine itself is linked with."""
import ctypes

import numpy as np
import pytest

import annie_amd
from tests.helpers import usable_N

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


@pytest.fixture(scope="module")
def hip(gpu):
    lib = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so.7")
    lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    lib.hipFree.argtypes = [ctypes.c_void_p]
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return lib


class DevBuf:
    def __init__(self, hip, words):
        self.hip, self.words, self.ptr = hip, words, ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), words * 4) == 0

    def write(self, arr):
        a = np.ascontiguousarray(np.asarray(arr, dtype=np.uint32))
        assert a.size == self.words
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D) == 0

    def read(self):
        a = np.empty(self.words, dtype=np.uint32)
        assert self.hip.hipMemcpy(a.ctypes.data, self.ptr, a.nbytes, D2H) == 0
        return a

    def free(self):
        self.hip.hipFree(self.ptr)


def _one_case(hip, bits, seed, steps):
    a = 3
    N = usable_N(bits, seed, a)
    t = annie_amd.Transform(N, a, min_ln=5)
    print(f"case no-alloc bits={bits} seed={seed} plan K={t.K} LN={t.LN}", flush=True)
    assert t.K * t.n <= annie_amd.NOALLOC_MAX_KN
    words = t.K * t.n
    x, y = DevBuf(hip, words), DevBuf(hip, words)
    try:
        t.set1()
        for _ in range(steps):
            t.square_mul(a)
        x.write(np.ctypeslib.as_array(t.read_register(0)))
        r1_before = np.ctypeslib.as_array(t.read_register(1)).copy()
        t.square_mul_noalloc(x.ptr.value, y.ptr.value, a)
        t.sync()
        t.square_mul(a)  # the existing entry point on the context's r0
        want = np.ctypeslib.as_array(t.read_register(0))
        got = y.read()
        assert np.array_equal(got, want), f"{int(np.sum(got != want))} of {words} words differ"
        # in place, from the new r0: one more squaring each way
        x.write(want)
        t.square_mul_noalloc(x.ptr.value, x.ptr.value, 1)
        t.sync()
        t.square_mul(1)
        assert np.array_equal(x.read(), np.ctypeslib.as_array(t.read_register(0)))
        assert np.array_equal(np.ctypeslib.as_array(t.read_register(1)), r1_before)
        assert t.get(0) == pow(a, 2 ** (steps + 2) - 2, N)
    finally:
        x.free()
        y.free()
        t.close()


def test_noalloc_matches_square_mul_twice_second_smaller(hip):
    _one_case(hip, 300000, 21, 4)   # K n = 3 * 2^13
    _one_case(hip, 20000, 22, 3)    # a smaller transform, the same static array


def test_noalloc_at_the_size_limit(hip):
    _one_case(hip, 6000000, 24, 2)  # K n = 4 * 2^17 = the limit


def test_noalloc_rejects_large_and_misaligned(hip):
    lib = annie_amd.load_library()
    N = usable_N(8000000, 23, 3)
    t = annie_amd.Transform(N, 3, min_ln=5)
    assert t.K * t.n > annie_amd.NOALLOC_MAX_KN  # K n = 5 * 2^17
    buf = DevBuf(hip, t.K * t.n + 4)
    try:
        rc = lib.annie_hip_square_mul_noalloc(t._h, buf.ptr.value, buf.ptr.value, 3)
        assert rc == annie_amd.ANNIE_ERR_SIZE
        t.close()
        t = annie_amd.Transform(usable_N(20000, 22, 3), 3, min_ln=5)
        rc = lib.annie_hip_square_mul_noalloc(t._h, buf.ptr.value + 4, buf.ptr.value, 3)
        assert rc == annie_amd.ANNIE_ERR_ARG
    finally:
        buf.free()
        t.close()
