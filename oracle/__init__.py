"""ORACLE -- test infrastructure only.

Python bindings of the CPU oracle (liboracle.so: GMP MMGW restatement,
independent CPU RNS-NTT, the annie::check protocol) and of the reference's own
CPU transforms built in place (_ref/libannie_ref.so).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
package, and only as the checker.
"""
import ctypes
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libannie_ref.so")

u64p = ctypes.POINTER(ctypes.c_uint64)
u32p = ctypes.POINTER(ctypes.c_uint32)
sz = ctypes.c_size_t
vp = ctypes.c_void_p
ip = ctypes.POINTER(ctypes.c_int)

_lib = None
_ref = None


def build():
    subprocess.run(["make", "-C", HERE], check=True)


def limbs(x):
    """Non-negative int -> (ctypes uint64 array, limb count), little-endian limbs."""
    if x < 0:
        raise ValueError("negative integer")
    n = max(1, (x.bit_length() + 63) // 64)
    arr = (ctypes.c_uint64 * n).from_buffer_copy(x.to_bytes(8 * n, "little"))
    return arr, n


def from_limbs(arr, n):
    """First n little-endian uint64 limbs of a ctypes array -> int."""
    if n <= 0:
        return 0
    return int.from_bytes(ctypes.string_at(ctypes.addressof(arr), 8 * n), "little")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.or_gmp_create.argtypes = [u64p, sz, ctypes.c_int32]
        L.or_gmp_create.restype = vp
        for f in ("or_gmp_failed", "or_gmp_range_ok", "or_gmp_set1", "or_gmp_destroy"):
            getattr(L, f).argtypes = [vp]
        L.or_gmp_copy.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.or_gmp_square_mul.argtypes = [vp, ctypes.c_int32]
        L.or_gmp_mul.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.or_gmp_get.argtypes = [vp, ctypes.c_int, u64p, sz]
        L.or_gmp_get.restype = ctypes.c_long
        L.or_gmp_raw.argtypes = [vp, ctypes.c_int, u64p, sz]
        L.or_gmp_raw.restype = ctypes.c_long
        L.or_powm.argtypes = [u64p, sz, u64p, sz, u64p, sz, u64p, sz]
        L.or_powm.restype = ctypes.c_long
        L.or_gcd.argtypes = [u64p, sz, u64p, sz, u64p, sz]
        L.or_gcd.restype = ctypes.c_long
        L.or_base.argtypes = [u64p, sz, ctypes.c_int, ctypes.c_int32]
        L.or_base.restype = ctypes.c_int32
        L.or_check.argtypes = [u64p, sz, ctypes.c_int, ctypes.c_int32, ip, u64p, ip, ctypes.c_long, u64p, sz,
                               ctypes.POINTER(ctypes.c_long)]
        L.or_gl_block.argtypes = [sz]
        L.on_plan.argtypes = [u64p, sz, ctypes.c_int32, ip, ip, ip, ip]
        L.on_plan_d.argtypes = [ctypes.c_double, ctypes.c_long, ctypes.c_int32, ip, ip, ip, ip]
        L.on_create.argtypes = [u64p, sz, ctypes.c_int32]
        L.on_create.restype = vp
        L.on_failed.argtypes = [vp]
        L.on_info.argtypes = [vp, ip, ip, ip, ip]
        L.on_set1.argtypes = [vp]
        L.on_copy.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.on_square_mul.argtypes = [vp, ctypes.c_uint32]
        L.on_mul.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.on_get.argtypes = [vp, ctypes.c_int, u64p, sz]
        L.on_get.restype = ctypes.c_long
        L.on_read_register.argtypes = [vp, ctypes.c_int, u32p]
        L.on_read_tq.argtypes = [vp, u32p]
        L.on_write_register.argtypes = [vp, ctypes.c_int, u32p]
        L.on_destroy.argtypes = [vp]
        L.on_debug_square_stage.argtypes = [vp, ctypes.c_uint32, ctypes.c_int, u32p, u32p]
        L.on_ntt.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, u32p]
        L.on_crt.argtypes = [ctypes.c_int, u32p, u64p]
        L.on_tables.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, u32p]
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_LIB)


def ref():
    global _ref
    if _ref is None:
        L = ctypes.CDLL(REF_LIB)
        L.ref_create.argtypes = [ctypes.c_int, u64p, sz, ctypes.c_int32]
        L.ref_create.restype = vp
        for f in ("ref_init_failed", "ref_set1", "ref_destroy"):
            getattr(L, f).argtypes = [vp]
        L.ref_copy.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.ref_square_mul.argtypes = [vp, ctypes.c_int32]
        L.ref_mul.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.ref_get.argtypes = [vp, ctypes.c_int, u64p, sz]
        L.ref_get.restype = ctypes.c_long
        L.ref_info.argtypes = [vp, ctypes.c_char_p, sz]
        L.ref_check.argtypes = [u64p, sz, ctypes.c_int, ctypes.c_int32, ctypes.c_char_p]
        _ref = L
    return _ref


def powm(b, e, N):
    L = lib()
    B, nb = limbs(b)
    E, ne = limbs(e)
    M, nn = limbs(N)
    cap = nn + 1
    out = (ctypes.c_uint64 * cap)()
    r = L.or_powm(B, nb, E, ne, M, nn, out, cap)
    return from_limbs(out, r)


def gcd(a, b):
    """gcd(a, b) by GMP (subquadratic; Python's math.gcd is quadratic)."""
    L = lib()
    A, na = limbs(a)
    B, nb = limbs(b)
    cap = min(na, nb) + 1
    out = (ctypes.c_uint64 * cap)()
    r = L.or_gcd(A, na, B, nb, out, cap)
    return from_limbs(out, r)


def base(N, is_euler, a_ref=0):
    arr, n = limbs(N)
    return lib().or_base(arr, n, 1 if is_euler else 0, a_ref)


def gl_block(size):
    return lib().or_gl_block(size)


def check(N, is_euler=False, a_ref=0, stop_after=-1):
    """annie::check protocol on the exact GMP transform.

    Returns dict(is_prp, res64, gl_ok, a) or, with stop_after >= 0, the
    canonical value of r0 after that many loop iterations.
    """
    L = lib()
    a = base(N, is_euler, a_ref)
    arr, n = limbs(N)
    is_prp, gl_ok = ctypes.c_int(), ctypes.c_int()
    res64 = ctypes.c_uint64()
    cap = n + 1
    out = (ctypes.c_uint64 * cap)()
    nout = ctypes.c_long()
    rc = L.or_check(arr, n, 1 if is_euler else 0, a, ctypes.byref(is_prp), ctypes.byref(res64),
                    ctypes.byref(gl_ok), stop_after, out, cap, ctypes.byref(nout))
    if rc == 1:
        return {"init_failed": True, "a": a}
    if rc == 2:
        return {"value": from_limbs(out, nout.value), "a": a}
    return {"is_prp": bool(is_prp.value), "res64": res64.value, "gl_ok": bool(gl_ok.value), "a": a}


def plan(N, a):
    L = lib()
    arr, n = limbs(N)
    K, pck, LN, lb = (ctypes.c_int() for _ in range(4))
    L.on_plan(arr, n, a, ctypes.byref(K), ctypes.byref(pck), ctypes.byref(LN), ctypes.byref(lb))
    return K.value, pck.value, LN.value, lb.value


class _Machine:
    """Common register-machine wrapper (set1/copy/square_mul/mul/get)."""

    def __init__(self, N):
        self.N = N

    def _cap(self):
        return (self.N.bit_length() + 63) // 64 + 1


class GmpTransform(_Machine):
    """Exact MMGW with GMP (restatement of src/transformGMP.h)."""

    def __init__(self, N, a):
        super().__init__(N)
        self.L = lib()
        arr, n = limbs(N)
        self.h = self.L.or_gmp_create(arr, n, a)

    def init_failed(self):
        return bool(self.L.or_gmp_failed(self.h))

    def range_ok(self):
        return bool(self.L.or_gmp_range_ok(self.h))

    def set1(self):
        self.L.or_gmp_set1(self.h)

    def copy(self, d, s):
        self.L.or_gmp_copy(self.h, d, s)

    def square_mul(self, a):
        self.L.or_gmp_square_mul(self.h, a)

    def mul(self, d, s):
        self.L.or_gmp_mul(self.h, d, s)

    def get(self, s):
        cap = self._cap()
        out = (ctypes.c_uint64 * cap)()
        return from_limbs(out, self.L.or_gmp_get(self.h, s, out, cap))

    def raw(self, s):
        cap = self._cap() + 2
        out = (ctypes.c_uint64 * cap)()
        return from_limbs(out, self.L.or_gmp_raw(self.h, s, out, cap))

    def close(self):
        if self.h:
            self.L.or_gmp_destroy(self.h)
            self.h = None

    __del__ = close


class NttTransform(_Machine):
    """CPU RNS-NTT MMGW mirroring the device representation register for register."""

    def __init__(self, N, a):
        super().__init__(N)
        self.L = lib()
        arr, n = limbs(N)
        self.h = self.L.on_create(arr, n, a)
        K, pck, LN, lb = (ctypes.c_int() for _ in range(4))
        self.L.on_info(self.h, ctypes.byref(K), ctypes.byref(pck), ctypes.byref(LN), ctypes.byref(lb))
        self.K, self.pck, self.LN, self.lbase = K.value, pck.value, LN.value, lb.value
        self.n = 1 << self.LN

    def init_failed(self):
        return bool(self.L.on_failed(self.h))

    def set1(self):
        self.L.on_set1(self.h)

    def copy(self, d, s):
        self.L.on_copy(self.h, d, s)

    def square_mul(self, a):
        self.L.on_square_mul(self.h, a)

    def mul(self, d, s):
        self.L.on_mul(self.h, d, s)

    def get(self, s):
        cap = self._cap()
        out = (ctypes.c_uint64 * cap)()
        return from_limbs(out, self.L.on_get(self.h, s, out, cap))

    def read_register(self, r):
        out = (ctypes.c_uint32 * (self.K * self.n))()
        self.L.on_read_register(self.h, r, out)
        return out

    def read_tq(self):
        out = (ctypes.c_uint32 * (3 * self.n))()
        self.L.on_read_tq(self.h, out)
        return out

    def write_register(self, r, words):
        buf = (ctypes.c_uint32 * len(words))(*words)
        self.L.on_write_register(self.h, r, buf)

    def debug_square_stage(self, a, stage):
        """(out_r, out_q) intermediates of one squaring of r0 (see ntt_oracle.cpp)."""
        out_r = (ctypes.c_uint32 * (self.K * self.n))()
        out_q = (ctypes.c_uint32 * (max(self.K, 3) * self.n))()
        self.L.on_debug_square_stage(self.h, a, stage, out_r, out_q)
        return out_r, out_q

    def close(self):
        if self.h:
            self.L.on_destroy(self.h)
            self.h = None

    __del__ = close


class RefTransform(_Machine):
    """The reference's own transformGMP (kind 0) or transformCPU (kind 1)."""

    def __init__(self, N, a, kind):
        super().__init__(N)
        self.L = ref()
        arr, n = limbs(N)
        self.h = self.L.ref_create(kind, arr, n, a)

    def init_failed(self):
        return bool(self.L.ref_init_failed(self.h))

    def set1(self):
        self.L.ref_set1(self.h)

    def copy(self, d, s):
        self.L.ref_copy(self.h, d, s)

    def square_mul(self, a):
        self.L.ref_square_mul(self.h, a)

    def mul(self, d, s):
        self.L.ref_mul(self.h, d, s)

    def get(self, s):
        cap = self._cap()
        out = (ctypes.c_uint64 * cap)()
        return from_limbs(out, self.L.ref_get(self.h, s, out, cap))

    def info(self):
        buf = ctypes.create_string_buffer(256)
        self.L.ref_info(self.h, buf, 256)
        return buf.value.decode()

    def close(self):
        if self.h:
            self.L.ref_destroy(self.h)
            self.h = None

    __del__ = close


def ntt_plane(LN, k, nega, inverse, values):
    """Oracle forward/inverse transform of one prime plane."""
    buf = (ctypes.c_uint32 * len(values))(*values)
    lib().on_ntt(LN, k, 1 if nega else 0, 1 if inverse else 0, buf)
    return list(buf)


def root_table(LN, k, t):
    """Oracle root table t (0 cyc fwd, 1 cyc inv, 2 nega fwd, 3 nega inv), plain residues."""
    out = (ctypes.c_uint32 * (1 << LN))()
    lib().on_tables(LN, k, t, out)
    return list(out)


def crt(K, residues):
    buf = (ctypes.c_uint32 * K)(*residues)
    out = (ctypes.c_uint64 * 3)()
    lib().on_crt(K, buf, out)
    v = from_limbs(out, 3)
    return v - (1 << 192) if v >> 191 else v
