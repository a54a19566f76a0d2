"""annie_amd -- MI355X (gfx950) engine for annie's modular-squaring loop.

Python view of the C ABI in include/annie_hip.h.  `Transform` mirrors the
reference's `transform` interface (src/transform.h:18-45) with the same method
names, argument meaning and error behaviour: sizes outside the supported
range raise RuntimeError (the reference throws std::runtime_error,
src/transformGPU.h:853-857), and gcd(N, R) > 1 yields an object whose
init_failed() is True and whose get_factor() returns the common factor
(src/transformGPU.h:652-661).

The library is the native HIP build in annie_amd/lib; importing this module
never falls back to another implementation: a missing library raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(_HERE, "lib")
HIP_LIB_PATH = os.path.join(LIBDIR, "libannie_hip.so")
HOST_LIB_PATH = os.path.join(LIBDIR, "libannie_host.so")

Reg0, Reg1, Reg2 = 0, 1, 2

ANNIE_OK = 0
ANNIE_INIT_FAILED = 1
ANNIE_ERR = -1
ANNIE_ERR_SIZE = -2
ANNIE_ERR_ARG = -3
ANNIE_ERR_NOMEM = -4
# Largest K * n words per register of annie_hip_square_mul_noalloc
# (ANNIE_HIP_NOALLOC_MAX_KN in include/annie_hip.h).
NOALLOC_MAX_KN = 1 << 19

_lib = None


def int_to_limbs(x):
    """Non-negative int -> (ctypes uint64 array, limb count), little-endian limbs."""
    if x < 0:
        raise ValueError("negative integer")
    n = max(1, (x.bit_length() + 63) // 64)
    arr = (ctypes.c_uint64 * n).from_buffer_copy(x.to_bytes(8 * n, "little"))
    return arr, n


def limbs_to_int(arr, n):
    """First n little-endian uint64 limbs of a ctypes array -> int."""
    if n <= 0:
        return 0
    return int.from_bytes(ctypes.string_at(ctypes.addressof(arr), 8 * n), "little")


def load_library(path=HIP_LIB_PATH):
    """Load libannie_hip.so (raises OSError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError(f"{path} is missing: build it with `python -m annie_amd.build`")
    lib = ctypes.CDLL(path)
    u64p = ctypes.POINTER(ctypes.c_uint64)
    ip = ctypes.POINTER(ctypes.c_int)
    sz = ctypes.c_size_t
    h = ctypes.c_void_p
    lib.annie_hip_select_plan.argtypes = [u64p, sz, ctypes.c_int32, ip, ip, ip, ip]
    lib.annie_hip_create.argtypes = [u64p, sz, ctypes.c_int32, ctypes.c_int, ctypes.c_int, ctypes.POINTER(h)]
    lib.annie_hip_get_factor.argtypes = [h, u64p, sz, ctypes.POINTER(sz)]
    lib.annie_hip_set1.argtypes = [h]
    lib.annie_hip_copy.argtypes = [h, sz, sz]
    lib.annie_hip_square_mul.argtypes = [h, ctypes.c_int32]
    lib.annie_hip_square_mul_bits.argtypes = [h, ctypes.POINTER(ctypes.c_uint8), sz, ctypes.c_int32]
    lib.annie_hip_square_mul_noalloc.argtypes = [h, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32]
    lib.annie_hip_mul.argtypes = [h, sz, sz]
    lib.annie_hip_get.argtypes = [h, sz, u64p, sz, ctypes.POINTER(sz)]
    lib.annie_hip_checkpoint_size.argtypes = [h]
    lib.annie_hip_checkpoint_size.restype = sz
    lib.annie_hip_save_checkpoint.argtypes = [h, ctypes.c_void_p, sz]
    lib.annie_hip_load_checkpoint.argtypes = [h, ctypes.c_void_p, sz]
    lib.annie_hip_info.argtypes = [h, ip, ip, ip, ip]
    lib.annie_hip_schedule_fits.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.annie_hip_set_schedule.argtypes = [h, ctypes.c_int, ctypes.c_int]
    lib.annie_hip_get_schedule.argtypes = [h, ip, ip]
    lib.annie_hip_sync.argtypes = [h]
    lib.annie_hip_read_register.argtypes = [h, sz, ctypes.POINTER(ctypes.c_uint32), sz]
    lib.annie_hip_write_register.argtypes = [h, sz, ctypes.POINTER(ctypes.c_uint32), sz]
    lib.annie_hip_debug_square_stage.argtypes = [h, ctypes.c_int32, ctypes.c_int]
    lib.annie_hip_read_tq.argtypes = [h, ctypes.POINTER(ctypes.c_uint32), sz]
    lib.annie_hip_time_square_mul_bits.argtypes = [h, ctypes.POINTER(ctypes.c_uint8), sz, ctypes.c_int32,
                                                    ctypes.POINTER(ctypes.c_float)]
    lib.annie_hip_profile_begin.argtypes = [h]
    lib.annie_hip_profile_end.argtypes = [h, ctypes.c_char_p, sz]
    u32p = ctypes.POINTER(ctypes.c_uint32)
    lib.annie_hip_plan_constants.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, u32p, sz]
    lib.annie_hip_twiddles.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, u32p, u32p, sz]
    lib.annie_hip_batch_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, sz, ctypes.c_int,
                                           ctypes.POINTER(h)]
    lib.annie_hip_batch_bytes.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(sz), ctypes.POINTER(sz)]
    lib.annie_hip_batch_load.argtypes = [h, sz, u64p, sz, ctypes.c_int32, u64p, sz, ctypes.POINTER(sz)]
    lib.annie_hip_batch_set1.argtypes = [h, u32p, sz]
    lib.annie_hip_batch_copy.argtypes = [h, sz, sz, u32p, sz]
    lib.annie_hip_batch_square_mul.argtypes = [h, u32p, u32p, sz]
    lib.annie_hip_batch_mul.argtypes = [h, sz, sz, u32p, sz]
    lib.annie_hip_batch_square_run.argtypes = [h, u32p, sz, u32p, sz]
    lib.annie_hip_batch_run_supported.argtypes = [h]
    lib.annie_hip_batch_get.argtypes = [h, sz, u32p, sz, u64p, sz, ctypes.POINTER(sz)]
    lib.annie_hip_batch_sync.argtypes = [h]
    lib.annie_hip_batch_set_graphs.argtypes = [h, ctypes.c_int]
    lib.annie_hip_destroy.argtypes = [h]
    lib.annie_hip_destroy.restype = None
    lib.annie_hip_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def _err(lib):
    return lib.annie_hip_last_error().decode(errors="replace")


def select_plan(N, a):
    """(K, pck, LN, lbase) chosen for N and base a (src/transformGPU.h:822-850)."""
    lib = load_library()
    arr, n = int_to_limbs(N)
    K, pck, LN, lb = (ctypes.c_int() for _ in range(4))
    rc = lib.annie_hip_select_plan(arr, n, a, ctypes.byref(K), ctypes.byref(pck), ctypes.byref(LN), ctypes.byref(lb))
    if rc != ANNIE_OK:
        raise RuntimeError(_err(lib))
    return K.value, pck.value, LN.value, lb.value


def plan_constants(LN, K, pck, lbase):
    """Per-prime constants of a plan (see annie_hip_plan_constants)."""
    lib = load_library()
    out = (ctypes.c_uint32 * 50)()
    if lib.annie_hip_plan_constants(LN, K, pck, lbase, out, 50) != ANNIE_OK:
        raise RuntimeError(_err(lib))
    v = list(out)
    return {"p": v[0:5], "q": v[5:10], "sq_scale": v[10:15], "c32": v[15:20], "c64": v[20:25],
            "ginv": [v[25 + 5 * i:30 + 5 * i] for i in range(5)]}


def schedule_fits(LN, row_lines, col_cols):
    """Whether an NTT launch schedule fits transforms of length 2^LN."""
    return bool(load_library().annie_hip_schedule_fits(int(LN), int(row_lines), int(col_cols)))


def schedule_default(LN):
    """The engine's default launch schedule (row lines, column count) for
    length 2^LN, 0 where a pass has no knob (rows of 64 threads or more,
    columns of 2^10 or more, or no column pass)."""
    m1 = 0 if LN <= 13 else LN // 2
    m2 = LN - m1
    tr = (1 << m2) // 16
    lines = 0 if tr >= 64 else min(64 // tr, 1 << m1)
    cols = 0
    if 0 < m1 < 10:
        cols = min(256 // ((1 << m1) // 16), 16, 1 << (LN - m1))
    return lines, cols


def schedules(LN):
    """The launch schedules (row lines, column count) worth timing at length
    2^LN: (0, 0) (the default) first, then every other pair that fits with
    at least one 64-lane wave per workgroup in each pass that has a knob."""
    m1 = 0 if LN <= 13 else LN // 2
    m2 = LN - m1
    tr, tc = (1 << m2) // 16, (1 << m1) // 16
    dl, dc = schedule_default(LN)
    pw = [1 << i for i in range(11)]
    ls = [x for x in pw if x * tr >= 64] if dl else [0]
    cs = [x for x in pw if x * tc >= 64] if dc else [0]
    out = [(0, 0)]
    for rl in ls:
        for cc in cs:
            if (rl, cc) != (dl, dc) and schedule_fits(LN, rl, cc):
                out.append((rl, cc))
    return out


def batch_bytes(K, LN):
    """(bytes per slot, bytes per batch) of device memory for a batch of plan (K, LN)."""
    lib = load_library()
    per, fixed = ctypes.c_size_t(), ctypes.c_size_t()
    if lib.annie_hip_batch_bytes(K, LN, ctypes.byref(per), ctypes.byref(fixed)) != ANNIE_OK:
        raise RuntimeError(_err(lib))
    return per.value, fixed.value


def twiddles(LN, k, table):
    """(w, wq) lists of a heap-indexed twiddle table (Montgomery form)."""
    lib = load_library()
    n = 1 << LN
    w = (ctypes.c_uint32 * n)()
    wq = (ctypes.c_uint32 * n)()
    if lib.annie_hip_twiddles(LN, k, table, w, wq, n) != ANNIE_OK:
        raise RuntimeError(_err(lib))
    return list(w), list(wq)


class Transform:
    """MMGW register machine for one N on one GPU (src/transform.h:18-45)."""

    Reg0, Reg1, Reg2 = Reg0, Reg1, Reg2

    def __init__(self, N, a, device=0, min_ln=10):
        lib = load_library()
        self._lib = lib
        self._h = ctypes.c_void_p()
        self.N = int(N)
        arr, n = int_to_limbs(self.N)
        rc = lib.annie_hip_create(arr, n, int(a), int(device), int(min_ln), ctypes.byref(self._h))
        if rc < 0:
            self._h = ctypes.c_void_p()
            raise RuntimeError(_err(lib))
        self._init_failed = rc == ANNIE_INIT_FAILED
        K, pck, LN, lb = (ctypes.c_int() for _ in range(4))
        lib.annie_hip_info(self._h, ctypes.byref(K), ctypes.byref(pck), ctypes.byref(LN), ctypes.byref(lb))
        self.K, self.pck, self.LN, self.lbase = K.value, pck.value, LN.value, lb.value
        self.n = 1 << self.LN

    def _check(self, rc):
        if rc != ANNIE_OK:
            raise RuntimeError(_err(self._lib))

    def init_failed(self):
        return self._init_failed

    def get_factor(self):
        cap = (self.N.bit_length() + 63) // 64 + 1
        out = (ctypes.c_uint64 * cap)()
        nout = ctypes.c_size_t()
        self._check(self._lib.annie_hip_get_factor(self._h, out, cap, ctypes.byref(nout)))
        return limbs_to_int(out, nout.value)

    def set1(self):
        self._check(self._lib.annie_hip_set1(self._h))

    def copy(self, dst, src):
        self._check(self._lib.annie_hip_copy(self._h, dst, src))

    def square_mul(self, a):
        self._check(self._lib.annie_hip_square_mul(self._h, int(a)))

    def square_mul_noalloc(self, x_ptr, y_ptr, a):
        """y = a * x^2 on caller-owned device buffers (K planes of n u32 words,
        plane-major) with the library's static work array; allocates nothing."""
        self._check(self._lib.annie_hip_square_mul_noalloc(self._h, int(x_ptr), int(y_ptr), int(a)))

    def square_mul_bits(self, bits, a):
        """r0 = (bit ? a : 1) * r0^2 for each bit of `bits` (bytes or list of 0/1)."""
        buf = (ctypes.c_uint8 * len(bits))(*bits)
        self._check(self._lib.annie_hip_square_mul_bits(self._h, buf, len(bits), int(a)))

    def time_square_mul_bits(self, bits, a):
        """As square_mul_bits; returns the GPU time of the sequence in ms (HIP events)."""
        buf = (ctypes.c_uint8 * len(bits))(*bits)
        ms = ctypes.c_float()
        self._check(self._lib.annie_hip_time_square_mul_bits(self._h, buf, len(bits), int(a), ctypes.byref(ms)))
        return ms.value

    def profile_begin(self):
        self._check(self._lib.annie_hip_profile_begin(self._h))

    def profile_end(self):
        """{kernel name: (launches, total ms)} since profile_begin()."""
        buf = ctypes.create_string_buffer(1 << 16)
        self._check(self._lib.annie_hip_profile_end(self._h, buf, 1 << 16))
        out = {}
        for line in buf.value.decode().splitlines():
            name, cnt, tot = line.split()
            out[name] = (int(cnt), float(tot))
        return out

    def mul(self, dst, src):
        self._check(self._lib.annie_hip_mul(self._h, dst, src))

    def get(self, src):
        cap = (self.N.bit_length() + 63) // 64 + 1
        out = (ctypes.c_uint64 * cap)()
        nout = ctypes.c_size_t()
        self._check(self._lib.annie_hip_get(self._h, src, out, cap, ctypes.byref(nout)))
        return limbs_to_int(out, nout.value)

    def save_checkpoint(self):
        size = self._lib.annie_hip_checkpoint_size(self._h)
        buf = ctypes.create_string_buffer(size)
        self._check(self._lib.annie_hip_save_checkpoint(self._h, buf, size))
        return buf.raw

    def read_checkpoint(self, data):
        buf = ctypes.create_string_buffer(bytes(data), len(data))
        rc = self._lib.annie_hip_load_checkpoint(self._h, buf, len(data))
        return rc == ANNIE_OK

    def has_checkpoint(self):
        return True

    def set_schedule(self, row_lines=0, col_cols=0):
        """NTT launch schedule (lines per row-pass workgroup for short rows,
        columns per column-pass workgroup for columns of up to 2^9); 0 keeps
        the default.  Results do not depend on it."""
        self._check(self._lib.annie_hip_set_schedule(self._h, int(row_lines), int(col_cols)))

    def get_schedule(self):
        """(row lines, column count) in effect; 0 where a pass has no knob."""
        a, b = ctypes.c_int(), ctypes.c_int()
        self._check(self._lib.annie_hip_get_schedule(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def sync(self):
        self._check(self._lib.annie_hip_sync(self._h))

    def read_register(self, reg):
        """Raw residues of a device register (K planes of n words, plane-major)."""
        words = self.K * self.n
        out = (ctypes.c_uint32 * words)()
        self._check(self._lib.annie_hip_read_register(self._h, reg, out, words))
        return out

    def write_register(self, reg, words):
        buf = (ctypes.c_uint32 * len(words))(*words)
        self._check(self._lib.annie_hip_write_register(self._h, reg, buf, len(words)))

    def debug_square_stage(self, a, stage):
        self._check(self._lib.annie_hip_debug_square_stage(self._h, int(a), int(stage)))

    def read_tq(self):
        words = 3 * self.n
        out = (ctypes.c_uint32 * words)()
        self._check(self._lib.annie_hip_read_tq(self._h, out, words))
        return out

    def get_info_short(self):
        return f"{self.K}-prime NTT of length {self.n}, packed-base = 2^{{{self.pck}x{self.lbase}}}"

    def close(self):
        if self._h:
            self._lib.annie_hip_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class CheckOpts(ctypes.Structure):
    _fields_ = [("is_euler", ctypes.c_int), ("a_ref", ctypes.c_int32), ("device", ctypes.c_int),
                ("gmp_threshold_digits", ctypes.c_int), ("min_ln", ctypes.c_int), ("workdir", ctypes.c_char_p),
                ("checkpoint_seconds", ctypes.c_double), ("quit_after", ctypes.c_long),
                ("corrupt_at", ctypes.c_long), ("verbose", ctypes.c_int), ("expression", ctypes.c_char_p)]


class CheckResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("is_prp", ctypes.c_int), ("res64", ctypes.c_uint64),
                ("gl_ok", ctypes.c_int), ("a", ctypes.c_int32), ("resumed", ctypes.c_int),
                ("iterations", ctypes.c_long), ("seconds", ctypes.c_double), ("message", ctypes.c_char * 1024)]


STATUS = {0: "success", 1: "failed", 2: "aborted", 3: "factor", 4: "decided", -1: "error"}
_host = None


def load_host_library(path=HOST_LIB_PATH):
    """Load libannie_host.so (the test protocol host over the engine)."""
    global _host
    if _host is None:
        load_library()
        if not os.path.exists(path):
            raise OSError(f"{path} is missing: build it with `python -m annie_amd.build`")
        lib = ctypes.CDLL(path)
        lib.annie_check.argtypes = [ctypes.POINTER(ctypes.c_uint64), ctypes.c_size_t, ctypes.POINTER(CheckOpts),
                                    ctypes.POINTER(CheckResult)]
        lib.annie_checkpoint_name.argtypes = [ctypes.POINTER(ctypes.c_uint64), ctypes.c_size_t, ctypes.c_char_p,
                                              ctypes.c_size_t]
        lib.annie_crc32.argtypes = [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_size_t]
        lib.annie_crc32.restype = ctypes.c_uint32
        _host = lib
    return _host


def check(N, is_euler=False, a=0, device=0, gmp_threshold_digits=10000, min_ln=10, workdir=None,
          checkpoint_seconds=600.0, quit_after=-1, corrupt_at=-1, verbose=False):
    """Run the probable-primality protocol of annie::check (src/annie.h:217-430) on N.

    Returns a dict: status, is_prp, res64, gl_ok, a, resumed, iterations, seconds, message.
    """
    lib = load_host_library()
    arr, n = int_to_limbs(N)
    o = CheckOpts(1 if is_euler else 0, a, device, gmp_threshold_digits, min_ln,
                  workdir.encode() if workdir else None, checkpoint_seconds, quit_after, corrupt_at,
                  1 if verbose else 0, None)
    r = CheckResult()
    lib.annie_check(arr, n, ctypes.byref(o), ctypes.byref(r))
    return {"status": STATUS.get(r.status, r.status), "is_prp": bool(r.is_prp), "res64": r.res64,
            "gl_ok": bool(r.gl_ok), "a": r.a, "resumed": bool(r.resumed), "iterations": r.iterations,
            "seconds": r.seconds, "message": r.message.decode(errors="replace")}


def request_quit():
    load_host_library().annie_request_quit()


def clear_quit():
    load_host_library().annie_clear_quit()


class BatchTest(ctypes.Structure):
    _fields_ = [("N", ctypes.POINTER(ctypes.c_uint64)), ("nlimbs", ctypes.c_size_t), ("is_euler", ctypes.c_int),
                ("a_ref", ctypes.c_int32), ("arrival", ctypes.c_double)]


class BatchOpts(ctypes.Structure):
    _fields_ = [("engine", ctypes.c_int), ("device", ctypes.c_int), ("max_slots", ctypes.c_size_t),
                ("hbm_budget_mb", ctypes.c_double), ("realtime", ctypes.c_int), ("time_limit", ctypes.c_double),
                ("verbose", ctypes.c_int), ("graphs", ctypes.c_int), ("flush_ticks", ctypes.c_int),
                ("run_steps", ctypes.c_int), ("run_min_active", ctypes.c_int), ("host_threads", ctypes.c_int),
                ("row_lines", ctypes.c_int), ("col_cols", ctypes.c_int)]


class BatchResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("is_prp", ctypes.c_int), ("res64", ctypes.c_uint64),
                ("gl_ok", ctypes.c_int), ("a", ctypes.c_int32), ("iterations", ctypes.c_long),
                ("arrival", ctypes.c_double), ("admitted", ctypes.c_double), ("finished", ctypes.c_double),
                ("message", ctypes.c_char * 256)]


class BatchStats(ctypes.Structure):
    _fields_ = [("seconds", ctypes.c_double), ("squarings", ctypes.c_double), ("classes", ctypes.c_int),
                ("slots", ctypes.c_size_t), ("setup_seconds", ctypes.c_double)]


def batch_run(tests, engine="gpu", device=0, max_slots=0, hbm_budget_mb=0.0, realtime=False, time_limit=0.0,
              verbose=False, graphs=False, flush_ticks=0, run_steps=0, run_min_active=0, host_threads=0,
              row_lines=0, col_cols=0):
    """Run many tests together, one engine batch per length class, with
    continuous admission (annie_batch_run in include/annie_host.h).

    `tests`: iterable of dicts {"N": int, "is_euler": bool, "a": int (Fermat
    base, 0 = 2), "arrival": seconds}.  engine "gpu" is the gfx950 engine;
    "gmp" runs the same driver on exact GMP transforms on the CPU.
    run_steps > 0: classes with at least run_min_active running tests take up
    to run_steps squarings per test in one launch (squaring runs, lengths
    2^9 .. 2^13); results are the same either way.  host_threads: host
    workers driving the classes (0 = one per class, at most 16).
    row_lines / col_cols: NTT launch schedule (Transform.set_schedule) of
    every class whose transform length it fits; 0 = default.
    Returns (results, stats): one dict per test (status, is_prp, res64, gl_ok,
    a, iterations, arrival, admitted, finished, message) and a stats dict.
    """
    lib = load_host_library()
    if not hasattr(lib, "_batch_ready"):
        lib.annie_batch_run.argtypes = [ctypes.POINTER(BatchTest), ctypes.c_size_t, ctypes.POINTER(BatchOpts),
                                        ctypes.POINTER(BatchResult), ctypes.POINTER(BatchStats)]
        lib._batch_ready = True
    tests = list(tests)
    keep = []
    arr = (BatchTest * max(len(tests), 1))()
    for i, t in enumerate(tests):
        limbs, n = int_to_limbs(int(t["N"]))
        keep.append(limbs)
        arr[i] = BatchTest(limbs, n, 1 if t.get("is_euler") else 0, int(t.get("a", 0)), float(t.get("arrival", 0.0)))
    opts = BatchOpts(0 if engine == "gpu" else 1, device, max_slots, hbm_budget_mb, 1 if realtime else 0,
                     time_limit, 1 if verbose else 0, 1 if graphs else 0, int(flush_ticks), int(run_steps),
                     int(run_min_active), int(host_threads), int(row_lines), int(col_cols))
    res = (BatchResult * max(len(tests), 1))()
    st = BatchStats()
    rc = lib.annie_batch_run(arr, len(tests), ctypes.byref(opts), res, ctypes.byref(st))
    out = []
    for i in range(len(tests)):
        r = res[i]
        out.append({"status": STATUS.get(r.status, r.status), "is_prp": bool(r.is_prp), "res64": r.res64,
                    "gl_ok": bool(r.gl_ok), "a": r.a, "iterations": r.iterations, "arrival": r.arrival,
                    "admitted": r.admitted, "finished": r.finished,
                    "message": r.message.decode(errors="replace")})
    stats = {"rc": rc, "seconds": st.seconds, "squarings": st.squarings, "classes": st.classes, "slots": st.slots,
             "setup_seconds": st.setup_seconds}
    return out, stats
