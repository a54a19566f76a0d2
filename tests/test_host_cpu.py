"""CPU tests of the native libraries: the C ABI surface, the protocol host on
its GMP path (N below 10,000 digits, as the reference), checkpoint files and
failure handling, and the command-line front end."""
import ctypes
import os
import re
import subprocess

import pytest

import annie_amd
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def _declared(header):
    text = open(os.path.join(INCLUDE, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(annie_\w+)\s*\(", text)))


@pytest.mark.parametrize("header,path", [("annie_hip.h", annie_amd.HIP_LIB_PATH),
                                         ("annie_host.h", annie_amd.HOST_LIB_PATH)])
def test_library_exports_every_declared_symbol(header, path):
    lib = ctypes.CDLL(path)
    names = _declared(header)
    assert names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_plan_entry_needs_no_gpu():
    assert annie_amd.select_plan((1 << 132899999) | 1, 3) == (5, 4, 21, 16)


def _fib(n):
    a, b = 0, 1
    for _ in range(n):
        a, b = b, a + b
    return a


@pytest.mark.parametrize("i", [359, 802, 1601, 2971])
def test_host_protocol_gmp_path(i):
    N = _fib(i)
    r = annie_amd.check(N)
    o = oracle.check(N)
    assert r["status"] == "success"
    assert r["is_prp"] == o["is_prp"]
    assert r["res64"] == o["res64"]
    assert r["gl_ok"]


def test_host_protocol_small_and_even():
    assert annie_amd.check(1000003)["status"] == "decided"
    assert annie_amd.check(1000003)["is_prp"]
    assert annie_amd.check(2 ** 70)["message"].endswith("is even.")


def test_host_protocol_euler_proth():
    # 3*2^189+1 is prime (Proth); Euler test with the reference's base choice.
    N = 3 * 2 ** 189 + 1
    r = annie_amd.check(N, is_euler=True)
    assert r["is_prp"] and r["gl_ok"]
    assert r["a"] == oracle.base(N, True)


def test_host_protocol_detects_corruption():
    """A register corrupted mid-loop must make the Gerbicz-Li check fail (src/annie.h:343-404)."""
    N = _fib(2971)
    r = annie_amd.check(N, corrupt_at=1000)
    assert r["status"] == "failed"
    assert not r["gl_ok"]


def test_host_protocol_abort(tmp_path):
    r = annie_amd.check(_fib(2971), workdir=str(tmp_path), quit_after=100)
    assert r["status"] == "aborted" and r["iterations"] == 100


def test_crc32_matches_reference_algorithm():
    """src/file.h:56-90 (reflected 0xedb88320); zlib's crc32 is the same function."""
    import zlib
    lib = annie_amd.load_host_library()
    for data in (b"", b"a", b"annie checkpoint", bytes(range(256)) * 3):
        assert lib.annie_crc32(0, data, len(data)) == zlib.crc32(data)
    # running value: crc(a + b) == crc(crc(a), b)
    a, b = b"first part", b"second part"
    assert lib.annie_crc32(lib.annie_crc32(0, a, len(a)), b, len(b)) == zlib.crc32(a + b)


def test_checkpoint_name_closed_form():
    """a_<16 hex>.chk with hash = floor(mantissa * 2^52) * digits + limb0 (src/annie.h:288-291)."""
    import math
    N = _fib(9311)
    lib = annie_amd.load_host_library()
    arr, n = annie_amd.int_to_limbs(N)
    buf = ctypes.create_string_buffer(64)
    assert lib.annie_checkpoint_name(arr, n, buf, 64) == 0
    m, e = math.frexp(float(N >> (N.bit_length() - 60)))  # mantissa in [0.5, 1)
    h = (int(m * 2 ** 52) * len(str(N)) + (N & (2 ** 64 - 1))) % 2 ** 64
    assert buf.value.decode() == f"a_{h:016X}.chk"


def test_cli_expression_front_end(tmp_path):
    exe = os.path.join(annie_amd.LIBDIR, "annie")
    w = str(tmp_path)
    out = subprocess.run([exe, "-q", "fibonacci(359)", "-w", w], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert "is a Fermat probable prime" in out.stdout
    out = subprocess.run([exe, "-q", "(2^127-1)*3+2*0", "-w", w], capture_output=True, text=True, timeout=120)
    assert "composite" in out.stdout
    out = subprocess.run([exe, "-q", "10!+1"], capture_output=True, text=True, timeout=60, cwd=w)
    assert "10!+1 is" in out.stdout  # the expression as typed (src/annie.h:219)
    out = subprocess.run([exe, "-q", "unknown(3)"], capture_output=True, text=True, timeout=60, cwd=w)
    assert out.returncode == 1 and "unknown function" in out.stderr


@pytest.mark.parametrize("expr,value", [("-2^2+11", 7), ("2*-3+13", 7), ("2^3^2-1", 511), ("-(2)^2+13", 9),
                                        ("3!-(-1)", 7), ("10-2-1", 7), ("2^2^0+2", 4)])
def test_cli_expression_precedence_as_pari(tmp_path, expr, value):
    """Operator precedence of PARI/GP, the reference's evaluator: unary minus
    below '^' (-2^2 = -4) and above '*', '^' right associative.  Small values
    are decided without the loop (src/annie.h:222-236), so the verdict names
    the value's primality (a wrong precedence gives 15, 17 or 3 in three of
    the cases, which flips the verdict)."""
    exe = os.path.join(annie_amd.LIBDIR, "annie")
    out = subprocess.run([exe, "-q", expr], capture_output=True, text=True, timeout=60, cwd=str(tmp_path))
    assert out.returncode == 0, out.stderr
    prime = value > 1 and all(value % d for d in range(2, int(value ** 0.5) + 1))
    assert f"{expr} is {'prime' if prime else 'composite'}." in out.stdout, out.stdout


@pytest.mark.parametrize("expr", ["2^-1", "-2^2", "2^(0-3)"])
def test_cli_rejects_non_positive_or_fractional(tmp_path, expr):
    exe = os.path.join(annie_amd.LIBDIR, "annie")
    out = subprocess.run([exe, "-q", expr], capture_output=True, text=True, timeout=60, cwd=str(tmp_path))
    assert out.returncode == 1 and "Error" in out.stderr



def test_bench_reads_committed_traffic_for_every_profiled_kernel():
    """bench.py takes the roofline `traffic` of the dominant kernel from the
    committed rocprofv3 PMC summary: every kernel of the 40M-digit squaring
    must have a parsable line there (HBM bytes at least its algorithmic bytes,
    a VALU utilisation in (0, 1])."""
    import bench
    U = 5 * 2 ** 21 * 4
    for kind in ("k_row<11,mulc>x1", "k_row<11,sqr>x2", "k_col_fwd<10>x1", "k_col_fwd<10>x2", "k_col_inv<10>x1",
                 "k_col_inv<10>x2", "k_carry<5>x1", "k_carry<5>x2"):
        c = bench.measured_counters(kind)
        assert c.get("traffic") is not None, kind
        assert c["traffic"] >= 0.95 * bench.alg_bytes(kind, U), kind
        assert c.get("valu_util") is not None and 0.0 < c["valu_util"] <= 1.0, kind


def test_schedule_fits_and_defaults():
    """The launch-schedule rules of annie_hip_schedule_fits (no GPU needed):
    powers of two only; row lines only for rows of fewer than 64 threads, at
    most the rows there are and 512 threads; columns only for column passes
    of up to 2^9, at most 1024 threads and 64 KiB of LDS; the defaults fit."""
    fits = annie_amd.schedule_fits
    assert fits(16, 4, 16) and fits(16, 32, 32) and fits(14, 64, 64)
    assert not fits(16, 3, 0) and not fits(16, 0, 24)      # not powers of two
    assert not fits(16, 64, 0)                             # 64 lines x 16 threads > 512
    assert not fits(14, 0, 128)                            # (128 + 1) x 2^7 words > 64 KiB
    assert not fits(13, 2, 0) and not fits(13, 0, 8)       # one-pass 2^13: no knobs
    assert not fits(21, 0, 8) and not fits(21, 1, 0)       # 2^10 columns, rows of 2^11
    assert fits(9, 1, 0) and not fits(9, 2, 0)             # one-pass 2^9: one row
    assert not fits(16, -1, 0)
    for LN in range(9, 23):
        d = annie_amd.schedule_default(LN)
        assert fits(LN, *d), LN
        sc = annie_amd.schedules(LN)
        assert sc[0] == (0, 0) and len(set(sc)) == len(sc) and d not in sc[1:]
        assert all(fits(LN, *x) for x in sc)
