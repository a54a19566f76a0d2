"""End-to-end tests on the GPU path: the native protocol host (restating
annie::check, src/annie.h:217-430) driving the gfx950 engine, against the
oracle's protocol on exact GMP arithmetic and mpz_powm; checkpoint / resume;
Gerbicz-Li failure detection; factor reporting; the 40M-digit size against a
committed GMP fixture."""
import hashlib
import json
import math
import os
import random

import pytest

import annie_amd
import oracle
from tests.helpers import proth, usable_N

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def gpu_check(N, **kw):
    # gmp_threshold_digits=1 forces the GPU transform at every size.
    kw.setdefault("gmp_threshold_digits", 1)
    kw.setdefault("min_ln", 5)
    return annie_amd.check(N, **kw)


@pytest.mark.parametrize("k,n", [(3, 34350), (3, 20909), (5, 30000), (1023, 25000)])
def test_proth_euler_matches_oracle(gpu, k, n):
    """Proth numbers k*2^n+1 (Euler test = Proth's theorem); 3*2^34350+1 and
    3*2^20909+1 are primes."""
    N = proth(k, n)
    r = gpu_check(N, is_euler=True)
    o = oracle.check(N, is_euler=True)
    print(f"case {k}*2^{n}+1: {r['message']}", flush=True)
    if o.get("init_failed"):
        assert r["status"] == "factor"
        return
    assert r["status"] == "success"
    assert r["gl_ok"]
    assert r["is_prp"] == o["is_prp"]
    assert r["res64"] == o["res64"]
    assert r["a"] == o["a"]


@pytest.mark.parametrize("bits,seed", [(12000, 1), (40000, 2), (70000, 3)])
def test_fermat_random_matches_powm(gpu, bits, seed):
    N = usable_N(bits, seed, 3)
    r = gpu_check(N, a=3)
    assert r["status"] == "success" and r["gl_ok"]
    assert r["res64"] == oracle.powm(3, N - 1, N) & (2 ** 64 - 1)
    assert not r["is_prp"]


def test_checkpoint_resume_gives_same_result(gpu, tmp_path):
    N = proth(3, 34350)
    full = gpu_check(N, is_euler=True)
    part = gpu_check(N, is_euler=True, workdir=str(tmp_path), quit_after=9000)
    assert part["status"] == "aborted"
    assert any(f.name.endswith(".chk") for f in tmp_path.iterdir())
    rest = gpu_check(N, is_euler=True, workdir=str(tmp_path))
    assert rest["resumed"]
    assert rest["iterations"] == full["iterations"] - 9000
    assert rest["res64"] == full["res64"] and rest["is_prp"] == full["is_prp"] and rest["gl_ok"]
    # the checkpoint files are removed at the end of the test (src/annie.h:428)
    assert not any(f.name.endswith(".chk") for f in tmp_path.iterdir())


def test_corrupted_checkpoint_falls_back_to_old(gpu, tmp_path):
    """A damaged current checkpoint is rejected by its CRC and the .old one is used."""
    N = proth(3, 34350)
    full = gpu_check(N, is_euler=True)
    assert gpu_check(N, is_euler=True, workdir=str(tmp_path), quit_after=4000)["status"] == "aborted"
    first = [f for f in tmp_path.iterdir() if f.name.endswith(".chk")][0]
    # resume, run a bit more and abort again: the first file becomes .old
    assert gpu_check(N, is_euler=True, workdir=str(tmp_path), quit_after=3000)["status"] == "aborted"
    data = bytearray(first.read_bytes())
    data[100] ^= 0xFF
    first.write_bytes(bytes(data))
    rest = gpu_check(N, is_euler=True, workdir=str(tmp_path))
    assert rest["resumed"]
    assert rest["iterations"] == full["iterations"] - 4000
    assert rest["res64"] == full["res64"] and rest["gl_ok"]


def test_gerbicz_li_detects_corruption(gpu):
    r = gpu_check(proth(3, 34350), is_euler=True, corrupt_at=5000)
    assert r["status"] == "failed"
    assert not r["gl_ok"]


def test_factor_reported_when_gcd_with_R(gpu):
    """gcd(N, R) > 1 (src/transformGPU.h:652-661): the transform reports the factor."""
    rng = random.Random(9)
    while True:
        N = 3 * (rng.getrandbits(60000) | 1)
        if N % 2:
            break
    K, pck, LN, lbase = annie_amd.select_plan(N, 3)
    R = (1 << ((pck * lbase) << LN)) - 1
    g = math.gcd(N, R)
    assert g > 1
    t = annie_amd.Transform(N, 3, min_ln=5)
    assert t.init_failed()
    assert t.get_factor() == g
    r = gpu_check(N, a=3)
    assert r["status"] == "factor"
    assert r["message"].startswith(str(g))


def test_gerbicz_li_and_checkpoint_at_40M_digits(gpu):
    """The benchmark's 132,900,000-bit N over the exponent made of the top
    10,000 bits of N - 1 (the timed depth): the hot loop with the Gerbicz-Li
    products every B = 79 iterations (src/annie.h:319-341), a checkpoint
    round trip of r0 and r1 through host bytes half-way (the device
    registers are overwritten in between), then d(t+1) = d(t) result and the
    verification a^(res - 2^B) d(t)^(2^B) a^(res mod 2^B) == d(t+1)
    (src/annie.h:343-404); the residues after 100 and 1,000 iterations
    against the GMP fixture (mpz_powm of the same prefixes)."""
    import bench
    fx = json.load(open(os.path.join(HERE, "golden", "cfg3_prefix.json")))
    N = bench.workload_N(skip=fx["candidate"])
    a = fx["base"]
    t = annie_amd.Transform(N, a)
    assert (t.K, t.LN) == (5, 21)
    L = 10_000
    e = (N - 1) >> ((N - 1).bit_length() - L)
    B = oracle.gl_block(L)
    assert B == 79
    prefix = {x["iterations"]: x for x in fx["entries"]}
    assert set(prefix) == {100, 1000}
    nbytes = (N.bit_length() + 7) // 8
    t.set1()
    t.copy(1, 0)
    seg = []
    for i in range(L - 1, -1, -1):  # the hot loop, squarings issued in segments
        seg.append((e >> i) & 1)
        done = L - i
        gl = i % B == 0 and i // B != 0
        if gl or done in (100, 1000, 5000):
            t.square_mul_bits(seg, a)
            seg = []
        if gl:
            t.mul(1, 0)
        if done in prefix:  # a^(top `done` bits of e) mod N from mpz_powm
            v = t.get(0)
            assert v & (2 ** 64 - 1) == prefix[done]["res64"], f"res64 after {done} iterations"
            assert hashlib.sha256(v.to_bytes(nbytes, "little")).hexdigest() == prefix[done]["sha256"]
        if done == 5000:  # checkpoint round trip half-way
            saved = t.save_checkpoint()
            t.set1()
            t.copy(1, 0)
            assert t.read_checkpoint(saved)
    t.square_mul_bits(seg, a)
    t.copy(2, 0)
    t.mul(2, 1)
    res, ee = 0, e
    while ee:
        res += ee & ((1 << B) - 1)
        ee >>= B
    rs = res.bit_length()
    t.set1()
    t.square_mul_bits([(res >> j) & 1 for j in range(rs - 1, B - 1, -1)], a)
    t.mul(0, 1)
    t.square_mul_bits([(res >> j) & 1 if j < rs else 0 for j in range(B - 1, -1, -1)], a)
    r1, r2 = t.get(0), t.get(2)
    assert r1 == r2, "Gerbicz-Li check failed"
    t.close()
