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


def test_prefix_at_40M_digits_matches_gmp_fixture(gpu):
    """The benchmark's 132,900,000-bit N: residue after the first t loop
    iterations == 3^(top t bits of N-1) mod N from mpz_powm (committed fixture)."""
    import bench
    path = os.path.join(HERE, "golden", "cfg3_prefix.json")
    fx = json.load(open(path))
    N = bench.workload_N(skip=fx["candidate"])
    t = annie_amd.Transform(N, fx["base"])
    assert not t.init_failed()
    assert (t.K, t.LN, t.lbase) == (5, 21, 16)
    e = N - 1
    done = 0
    t.set1()
    for ent in sorted(fx["entries"], key=lambda d: d["iterations"]):
        it = ent["iterations"]
        bits = bench.exponent_bits(N, it)[done:]
        t.square_mul_bits(bits, fx["base"])
        done = it
        v = t.get(0)
        assert v & (2 ** 64 - 1) == ent["res64"], f"res64 after {it} iterations"
        nbytes = (N.bit_length() + 7) // 8
        assert hashlib.sha256(v.to_bytes(nbytes, "little")).hexdigest() == ent["sha256"]
    t.close()
