"""GPU parity: the HIP path against the CPU oracle, bit for bit.

Registers are compared residue by residue with the CPU RNS-NTT oracle (which
mirrors the device representation), and canonical values against mpz_powm /
Python pow and the exact GMP MMGW transform (restating src/transformGMP.h).
"""
import random

import pytest

import annie_amd
import oracle
from tests.helpers import assert_same_words, usable_N

pytestmark = pytest.mark.gpu

# (bits, seed): single-pass lengths 2^5..2^13 and two-pass lengths 2^14..2^17.
SIZES = [(2000, 1), (5000, 2), (20000, 3), (60000, 4), (150000, 5), (300000, 6), (500000, 7),
         (1000000, 8), (2000000, 9), (4000000, 10), (332000, 12), (3000020, 11)]
# Plans the reference selects at the sizes of the configurations, asserted
# where they are the point of the case: the 332,000-bit Fermat test and the
# Proth k*2^3,000,000+1 test run on K = 4 (one-pass 2^13, two-pass 2^16).
PLAN_AT = {332000: (4, 2, 13, 28), 3000020: (4, 2, 16, 27)}


def _flush(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("bits,seed", SIZES)
def test_square_mul_registers_match_oracle(gpu, bits, seed):
    a = 3
    N = usable_N(bits, seed, a)
    _flush(f"case bits={bits} seed={seed}")
    t = annie_amd.Transform(N, a, min_ln=5)
    o = oracle.NttTransform(N, a)
    assert (t.K, t.pck, t.LN, t.lbase) == (o.K, o.pck, o.LN, o.lbase)
    if bits in PLAN_AT:
        assert (t.K, t.pck, t.LN, t.lbase) == PLAN_AT[bits]
    t.set1()
    o.set1()
    assert_same_words(t.read_register(0), o.read_register(0), "set1", t.n)
    rng = random.Random(seed)
    e = 1
    for step in range(3):
        bit = rng.getrandbits(1) if step else 1
        t.square_mul(a if bit else 1)
        o.square_mul(a if bit else 1)
        e = 2 * e + bit if step else 1
        assert_same_words(t.read_register(4), o.read_register(4), f"m'+2 after step {step}", t.n)
        assert_same_words(t.read_register(0), o.read_register(0), f"r0 after step {step}", t.n)
    assert t.get(0) == pow(a, e, N)


@pytest.mark.parametrize("bits,seed", SIZES[:6])
def test_mul_and_get(gpu, bits, seed):
    a = 3
    N = usable_N(bits, seed, a)
    t = annie_amd.Transform(N, a, min_ln=5)
    o = oracle.NttTransform(N, a)
    for m in (t, o):
        m.set1()
        m.square_mul(a)
        m.square_mul(a)      # 3^3
        m.copy(1, 0)
        m.square_mul(1)      # 3^6
        m.mul(0, 1)          # 3^9
        m.copy(2, 0)
        m.mul(2, 2)          # 3^18
    assert_same_words(t.read_register(0), o.read_register(0), "r0", t.n)
    assert_same_words(t.read_register(2), o.read_register(2), "r2", t.n)
    assert t.get(0) == pow(3, 9, N)
    assert t.get(1) == pow(3, 3, N)
    assert t.get(2) == pow(3, 18, N)
