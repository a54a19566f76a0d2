"""The largest size the reference claims (~48 million digits, README.md:26):
a 159,000,000-bit N selects LN = 22 (n = 2^22).  Exact known answers while
3^e < N (no reduction: the value is the integer 3^e), the canonical get()
after the values wrap past N against one GMP modular squaring, and square_mul
against mul on the same register."""
import math
import random

import pytest

import annie_amd
import oracle

pytestmark = pytest.mark.gpu


def _candidate(bits, seed):
    """Seeded odd `bits`-bit N with no factor in common with 2^d - 1 for the
    small divisors d of the transform's exponent cbits * n (so that the
    transform initialises; the engine re-checks gcd(N, R) exactly)."""
    K, pck, LN, lbase = oracle.plan((1 << bits) - 1, 3)
    E = pck * lbase << LN
    ds = [d for d in range(1, 4097) if E % d == 0]
    rng = random.Random(seed)
    while True:
        N = rng.getrandbits(bits) | (1 << (bits - 1)) | 1
        if all(math.gcd(N % ((1 << d) - 1), (1 << d) - 1) == 1 for d in ds):
            return N


def test_largest_reference_size(gpu):
    N = _candidate(159_000_000, 7)
    print("case 159,000,000-bit N, seed 7", flush=True)
    t = annie_amd.Transform(N, 3)
    try:
        assert not t.init_failed()
        assert t.LN == 22
        t.set1()
        bits = [1, 0, 1, 1, 0, 0, 1, 0, 1, 1, 1, 0, 1, 0, 0, 1, 1, 0, 1, 1]
        t.square_mul_bits(bits, 3)
        e = int("".join(map(str, bits)), 2)
        assert t.get(0) == 3 ** e  # < 2^1.6M bits: no reduction mod N
        # wrap past N: 8 more squarings with a = 3 (3^e grows to ~4.2e8 bits > N)
        t.square_mul_bits([1] * 8, 3)
        x = t.get(0)
        assert 0 < x < N
        # one more squaring against GMP
        t.copy(1, 0)
        t.square_mul(1)
        want = oracle.powm(x, 2, N)
        assert t.get(0) == want
        # mul(1, 1) = x^2 as well
        t.mul(1, 1)
        assert t.get(1) == want
    finally:
        t.close()
