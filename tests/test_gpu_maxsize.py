"""The largest size the reference claims (~48 million digits, README.md:26):
a 159,000,000-bit N selects LN = 22 (n = 2^22); and the lengths 2^20 for
K = 3, 4 and 5 (36, 42 and 60 million bits), whose column passes run 2^10
elements per column as the 40M-digit size does.  Exact known answers while
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


@pytest.mark.parametrize("bits,seed,K,LN", [(36_000_000, 5, 3, 20), (42_000_000, 6, 4, 20), (60_000_000, 8, 5, 20),
                                           (159_000_000, 7, 4, 22)])
def test_largest_reference_size(gpu, bits, seed, K, LN):
    N = _candidate(bits, seed)
    print(f"case {bits}-bit N, seed {seed}", flush=True)
    t = annie_amd.Transform(N, 3)
    try:
        assert not t.init_failed()
        assert (t.K, t.LN) == (K, LN)
        t.set1()
        ebits = [1, 0, 1, 1, 0, 0, 1, 0, 1, 1, 1, 0, 1, 0, 0, 1, 1, 0, 1, 1]
        t.square_mul_bits(ebits, 3)
        e = int("".join(map(str, ebits)), 2)
        assert t.get(0) == 3 ** e  # < 2^1.6M bits: no reduction mod N
        # wrap past N: 8 more squarings with a = 3 (3^e grows to ~3e8 bits > N)
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
