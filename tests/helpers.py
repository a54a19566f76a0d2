"""Shared input generators and comparison helpers for the tests."""
import math
import random

import numpy as np

import oracle


def usable_N(bits, seed, a=3, min_ln=4):
    """Seeded random odd `bits`-bit N (top bit set) whose transform initialises.

    gcd(N, R) > 1 makes the reference report a factor instead of testing
    (src/transformGPU.h:652-661); such N are skipped here and tested apart.
    """
    rng = random.Random(seed)
    while True:
        N = rng.getrandbits(bits) | (1 << (bits - 1)) | 1
        K, pck, LN, lbase = oracle.plan(N, a)
        if LN < min_ln:
            raise ValueError("size below the engine's minimum length")
        R = (1 << (pck * lbase << LN)) - 1
        if oracle.gcd(N, R) == 1:
            return N


def proth(k, n):
    return k * (1 << n) + 1


def bits_of(e):
    return [(e >> i) & 1 for i in range(e.bit_length() - 1, -1, -1)]


def assert_same_words(got, want, what, n=None):
    """Word-exact comparison of two uint32 buffers with a compact report."""
    g = np.ctypeslib.as_array(got) if not isinstance(got, np.ndarray) else got
    w = np.ctypeslib.as_array(want) if not isinstance(want, np.ndarray) else want
    assert g.shape == w.shape, f"{what}: shape {g.shape} != {w.shape}"
    bad = np.nonzero(g != w)[0]
    if bad.size:
        i = int(bad[0])
        where = f"plane {i // n} index {i % n}" if n else f"index {i}"
        raise AssertionError(f"{what}: {bad.size} of {g.size} words differ; first at {where}: "
                             f"got {int(g[i]):#x} want {int(w[i]):#x}")
