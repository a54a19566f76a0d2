// Host-side planning for one modulus N: transform size selection, twiddle
// tables and the per-prime constants the kernels take as arguments.
#pragma once
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "arith.h"

namespace annie {

struct Plan {
  int K = 0;      // number of RNS primes (2..5)
  int pck = 0;    // digits packed per coefficient
  int LN = 0;     // log2 of the transform length
  int lbase = 0;  // bits per packed digit
  size_t n() const { return size_t(1) << LN; }
  int cbits() const { return pck * lbase; }  // bits per coefficient, B = 2^cbits
};

// Number of digits packed per coefficient for K primes (src/transformGPU.h:816-820):
// 1, 2, 2, 4 for K = 2, 3, 4, 5.
static inline int pack_factor(int K) {
  int s = -1;
  for (int k = K - 1; k != 0; k /= 2) ++s;
  return 1 << s;
}

// Largest lbase (31 down to 2) whose smallest power-of-two length covers
// log2(16 a N) bits while keeping LN + 2*log2(B) below log2(P0...P_{K-1}) - log2(2a),
// i.e. a * n * B^2 < (P0...P_{K-1}) / 2 so that every convolution coefficient
// is recovered exactly by the CRT (src/transformGPU.h:798-814).
static inline void length_and_base(int& rln, int& rlbase, int pck, double log2_16aN, double log2_p_2a) {
  for (int lbase = 31; lbase >= 2; --lbase) {
    int ln = 4;
    while (double(lbase * pck) * double(1 << ln) < log2_16aN) ++ln;
    const int log2_base = lbase * pck;
    if (ln + 2 * log2_base < log2_p_2a) {
      rln = ln;
      rlbase = lbase;
      return;
    }
  }
}

// Size selection of src/transformGPU.h:822-850: for K = 2..5 pick (LN, lbase) and
// keep the K minimising K * n * LN (first minimum on ties).  `mant` and `exp2`
// are N = mant * 2^exp2 with mant in [0.5, 1), as mpz_get_d_2exp returns; the
// floating-point expressions are evaluated in the same order as the reference
// so that class boundaries agree bit for bit.
static inline Plan select_plan(double mant, long exp2, int32_t a) {
  const double log2_16aN = std::log2(16.0 * a) + std::log2(mant) + double(exp2);
  double log2_p_2a = std::log2(double(PRIMES[0])) - std::log2(2.0 * a);
  uint64_t c_min = ~uint64_t(0);
  int K_min = 0;
  for (int K = 2; K < 6; ++K) {
    log2_p_2a += std::log2(double(PRIMES[K - 1]));
    int ln = 0, lbase = 0;
    length_and_base(ln, lbase, pack_factor(K), log2_16aN, log2_p_2a);
    const uint64_t c = uint64_t(K) * (uint64_t(1) << ln) * uint64_t(ln);
    if (c < c_min) {
      c_min = c;
      K_min = K;
    }
  }
  Plan pl;
  pl.K = K_min;
  pl.pck = pack_factor(K_min);
  double lp = std::log2(double(PRIMES[0])) - std::log2(2.0 * a);
  for (int k = 1; k < K_min; ++k) lp += std::log2(double(PRIMES[k]));
  length_and_base(pl.LN, pl.lbase, pl.pck, log2_16aN, lp);
  return pl;
}

static inline uint32_t bitrev(uint32_t x, int bits) {
  uint32_t r = 0;
  for (int i = 0; i < bits; ++i) {
    r = (r << 1) | (x & 1);
    x >>= 1;
  }
  return r;
}

// A twiddle in Montgomery form with its precomputed w * p^-1 mod 2^32.
struct Tw {
  uint32_t w, wq;
};

// Twiddle tables of one prime for length n = 2^LN, heap-indexed: entry 2^l + b
// (l in [0, LN), b in [0, 2^l)) is the root r(l, b) that splits node b of level
// l, X^{2h} - r^2 = (X^h - r)(X^h + r) with h = n / 2^{l+1}:
//   cyclic (mod X^n - 1):     r(l, b) = w^{h * brv_l(b)},           w primitive n-th root
//   negacyclic (mod X^n + 1): r(l, b) = psi^{h * (2 brv_l(b) + 1)}, psi primitive 2n-th root
// These are the roots of src/transformGPU.h:116-160 (set_roots_R / set_roots_Q)
// arranged level by level for the forward CT and inverse GS butterflies.
// Forward tables hold r, inverse tables r^-1, all in Montgomery form so that
// mont_c(x, w, wq) = x * r mod p.
struct PrimeTables {
  std::vector<Tw> fwdR, invR, fwdQ, invQ;
};

static inline PrimeTables make_tables(int LN, int k) {
  const uint32_t p = PRIMES[k], q = inv32(p);
  const size_t n = size_t(1) << LN;
  const uint32_t psi = powmod_plain(PRIMROOTS[k], (uint64_t(p) - 1) >> (LN + 1), p);
  PrimeTables t;
  t.fwdR.resize(n);
  t.invR.resize(n);
  t.fwdQ.resize(n);
  t.invQ.resize(n);
  auto mk = [&](uint32_t r) {
    Tw e;
    e.w = to_mont(r, p);
    e.wq = e.w * q;
    return e;
  };
  // psi^e for e in [0, 2n): table of powers.
  std::vector<uint32_t> pw(2 * n);
  pw[0] = 1;
  for (size_t i = 1; i < 2 * n; ++i) pw[i] = mulmod_plain(pw[i - 1], psi, p);
  t.fwdR[0] = t.invR[0] = t.fwdQ[0] = t.invQ[0] = mk(1);
  for (int l = 0; l < LN; ++l) {
    const size_t h = n >> (l + 1);
    for (size_t b = 0; b < (size_t(1) << l); ++b) {
      const size_t br = bitrev(uint32_t(b), l);
      const size_t eR = (2 * h * br) % (2 * n);        // w = psi^2
      const size_t eQ = (h * (2 * br + 1)) % (2 * n);
      const size_t idx = (size_t(1) << l) + b;
      t.fwdR[idx] = mk(pw[eR]);
      t.invR[idx] = mk(pw[(2 * n - eR) % (2 * n)]);
      t.fwdQ[idx] = mk(pw[eQ]);
      t.invQ[idx] = mk(pw[(2 * n - eQ) % (2 * n)]);
    }
  }
  return t;
}

// Per-prime constants passed to the kernels (plain structs: kernel arguments).
struct PrimeConsts {
  uint32_t p[KMAX], q[KMAX];
  uint32_t sq_scale[KMAX], sq_scale_q[KMAX];  // n^-1 * 2^64: mont(mont(x,x), s) = x^2 / n
  uint32_t c32[KMAX], c32q[KMAX];             // 2^64 mod p: mont(d, c32) = d * 2^32
  uint32_t c64[KMAX], c64q[KMAX];             // 2^96 mod p: mont(d, c64) = d * 2^64
  // Garner: ginv[i][j] = p_i^-1 mod p_j (Montgomery form) for i > j.
  uint32_t ginv[KMAX][KMAX], ginvq[KMAX][KMAX];
};

static inline PrimeConsts make_consts(const Plan& pl) {
  PrimeConsts c{};
  const size_t n = pl.n();
  for (int k = 0; k < KMAX; ++k) {
    const uint32_t p = PRIMES[k], q = inv32(p);
    c.p[k] = p;
    c.q[k] = q;
    const uint32_t ninv = invmod_plain(uint32_t(n % p), p);
    c.sq_scale[k] = mulmod_plain(ninv, pow2mod(64, p), p);
    c.sq_scale_q[k] = c.sq_scale[k] * q;
    c.c32[k] = pow2mod(64, p);
    c.c32q[k] = c.c32[k] * q;
    c.c64[k] = pow2mod(96, p);
    c.c64q[k] = c.c64[k] * q;
    for (int j = 0; j < KMAX; ++j) {
      if (j == k) continue;
      const uint32_t pj = PRIMES[j];
      c.ginv[k][j] = to_mont(invmod_plain(p % pj, pj), pj);
      c.ginvq[k][j] = c.ginv[k][j] * inv32(pj);
    }
  }
  return c;
}

}  // namespace annie
