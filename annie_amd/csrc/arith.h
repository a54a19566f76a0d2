// Modular arithmetic shared by the host planner and the gfx950 kernels.
//
// Residues live in Z/pZ for the five 32-bit primes of the RNS (the first five
// of the prime table at src/transformGPU.h:79-83).  Products use Montgomery
// reduction with radix 2^32 in the "signed" form: for a, b < 2^32 and p odd
// with p > 2^31, mont(a, b) = a*b*2^-32 mod p is hi(a*b) - hi(lo(a*b*q)*p) with
// q = p^-1 mod 2^32, a value in (-p, p) that one conditional add normalises.
// Multiplying by a fixed w uses the precomputed wq = w*q mod 2^32, so each
// product costs three 32-bit multiplies (mul_hi, mul_lo, mul_hi).  This is
// the same Montgomery variant as src/transformGPU.h:51-59 and
// ocl/kernel.cl:132-149; it is the textbook signed REDC and is restated here.
//
// The header compiles under g++ (host, oracle tests) and hipcc (device).
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define ANNIE_HD __host__ __device__ __forceinline__
#else
#define ANNIE_HD static inline
#endif

namespace annie {

// Largest RNS size the engine supports (the reference selects K in 2..5,
// src/transformGPU.h:830).
constexpr int KMAX = 5;

// 2^32 > P0 > P1 > ... > P4 > 2^31, p_k = c_k * 2^s_k + 1 with s_k >= 24, so
// every prime has 2^24-th roots of unity (transform lengths up to 2^23 need
// primitive 2n-th roots for the negacyclic transform).
constexpr uint32_t PRIMES[KMAX] = {
    (125u << 25) + 1, (243u << 24) + 1, (235u << 24) + 1, (29u << 27) + 1, (13u << 28) + 1};
// Primitive roots modulo each prime (src/transformGPU.h:89; checked by the tests
// against the factorisation of p - 1).
constexpr uint32_t PRIMROOTS[KMAX] = {3, 7, 3, 3, 3};

#if defined(__HIP_DEVICE_COMPILE__)
ANNIE_HD uint32_t mulhi32(uint32_t a, uint32_t b) { return __umulhi(a, b); }
#else
ANNIE_HD uint32_t mulhi32(uint32_t a, uint32_t b) { return uint32_t((uint64_t(a) * b) >> 32); }
#endif

// p^-1 mod 2^32 by Newton iteration (each step doubles the correct low bits).
ANNIE_HD uint32_t inv32(uint32_t p) {
  uint32_t q = p;  // correct to 3 bits for odd p
  for (int i = 0; i < 5; ++i) q *= 2u - p * q;
  return q;
}

// The corrections below are written with the carry / borrow builtins so that
// the device code is a subtract-with-borrow followed by a select of p or 0.

// a - b mod p for a, b in [0, p).
ANNIE_HD uint32_t submod(uint32_t a, uint32_t b, uint32_t p) {
  uint32_t d;
  const bool borrow = __builtin_sub_overflow(a, b, &d);
  return d + (borrow ? p : 0u);
}

// a + b mod p for a, b in [0, p), p < 2^32: a - (p - b) mod p.
ANNIE_HD uint32_t addmod(uint32_t a, uint32_t b, uint32_t p) { return submod(a, p - b, p); }

// Montgomery product by a constant: a * w * 2^-32 mod p, with wq = w * p^-1 mod 2^32.
// Valid for any a < 2^32 and w < p.
ANNIE_HD uint32_t mont_c(uint32_t a, uint32_t w, uint32_t wq, uint32_t p) {
  const uint32_t h = mulhi32(a, w);
  const uint32_t m = a * wq;
  const uint32_t g = mulhi32(m, p);
  return submod(h, g, p);
}

// General Montgomery product a * b * 2^-32 mod p (q = p^-1 mod 2^32).
ANNIE_HD uint32_t mont(uint32_t a, uint32_t b, uint32_t p, uint32_t q) { return mont_c(a, b, b * q, p); }

// (a - b) * w * 2^-32 mod p: one Garner step (src/transformGPU.h:107, ocl/kernel.cl:146-149).
ANNIE_HD uint32_t submul_c(uint32_t a, uint32_t b, uint32_t w, uint32_t wq, uint32_t p) {
  return mont_c(submod(a, b, p), w, wq, p);
}

// Constants fixed by the five primes, evaluated at compile time so that the
// CRT / carry kernels (whose prime indices are compile-time after unrolling)
// take them as instruction literals instead of kernel-argument registers:
// q = p^-1 mod 2^32; c32 = 2^64 mod p and c64 = 2^96 mod p (mont by them
// multiplies by 2^32 and 2^64); c32n = p - c32; the products of each with q;
// ginv[i][j] = p_i^-1 mod p_j in Montgomery form (i != j) and its product
// with p_j^-1 mod 2^32.  The same values as the host's make_consts.
struct PrimeLiterals {
  uint32_t q[KMAX], c32[KMAX], c32q[KMAX], c64[KMAX], c64q[KMAX], c32n[KMAX], c32nq[KMAX];
  uint32_t ginv[KMAX][KMAX], ginvq[KMAX][KMAX];
};

constexpr uint32_t cx_powmod(uint32_t b, uint64_t e, uint32_t p) {
  uint64_t r = 1 % p, x = b % p;
  while (e) {
    if (e & 1) r = (r * x) % p;
    x = (x * x) % p;
    e >>= 1;
  }
  return uint32_t(r);
}

constexpr uint32_t cx_inv32(uint32_t p) {
  uint32_t q = p;
  for (int i = 0; i < 5; ++i) q *= 2u - p * q;
  return q;
}

constexpr PrimeLiterals make_prime_literals() {
  PrimeLiterals c{};
  for (int k = 0; k < KMAX; ++k) {
    const uint32_t p = PRIMES[k], q = cx_inv32(p);
    c.q[k] = q;
    c.c32[k] = cx_powmod(2, 64, p);
    c.c32q[k] = c.c32[k] * q;
    c.c64[k] = cx_powmod(2, 96, p);
    c.c64q[k] = c.c64[k] * q;
    c.c32n[k] = p - c.c32[k];
    c.c32nq[k] = c.c32n[k] * q;
    for (int j = 0; j < KMAX; ++j) {
      if (j == k) continue;
      const uint32_t pj = PRIMES[j];
      const uint32_t inv = cx_powmod(p % pj, uint64_t(pj) - 2, pj);
      c.ginv[k][j] = uint32_t((uint64_t(inv) << 32) % pj);
      c.ginvq[k][j] = c.ginv[k][j] * cx_inv32(pj);
    }
  }
  return c;
}

constexpr PrimeLiterals PLIT = make_prime_literals();

// Host-side helpers (plain integer arithmetic; used for tables and constants).
static inline uint32_t powmod_plain(uint32_t b, uint64_t e, uint32_t p) {
  uint64_t r = 1 % p, x = b % p;
  while (e) {
    if (e & 1) r = (r * x) % p;
    x = (x * x) % p;
    e >>= 1;
  }
  return uint32_t(r);
}
static inline uint32_t invmod_plain(uint32_t a, uint32_t p) { return powmod_plain(a, uint64_t(p) - 2, p); }
static inline uint32_t mulmod_plain(uint32_t a, uint32_t b, uint32_t p) { return uint32_t((uint64_t(a) * b) % p); }
// Montgomery form of x: x * 2^32 mod p.
static inline uint32_t to_mont(uint32_t x, uint32_t p) { return uint32_t(((uint64_t(x % p)) << 32) % p); }
// 2^e mod p.
static inline uint32_t pow2mod(unsigned e, uint32_t p) { return powmod_plain(2, e, p); }

}  // namespace annie
