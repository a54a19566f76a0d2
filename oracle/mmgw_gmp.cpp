// ORACLE -- test infrastructure only.  Imported, linked or run only by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the checker.
//
// Exact Montgomery-McLaughlin-Gallot-Woltman (MMGW) arithmetic with GMP, a
// restatement of src/transformGMP.h, and the Fermat/Euler test protocol with
// the Gerbicz-Li proof check, a restatement of annie::check (src/annie.h:217-430).
//
// Pinning (see tests/test_oracle.py): get() after k squarings equals
// a^(2^k...) computed with mpz_powm; the protocol's res64 equals the low 64
// bits of a^(N-1) mod N from mpz_powm; verdicts match the QUICK_TEST Fibonacci
// list of src/main.cpp:163 and known primes; the MMGW range theorem
// 0 <= r < 4N (doc/annie.tex:151-164) is asserted on every reduction.
#include <gmp.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct Gmp {
  mp_bitcnt_t nb = 0;  // R = 2^nb - 1, Q = 2^nb + 1
  bool failed = false;
  bool range_ok = true;  // every reduction stayed in [0, 4N)
  mpz_t N, Ninv, R, Q, one, tq, tmp;
  mpz_t reg[3];

  // z mod R by folding the bits above nb (src/transformGMP.h:30-40); result in [0, R].
  void mod_R(mpz_t z) {
    while (mpz_sizeinbase(z, 2) > nb) {
      mpz_tdiv_q_2exp(tmp, z, nb);
      mpz_tdiv_r_2exp(z, z, nb);
      mpz_add(z, z, tmp);
    }
  }
  // z mod Q for z >= 0 (src/transformGMP.h:42-60): fold 2nb-bit chunks (2^2nb = 1
  // mod Q), then 2^nb = -1 gives low - high, lifted into [0, Q).
  void mod_Q(mpz_t z) {
    while (mpz_sizeinbase(z, 2) > 2 * nb) {
      mpz_tdiv_q_2exp(tmp, z, 2 * nb);
      mpz_tdiv_r_2exp(z, z, 2 * nb);
      mpz_add(z, z, tmp);
    }
    mpz_tdiv_q_2exp(tmp, z, nb);
    mpz_tdiv_r_2exp(z, z, nb);
    mpz_sub(z, z, tmp);
    if (mpz_sgn(z) < 0) mpz_add(z, z, Q);
  }
  // x <- ((x N^-1 mod R) + 2) N - T_Q mod Q  (src/transformGMP.h:95-115).
  void reduce(mpz_t x) {
    mpz_mul(x, x, Ninv);
    mod_R(x);
    mpz_add_ui(x, x, 2);
    mpz_mul(x, x, N);
    mod_Q(x);
    mpz_sub(x, x, tq);
    if (mpz_sgn(x) < 0) mpz_add(x, x, Q);
    // Theorem (doc/annie.tex:151-164): 0 <= r' < 4N.
    mpz_mul_2exp(tmp, N, 2);
    if (mpz_sgn(x) < 0 || mpz_cmp(x, tmp) >= 0) range_ok = false;
  }
  void square_mul(int32_t a) {  // src/transformGMP.h:124-129
    mpz_t& x = reg[0];
    mpz_mul_si(tq, x, a);
    mpz_mul(tq, tq, x);
    mod_Q(tq);
    mpz_mul(x, x, x);
    mpz_mul_si(x, x, a);
    mod_R(x);
    reduce(x);
  }
  void mul(int d, int s) {  // src/transformGMP.h:117-122
    mpz_mul(tq, reg[d], reg[s]);
    mod_Q(tq);
    mpz_mul(reg[d], reg[d], reg[s]);
    mod_R(reg[d]);
    reduce(reg[d]);
  }
  void get(mpz_t z, int s) {  // src/transformGMP.h:138-145
    mpz_set(tq, reg[s]);
    mpz_set(z, reg[s]);
    reduce(z);
    mpz_mod(z, z, N);
  }
};

void from_limbs(mpz_t z, const uint64_t* l, size_t n) { mpz_import(z, n, -1, 8, 0, 0, l); }
size_t to_limbs(const mpz_t z, uint64_t* out, size_t cap) {
  const size_t need = mpz_sgn(z) == 0 ? 0 : (mpz_sizeinbase(z, 2) + 63) / 64;
  if (need > cap) return size_t(-1);
  size_t cnt = 0;
  if (need) mpz_export(out, &cnt, -1, 8, 0, 0, z);
  return need;
}

// B of the Gerbicz-Li proof for an exponent of `size` bits (src/annie.h:97-101).
int gl_block(size_t size) {
  int lg = -1;
  for (uint32_t v = uint32_t(size); v; v >>= 1) ++lg;
  const size_t L = size_t(2) << (lg / 2);
  return int((size - 1) / L) + 1;
}

}  // namespace

extern "C" {

void* or_gmp_create(const uint64_t* Nl, size_t nl, int32_t a) {
  Gmp* g = new Gmp();
  mpz_inits(g->N, g->Ninv, g->R, g->Q, g->one, g->tq, g->tmp, g->reg[0], g->reg[1], g->reg[2], nullptr);
  from_limbs(g->N, Nl, nl);
  // 2^nb > 16 a N (src/transformGMP.h:176)
  int abits = 0;
  for (uint32_t v = uint32_t(a); v; v >>= 1) ++abits;
  g->nb = mpz_sizeinbase(g->N, 2) + size_t(abits) + 4;
  mpz_set_ui(g->R, 1);
  mpz_mul_2exp(g->R, g->R, g->nb);
  mpz_add_ui(g->Q, g->R, 1);
  mpz_sub_ui(g->R, g->R, 1);
  if (mpz_invert(g->Ninv, g->N, g->R) == 0) {
    mpz_gcd(g->Ninv, g->N, g->R);
    g->failed = true;
    return g;
  }
  // MMGW form of 1: R/2 mod N (src/transformGMP.h:185-189)
  mpz_mod(g->one, g->R, g->N);
  if (mpz_odd_p(g->one)) mpz_add(g->one, g->one, g->N);
  mpz_tdiv_q_2exp(g->one, g->one, 1);
  return g;
}

int or_gmp_failed(void* h) { return static_cast<Gmp*>(h)->failed ? 1 : 0; }
int or_gmp_range_ok(void* h) { return static_cast<Gmp*>(h)->range_ok ? 1 : 0; }
void or_gmp_set1(void* h) { Gmp* g = static_cast<Gmp*>(h); mpz_set(g->reg[0], g->one); }
void or_gmp_copy(void* h, int d, int s) { Gmp* g = static_cast<Gmp*>(h); mpz_set(g->reg[d], g->reg[s]); }
void or_gmp_square_mul(void* h, int32_t a) { static_cast<Gmp*>(h)->square_mul(a); }
void or_gmp_mul(void* h, int d, int s) { static_cast<Gmp*>(h)->mul(d, s); }

long or_gmp_get(void* h, int s, uint64_t* out, size_t cap) {
  Gmp* g = static_cast<Gmp*>(h);
  mpz_t z;
  mpz_init(z);
  g->get(z, s);
  const size_t r = to_limbs(z, out, cap);
  mpz_clear(z);
  return long(r);
}

// Raw register content (MMGW form, not reduced mod N).
long or_gmp_raw(void* h, int s, uint64_t* out, size_t cap) { return long(to_limbs(static_cast<Gmp*>(h)->reg[s], out, cap)); }

void or_gmp_destroy(void* h) {
  Gmp* g = static_cast<Gmp*>(h);
  mpz_clears(g->N, g->Ninv, g->R, g->Q, g->one, g->tq, g->tmp, g->reg[0], g->reg[1], g->reg[2], nullptr);
  delete g;
}

// b^e mod N with mpz_powm (the textbook oracle).
long or_powm(const uint64_t* b, size_t nb, const uint64_t* e, size_t ne, const uint64_t* N, size_t nn, uint64_t* out,
             size_t cap) {
  mpz_t B, E, M, z;
  mpz_inits(B, E, M, z, nullptr);
  from_limbs(B, b, nb);
  from_limbs(E, e, ne);
  from_limbs(M, N, nn);
  mpz_powm(z, B, E, M);
  const size_t r = to_limbs(z, out, cap);
  mpz_clears(B, E, M, z, nullptr);
  return long(r);
}

// gcd(a, b) with GMP's subquadratic algorithm (Python's is quadratic: tens
// of seconds at millions of bits); pinned against math.gcd on small inputs.
long or_gcd(const uint64_t* a, size_t na, const uint64_t* b, size_t nb, uint64_t* out, size_t cap) {
  mpz_t A, B, z;
  mpz_inits(A, B, z, nullptr);
  from_limbs(A, a, na);
  from_limbs(B, b, nb);
  mpz_gcd(z, A, B);
  const size_t r = to_limbs(z, out, cap);
  mpz_clears(A, B, z, nullptr);
  return long(r);
}

// Base of the test (src/annie.h:249-254): 2 for Fermat unless a_ref != 0; for
// Euler the smallest a >= 2, skipping squares, with Jacobi (a/N) = -1.
int32_t or_base(const uint64_t* Nl, size_t nl, int is_euler, int32_t a_ref) {
  mpz_t N;
  mpz_init(N);
  from_limbs(N, Nl, nl);
  int32_t a = 2;
  if (is_euler) {
    while (mpz_si_kronecker(a, N) != -1) {
      ++a;
      const int32_t s = int32_t(std::lround(std::sqrt(double(a))));
      if (s * s == a) ++a;
    }
  } else if (a_ref != 0) {
    a = a_ref;
  }
  mpz_clear(N);
  return a;
}

// The protocol of annie::check (src/annie.h:293-404) on the exact GMP
// transform.  Returns 0 on success with *is_prp, *res64, *gl_ok filled, or 1
// when gcd(N, R) > 1 (factor not reported).  stop_after >= 0 truncates the hot
// loop after that many iterations and returns the canonical value of r0 in
// `out` (for mid-test parity), with return value 2.
int or_check(const uint64_t* Nl, size_t nl, int is_euler, int32_t a, int* is_prp, uint64_t* res64, int* gl_ok,
             long stop_after, uint64_t* out, size_t cap, long* nout) {
  mpz_t N, e, t, res, r1, r2;
  mpz_inits(N, e, t, res, r1, r2, nullptr);
  from_limbs(N, Nl, nl);
  Gmp* g = static_cast<Gmp*>(or_gmp_create(Nl, nl, a));
  if (g->failed) {
    or_gmp_destroy(g);
    mpz_clears(N, e, t, res, r1, r2, nullptr);
    return 1;
  }
  mpz_sub_ui(e, N, 1);
  if (is_euler) mpz_tdiv_q_2exp(e, e, 1);
  const int B = gl_block(mpz_sizeinbase(e, 2));
  mpz_set(g->reg[0], g->one);
  mpz_set(g->reg[1], g->reg[0]);  // d(t)
  const long i0 = long(mpz_sizeinbase(e, 2)) - 1;
  long done = 0;
  for (long i = i0; i >= 0; --i) {
    g->square_mul(mpz_tstbit(e, mp_bitcnt_t(i)) ? a : 1);
    if ((i % B == 0) && (i / B != 0)) g->mul(1, 0);
    if (stop_after >= 0 && ++done == stop_after) {
      g->get(t, 0);
      const size_t r = to_limbs(t, out, cap);
      if (nout) *nout = long(r);
      or_gmp_destroy(g);
      mpz_clears(N, e, t, res, r1, r2, nullptr);
      return 2;
    }
  }
  mpz_set(g->reg[2], g->reg[0]);  // d(t+1) = d(t) * result
  g->mul(2, 1);
  g->get(t, 0);
  if (is_euler) {
    mpz_add_ui(t, t, 1);
    if (mpz_cmp(t, N) >= 0) mpz_sub(t, t, N);
    g->square_mul(1);  // res64 is taken from a^(N-1) (src/annie.h:349-356)
  } else {
    mpz_sub_ui(t, t, 1);
  }
  *is_prp = mpz_sgn(t) == 0 ? 1 : 0;
  g->get(t, 0);
  *res64 = mpz_size(t) ? uint64_t(mpz_getlimbn(t, 0)) : 0;
  // res = sum of the B-bit chunks of e (src/annie.h:368-375)
  mpz_set_ui(res, 0);
  while (mpz_sgn(e) != 0) {
    mpz_tdiv_r_2exp(t, e, mp_bitcnt_t(B));
    mpz_add(res, res, t);
    mpz_tdiv_q_2exp(e, e, mp_bitcnt_t(B));
  }
  const long rs = long(mpz_sizeinbase(res, 2));
  mpz_set(g->reg[0], g->one);
  for (long i = rs - 1; i >= B; --i) g->square_mul(mpz_tstbit(res, mp_bitcnt_t(i)) ? a : 1);
  g->mul(0, 1);
  for (long i = B - 1; i >= 0; --i) g->square_mul((i < rs && mpz_tstbit(res, mp_bitcnt_t(i))) ? a : 1);
  g->get(r1, 0);
  g->get(r2, 2);
  *gl_ok = (mpz_cmp(r1, r2) == 0) ? 1 : 0;
  or_gmp_destroy(g);
  mpz_clears(N, e, t, res, r1, r2, nullptr);
  return 0;
}

int or_gl_block(size_t size) { return gl_block(size); }

}  // extern "C"
