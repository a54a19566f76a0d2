// ORACLE -- test infrastructure only.  Imported, linked or run only by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the checker.
//
// CPU restatement of the reference's RNS-NTT MMGW transform
// (src/transformGPU.h:286-796, src/transformCPU.h:553-1618, ocl/kernel.cl),
// written with plain 64-bit modular arithmetic and GMP for the CRT so that it
// shares no code with the product.  It follows the product's representation
// choices (transform order, lazy carry blocks of 4, canonical reduction rule)
// so that every intermediate register can be compared bit for bit:
//   * size selection            src/transformGPU.h:798-857
//   * roots of X^n -+ 1         src/transformGPU.h:116-160
//   * RNS encode of digits      src/transformGPU.h:323-360
//   * convolution               ocl/kernel.cl:206-284 (CT forward / GS inverse)
//   * CRT + carry + re-encode   ocl/kernel.cl:1418-1613
//   * mod R + 2 / mod Q         ocl/kernel.cl:1653-1683
//   * MMGW square / mul / get   src/transformGPU.h:523-586
//
// Pinning (tests/test_oracle.py): the transform against an O(n^2) brute-force
// cyclic/negacyclic convolution; the CRT against exact integers; the carry
// against the value identity sum d_i B^i = sum v_i B^i mod B^n -+ 1; get()
// against the GMP MMGW oracle and mpz_powm; the plan against the class table
// derived from the reference's selection code.
#include <gmp.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

typedef __int128 i128;

const uint32_t P[5] = {4194304001u, 4076863489u, 3942645761u, 3892314113u, 3489660929u};
const uint32_t G[5] = {3, 7, 3, 3, 3};

uint32_t pw(uint64_t b, uint64_t e, uint32_t p) {
  uint64_t r = 1;
  b %= p;
  while (e) {
    if (e & 1) r = r * b % p;
    b = b * b % p;
    e >>= 1;
  }
  return uint32_t(r);
}

void plan(double mant, long ex, int a, int& K, int& pck, int& LN, int& lbase) {
  // Restated from src/transformGPU.h:798-850 (same floating-point order).
  auto gets = [](int k) { int s = -1; for (int q = k - 1; q != 0; q /= 2) ++s; return 1 << s; };
  auto nb = [](int& rln, int& rlb, int pk, double l16, double lp) {
    for (int lb = 31; lb >= 2; --lb) {
      int ln = 4;
      while (double(lb * pk) * (1 << ln) < l16) ++ln;
      if (ln + 2 * (lb * pk) < lp) { rln = ln; rlb = lb; return; }
    }
  };
  const double l16 = std::log2(16.0 * a) + std::log2(mant) + double(ex);
  double lp = std::log2(double(P[0])) - std::log2(2.0 * a);
  uint64_t cmin = ~0ull;
  int kmin = 0;
  for (int k = 2; k < 6; ++k) {
    lp += std::log2(double(P[k - 1]));
    int ln = 0, lb = 0;
    nb(ln, lb, gets(k), l16, lp);
    const uint64_t c = uint64_t(k) * (1ull << ln) * uint64_t(ln);
    if (c < cmin) { cmin = c; kmin = k; }
  }
  K = kmin;
  pck = gets(K);
  double lq = std::log2(double(P[0])) - std::log2(2.0 * a);
  for (int k = 1; k < K; ++k) lq += std::log2(double(P[k]));
  nb(LN, lbase, pck, l16, lq);
}

uint32_t brv(uint32_t x, int bits) {
  uint32_t r = 0;
  for (int i = 0; i < bits; ++i) { r = (r << 1) | (x & 1); x >>= 1; }
  return r;
}

struct Oracle {
  int K, pck, LN, lbase, cbits;
  size_t n;
  bool failed = false;
  mpz_t N, M, Mh, factor;
  std::vector<uint32_t> tw[4][5];  // fwdR, invR, fwdQ, invQ (plain residues, heap-indexed)
  std::vector<uint32_t> NQ[5], NIR[5];  // forward transforms (unscaled)
  std::vector<i128> reg[3], one, TR, TQ, MP;  // MP: m' + 2 of the last reduction
  std::vector<mpz_t*> crt_coef;

  uint32_t ninv(int k) const { return pw(uint32_t(n % P[k]), P[k] - 2, P[k]); }

  void tables() {
    for (int k = 0; k < K; ++k) {
      const uint32_t p = P[k];
      const uint32_t psi = pw(G[k], (uint64_t(p) - 1) >> (LN + 1), p);
      std::vector<uint32_t> pp(2 * n);
      pp[0] = 1;
      for (size_t i = 1; i < 2 * n; ++i) pp[i] = uint32_t(uint64_t(pp[i - 1]) * psi % p);
      for (int t = 0; t < 4; ++t) tw[t][k].assign(n, 1);
      for (int l = 0; l < LN; ++l) {
        const size_t h = n >> (l + 1);
        for (size_t b = 0; b < (size_t(1) << l); ++b) {
          const size_t r = brv(uint32_t(b), l);
          const size_t eR = (2 * h * r) % (2 * n), eQ = (h * (2 * r + 1)) % (2 * n);
          const size_t ix = (size_t(1) << l) + b;
          tw[0][k][ix] = pp[eR];
          tw[1][k][ix] = pp[(2 * n - eR) % (2 * n)];
          tw[2][k][ix] = pp[eQ];
          tw[3][k][ix] = pp[(2 * n - eQ) % (2 * n)];
        }
      }
    }
  }

  void fwd(std::vector<uint32_t>& x, int k, bool nega) const {
    const uint32_t p = P[k];
    const std::vector<uint32_t>& T = tw[nega ? 2 : 0][k];
    for (int l = 0; l < LN; ++l) {
      const size_t h = n >> (l + 1);
      for (size_t b = 0; b < (size_t(1) << l); ++b) {
        const uint64_t r = T[(size_t(1) << l) + b];
        for (size_t j = b * 2 * h; j < b * 2 * h + h; ++j) {
          const uint32_t t = uint32_t(x[j + h] * r % p);
          x[j + h] = uint32_t((uint64_t(x[j]) + p - t) % p);
          x[j] = uint32_t((uint64_t(x[j]) + t) % p);
        }
      }
    }
  }

  void inv(std::vector<uint32_t>& x, int k, bool nega) const {
    const uint32_t p = P[k];
    const std::vector<uint32_t>& T = tw[nega ? 3 : 1][k];
    for (int l = LN - 1; l >= 0; --l) {
      const size_t h = n >> (l + 1);
      for (size_t b = 0; b < (size_t(1) << l); ++b) {
        const uint64_t r = T[(size_t(1) << l) + b];
        for (size_t j = b * 2 * h; j < b * 2 * h + h; ++j) {
          const uint32_t u = x[j], v = x[j + h];
          x[j] = uint32_t((uint64_t(u) + v) % p);
          x[j + h] = uint32_t((uint64_t(u) + p - v) % p * r % p);
        }
      }
    }
  }

  uint32_t res_of(i128 d, int k) const {
    const uint32_t p = P[k];
    i128 r = d % i128(p);
    if (r < 0) r += p;
    return uint32_t(r);
  }

  // Coefficient values mod p_k of the convolution of digit vectors x and y
  // (y == nullptr: y = x; yhat given: y's forward transform).
  void conv(const std::vector<i128>& x, const std::vector<i128>* y, const std::vector<uint32_t>* yhat, bool nega,
            std::vector<std::vector<uint32_t>>& out) const {
    out.assign(K, std::vector<uint32_t>(n));
    for (int k = 0; k < K; ++k) {
      const uint32_t p = P[k];
      std::vector<uint32_t> a(n), b;
      for (size_t i = 0; i < n; ++i) a[i] = res_of(x[i], k);
      fwd(a, k, nega);
      if (yhat) {
        b = yhat[k];
      } else if (y) {
        b.resize(n);
        for (size_t i = 0; i < n; ++i) b[i] = res_of((*y)[i], k);
        fwd(b, k, nega);
      } else {
        b = a;
      }
      const uint64_t ni = ninv(k);
      for (size_t i = 0; i < n; ++i) a[i] = uint32_t(uint64_t(a[i]) * b[i] % p * ni % p);
      inv(a, k, nega);
      out[k] = a;
    }
  }

  // Signed value in (-M/2, M/2] from residues (explicit CRT with GMP).
  void crt(const std::vector<std::vector<uint32_t>>& r, size_t i, mpz_t v) const {
    mpz_set_ui(v, 0);
    mpz_t t;
    mpz_init(t);
    for (int k = 0; k < K; ++k) {
      mpz_mul_ui(t, *crt_coef[k], r[k][i]);
      mpz_add(v, v, t);
    }
    mpz_mod(v, v, M);
    if (mpz_cmp(v, Mh) > 0) mpz_sub(v, v, M);
    mpz_clear(t);
  }

  // Balanced digit of f: d = f mod B taken in [-B/2, B/2), f <- (f - d) / B.
  void split(mpz_t f, mpz_t d) const {
    mpz_fdiv_r_2exp(d, f, cbits);
    if (mpz_tstbit(d, cbits - 1)) {
      mpz_t B;
      mpz_init_set_ui(B, 1);
      mpz_mul_2exp(B, B, cbits);
      mpz_sub(d, d, B);
      mpz_clear(B);
    }
    mpz_sub(f, f, d);
    mpz_fdiv_q_2exp(f, f, cbits);
  }

  // CRT values -> lazy balanced digits.  Blocks of 4 coefficients: phase 1
  // splits each block with zero carry-in, phase 2 adds the previous block's
  // carry-out (block 0: the last block's, negated mod Q) to the first three
  // digits and leaves the rest in the fourth.  Every digit ends in
  // [-B/2 - 1, B/2), so the number lies in (-B^n/2, B^n/2).
  // canon 1: reduce into [0, R) then add 2; canon 2: reduce into [0, Q).
  // The sign is that of the most significant nonzero digit (the scan of
  // src/transformCPU.h:1261-1281).
  std::vector<i128> carry(const std::vector<std::vector<uint32_t>>& r, uint32_t a, bool nega, int canon,
                          const std::vector<i128>* sub) const {
    const size_t nb = n / 4;
    std::vector<mpz_t> fo(nb);
    std::vector<i128> rd(n), d(n);
    mpz_t v, f, s;
    mpz_inits(v, f, s, nullptr);
    for (size_t b = 0; b < nb; ++b) {
      mpz_set_ui(f, 0);
      for (int j = 0; j < 4; ++j) {
        const size_t i = 4 * b + j;
        crt(r, i, v);
        mpz_mul_ui(v, v, a);
        if (sub) {
          set_i128(s, (*sub)[i]);
          mpz_sub(v, v, s);
        }
        mpz_add(f, f, v);
        split(f, v);
        rd[i] = get_i128(v);
      }
      mpz_init_set(fo[b], f);
    }
    for (size_t b = 0; b < nb; ++b) {
      if (b > 0) mpz_set(f, fo[b - 1]);
      else if (nega) mpz_neg(f, fo[nb - 1]);
      else mpz_set(f, fo[nb - 1]);
      for (int j = 0; j < 3; ++j) {
        set_i128(v, rd[4 * b + j]);
        mpz_add(f, f, v);
        split(f, v);
        d[4 * b + j] = get_i128(v);
      }
      set_i128(v, rd[4 * b + 3]);
      mpz_add(f, f, v);
      d[4 * b + 3] = get_i128(f);
    }
    for (size_t b = 0; b < nb; ++b) mpz_clear(fo[b]);
    mpz_clears(v, f, s, nullptr);
    if (canon) {
      int sign = 0;
      for (size_t i = n; i-- > 0;)
        if (d[i] != 0) {
          sign = d[i] < 0 ? -1 : 1;
          break;
        }
      if (sign < 0) {  // + R = B^n - 1, or + Q = B^n + 1
        d[n - 1] += i128(1) << cbits;
        d[0] += (canon == 1) ? -1 : 1;
      }
      if (canon == 1) d[0] += 2;
    }
    return d;
  }

  static void set_i128(mpz_t z, i128 x) {
    const bool neg = x < 0;
    unsigned __int128 u = neg ? (unsigned __int128)(-x) : (unsigned __int128)x;
    const uint64_t w[2] = {uint64_t(u), uint64_t(u >> 64)};
    mpz_import(z, 2, -1, 8, 0, 0, w);
    if (neg) mpz_neg(z, z);
  }
  static i128 get_i128(const mpz_t z) {
    uint64_t w[3] = {0, 0, 0};
    size_t cnt = 0;
    mpz_export(w, &cnt, -1, 8, 0, 0, z);
    unsigned __int128 u = (unsigned __int128)w[0] | ((unsigned __int128)w[1] << 64);
    return mpz_sgn(z) < 0 ? -(i128)u : (i128)u;
  }

  std::vector<i128> digits_of(const mpz_t x) const {
    std::vector<i128> d(n);
    mpz_t t;
    mpz_init(t);
    for (size_t i = 0; i < n; ++i) {
      mpz_tdiv_q_2exp(t, x, mp_bitcnt_t(i) * cbits);
      mpz_tdiv_r_2exp(t, t, cbits);
      d[i] = get_i128(t);
    }
    mpz_clear(t);
    return d;
  }

  void value_of(const std::vector<i128>& d, mpz_t z) const {
    mpz_set_ui(z, 0);
    mpz_t t;
    mpz_init(t);
    for (size_t i = n; i-- > 0;) {
      mpz_mul_2exp(z, z, cbits);
      set_i128(t, d[i]);
      mpz_add(z, z, t);
    }
    mpz_clear(t);
  }

  std::vector<std::vector<uint32_t>> residues(const std::vector<i128>& d) const {
    std::vector<std::vector<uint32_t>> r(K, std::vector<uint32_t>(n));
    for (int k = 0; k < K; ++k)
      for (size_t i = 0; i < n; ++i) r[k][i] = res_of(d[i], k);
    return r;
  }

  std::vector<i128> reduce(const std::vector<i128>& tr, const std::vector<i128>& tq) {
    std::vector<std::vector<uint32_t>> c;
    conv(tr, nullptr, NIR, false, c);
    MP = carry(c, 1, false, 1, nullptr);
    conv(MP, nullptr, NQ, true, c);
    return carry(c, 1, true, 2, &tq);
  }

  void square_mul(uint32_t a) {
    std::vector<std::vector<uint32_t>> c;
    conv(reg[0], nullptr, nullptr, false, c);
    TR = carry(c, a, false, 0, nullptr);
    conv(reg[0], nullptr, nullptr, true, c);
    TQ = carry(c, a, true, 0, nullptr);
    reg[0] = reduce(TR, TQ);
  }

  void mul(int dst, int src) {
    std::vector<std::vector<uint32_t>> c;
    conv(reg[dst], &reg[src], nullptr, false, c);
    TR = carry(c, 1, false, 0, nullptr);
    conv(reg[dst], &reg[src], nullptr, true, c);
    TQ = carry(c, 1, true, 0, nullptr);
    reg[dst] = reduce(TR, TQ);
  }

  void get(int src, mpz_t z) {
    TQ = carry(residues(reg[src]), 1, true, 0, nullptr);
    const std::vector<i128> r = reduce(reg[src], TQ);
    value_of(r, z);
    mpz_mod(z, z, N);
  }
};

void from_limbs(mpz_t z, const uint64_t* l, size_t n) { mpz_import(z, n, -1, 8, 0, 0, l); }

long to_limbs(const mpz_t z, uint64_t* out, size_t cap) {
  const size_t need = mpz_sgn(z) == 0 ? 0 : (mpz_sizeinbase(z, 2) + 63) / 64;
  if (need > cap) return -1;
  size_t cnt = 0;
  if (need) mpz_export(out, &cnt, -1, 8, 0, 0, z);
  return long(need);
}

}  // namespace

extern "C" {

void on_plan(const uint64_t* Nl, size_t nl, int32_t a, int* K, int* pck, int* LN, int* lbase) {
  mpz_t N;
  mpz_init(N);
  from_limbs(N, Nl, nl);
  long e = 0;
  const double m = mpz_get_d_2exp(&e, N);
  plan(m, e, a, *K, *pck, *LN, *lbase);
  mpz_clear(N);
}

void on_plan_d(double mant, long ex, int32_t a, int* K, int* pck, int* LN, int* lbase) {
  plan(mant, ex, a, *K, *pck, *LN, *lbase);
}

void* on_create(const uint64_t* Nl, size_t nl, int32_t a) {
  Oracle* o = new Oracle();
  mpz_inits(o->N, o->M, o->Mh, o->factor, nullptr);
  from_limbs(o->N, Nl, nl);
  long e = 0;
  const double m = mpz_get_d_2exp(&e, o->N);
  plan(m, e, a, o->K, o->pck, o->LN, o->lbase);
  o->cbits = o->pck * o->lbase;
  o->n = size_t(1) << o->LN;
  mpz_t R, t;
  mpz_inits(R, t, nullptr);
  mpz_set_ui(R, 1);
  mpz_mul_2exp(R, R, mp_bitcnt_t(o->cbits) * o->n);
  mpz_sub_ui(R, R, 1);
  if (mpz_invert(t, o->N, R) == 0) {
    mpz_gcd(o->factor, o->N, R);
    o->failed = true;
    mpz_clears(R, t, nullptr);
    return o;
  }
  mpz_set_ui(o->M, 1);
  for (int k = 0; k < o->K; ++k) mpz_mul_ui(o->M, o->M, P[k]);
  mpz_tdiv_q_2exp(o->Mh, o->M, 1);
  for (int k = 0; k < o->K; ++k) {
    mpz_t* c = new mpz_t[1];
    mpz_init(*c);
    mpz_t Mk, inv, pk;
    mpz_inits(Mk, inv, pk, nullptr);
    mpz_divexact_ui(Mk, o->M, P[k]);
    mpz_set_ui(pk, P[k]);
    mpz_invert(inv, Mk, pk);
    mpz_mul(*c, Mk, inv);
    mpz_clears(Mk, inv, pk, nullptr);
    o->crt_coef.push_back(c);
  }
  o->tables();
  std::vector<i128> d = o->digits_of(o->N);
  for (int k = 0; k < o->K; ++k) {
    o->NQ[k].resize(o->n);
    for (size_t i = 0; i < o->n; ++i) o->NQ[k][i] = o->res_of(d[i], k);
    o->fwd(o->NQ[k], k, true);
  }
  d = o->digits_of(t);
  for (int k = 0; k < o->K; ++k) {
    o->NIR[k].resize(o->n);
    for (size_t i = 0; i < o->n; ++i) o->NIR[k][i] = o->res_of(d[i], k);
    o->fwd(o->NIR[k], k, false);
  }
  mpz_mod(t, R, o->N);
  if (mpz_odd_p(t)) mpz_add(t, t, o->N);
  mpz_tdiv_q_2exp(t, t, 1);
  o->one = o->digits_of(t);
  for (int r = 0; r < 3; ++r) o->reg[r].assign(o->n, 0);
  mpz_clears(R, t, nullptr);
  return o;
}

int on_failed(void* h) { return static_cast<Oracle*>(h)->failed ? 1 : 0; }

void on_info(void* h, int* K, int* pck, int* LN, int* lbase) {
  Oracle* o = static_cast<Oracle*>(h);
  *K = o->K;
  *pck = o->pck;
  *LN = o->LN;
  *lbase = o->lbase;
}

// Every operation is a no-op on a context whose init failed (gcd(N, R) > 1).
#define OK_OR_RETURN(o, ...) \
  if ((o)->failed) return __VA_ARGS__;
void on_set1(void* h) { Oracle* o = static_cast<Oracle*>(h); OK_OR_RETURN(o); o->reg[0] = o->one; }
void on_copy(void* h, int d, int s) { Oracle* o = static_cast<Oracle*>(h); OK_OR_RETURN(o); o->reg[d] = o->reg[s]; }
void on_square_mul(void* h, uint32_t a) { Oracle* o = static_cast<Oracle*>(h); OK_OR_RETURN(o); o->square_mul(a); }
void on_mul(void* h, int d, int s) { Oracle* o = static_cast<Oracle*>(h); OK_OR_RETURN(o); o->mul(d, s); }

long on_get(void* h, int s, uint64_t* out, size_t cap) {
  Oracle* o = static_cast<Oracle*>(h);
  OK_OR_RETURN(o, -1);
  mpz_t z;
  mpz_init(z);
  o->get(s, z);
  const long r = to_limbs(z, out, cap);
  mpz_clear(z);
  return r;
}

// Residues (K planes of n) of register r as the device stores them: 0-2 user
// registers, 3 the MMGW one, 4 m' + 2 of the last reduction.
void on_read_register(void* h, int r, uint32_t* out) {
  Oracle* o = static_cast<Oracle*>(h);
  OK_OR_RETURN(o);
  const std::vector<i128>& d = (r == 3) ? o->one : (r == 4 ? o->MP : o->reg[r]);
  for (int k = 0; k < o->K; ++k)
    for (size_t i = 0; i < o->n; ++i) out[size_t(k) * o->n + i] = o->res_of(d[i], k);
}

// Signed digits of T_Q (last square_mul or mul) as three 32-bit limb planes.
void on_read_tq(void* h, uint32_t* out) {
  Oracle* o = static_cast<Oracle*>(h);
  OK_OR_RETURN(o);
  for (size_t i = 0; i < o->n; ++i) {
    const unsigned __int128 u = (unsigned __int128)o->TQ[i];
    out[i] = uint32_t(u);
    out[o->n + i] = uint32_t(u >> 32);
    out[2 * o->n + i] = uint32_t(u >> 64);
  }
}

// Set user register r from residues (K planes), as device writes do; the
// digits are recovered by CRT (they are small).
void on_write_register(void* h, int r, const uint32_t* in) {
  Oracle* o = static_cast<Oracle*>(h);
  OK_OR_RETURN(o);
  std::vector<std::vector<uint32_t>> res(o->K, std::vector<uint32_t>(o->n));
  for (int k = 0; k < o->K; ++k)
    for (size_t i = 0; i < o->n; ++i) res[k][i] = in[size_t(k) * o->n + i];
  mpz_t v;
  mpz_init(v);
  for (size_t i = 0; i < o->n; ++i) {
    o->crt(res, i, v);
    o->reg[r][i] = Oracle::get_i128(v);
  }
  mpz_clear(v);
}

// Intermediates of one squaring of r0 (r0 not updated), matching the device's
// annie_hip_debug_square_stage: stage 1 cyclic / negacyclic convolution
// residues (out_r, out_q: K planes); 2 T_R residues (out_r) and T_Q digits
// (out_q: 3 limb planes); 3 the m' convolution (out_r); 4 m' + 2 (out_r).
void on_debug_square_stage(void* h, uint32_t a, int stage, uint32_t* out_r, uint32_t* out_q) {
  Oracle* o = static_cast<Oracle*>(h);
  OK_OR_RETURN(o);
  const size_t n = o->n;
  std::vector<std::vector<uint32_t>> cr, cq;
  o->conv(o->reg[0], nullptr, nullptr, false, cr);
  o->conv(o->reg[0], nullptr, nullptr, true, cq);
  auto put = [&](const std::vector<std::vector<uint32_t>>& v, uint32_t* out) {
    for (int k = 0; k < o->K; ++k)
      for (size_t i = 0; i < n; ++i) out[size_t(k) * n + i] = v[k][i];
  };
  if (stage == 1) {
    put(cr, out_r);
    put(cq, out_q);
    return;
  }
  const std::vector<i128> tr = o->carry(cr, a, false, 0, nullptr);
  const std::vector<i128> tq = o->carry(cq, a, true, 0, nullptr);
  if (stage == 2) {
    put(o->residues(tr), out_r);
    for (size_t i = 0; i < n; ++i) {
      const unsigned __int128 u = (unsigned __int128)tq[i];
      out_q[i] = uint32_t(u);
      out_q[n + i] = uint32_t(u >> 32);
      out_q[2 * n + i] = uint32_t(u >> 64);
    }
    return;
  }
  std::vector<std::vector<uint32_t>> cm;
  o->conv(tr, nullptr, o->NIR, false, cm);
  if (stage == 3) {
    put(cm, out_r);
    return;
  }
  put(o->residues(o->carry(cm, 1, false, 1, nullptr)), out_r);
}

void on_destroy(void* h) {
  Oracle* o = static_cast<Oracle*>(h);
  for (mpz_t* c : o->crt_coef) {
    mpz_clear(*c);
    delete[] c;
  }
  mpz_clears(o->N, o->M, o->Mh, o->factor, nullptr);
  delete o;
}

// ---- building blocks for unit tests -------------------------------------

// In-place forward / inverse transform of one prime plane (length 2^LN) with
// the oracle's root tables.
void on_ntt(int LN, int k, int nega, int inverse, uint32_t* x) {
  Oracle o;
  o.K = k + 1;
  o.LN = LN;
  o.n = size_t(1) << LN;
  o.tables();
  std::vector<uint32_t> v(x, x + o.n);
  if (inverse) o.inv(v, k, nega != 0);
  else o.fwd(v, k, nega != 0);
  std::memcpy(x, v.data(), o.n * 4);
}

// Root table `t` (0 cyclic fwd, 1 cyclic inv, 2 negacyclic fwd, 3 negacyclic
// inv) of prime k, plain residues, heap-indexed.
void on_tables(int LN, int k, int t, uint32_t* out) {
  Oracle o;
  o.K = k + 1;
  o.LN = LN;
  o.n = size_t(1) << LN;
  o.tables();
  std::memcpy(out, o.tw[t][k].data(), o.n * 4);
}

// Signed CRT value of K residues as 3 64-bit limbs (two's complement, 192 bits).
void on_crt(int K, const uint32_t* r, uint64_t* out) {
  mpz_t M, Mh, v, t, inv, pk, Mk;
  mpz_inits(M, Mh, v, t, inv, pk, Mk, nullptr);
  mpz_set_ui(M, 1);
  for (int k = 0; k < K; ++k) mpz_mul_ui(M, M, P[k]);
  mpz_tdiv_q_2exp(Mh, M, 1);
  for (int k = 0; k < K; ++k) {
    mpz_divexact_ui(Mk, M, P[k]);
    mpz_set_ui(pk, P[k]);
    mpz_invert(inv, Mk, pk);
    mpz_mul(t, Mk, inv);
    mpz_mul_ui(t, t, r[k]);
    mpz_add(v, v, t);
  }
  mpz_mod(v, v, M);
  if (mpz_cmp(v, Mh) > 0) mpz_sub(v, v, M);
  // two's complement in 192 bits
  if (mpz_sgn(v) < 0) {
    mpz_set_ui(t, 1);
    mpz_mul_2exp(t, t, 192);
    mpz_add(v, v, t);
  }
  out[0] = out[1] = out[2] = 0;
  size_t cnt = 0;
  mpz_export(out, &cnt, -1, 8, 0, 0, v);
  mpz_clears(M, Mh, v, t, inv, pk, Mk, nullptr);
}

}  // extern "C"
