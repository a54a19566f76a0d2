// gfx950 kernels of the modular-squaring hot path: RNS number-theoretic
// transforms staged through LDS, fused pointwise products, and the CRT/carry
// pass that turns convolution residues back into packed digits.
//
// Transform definition (per prime p, length n = 2^LN).  The forward transform
// runs the Cooley-Tukey butterfly (a, b) -> (a + r b, a - r b) level by level,
// level l pairing elements h = n / 2^{l+1} apart inside node b of that level,
// with the root r(l, b) from the heap-indexed table (entry 2^l + b).  Its output
// is the evaluation of the input at the roots of X^n - 1 (cyclic, the mod-R
// product) or X^n + 1 (negacyclic, the mod-Q product) in a fixed permuted
// order; products are taken elementwise and the inverse Gentleman-Sande
// butterflies (a, b) -> (a + b, (a - b) r^-1) undo the levels in reverse,
// giving n times the cyclic/negacyclic convolution.  This is the transform of
// ocl/kernel.cl:206-284 (FWD2/BCK2 over the roots of src/transformGPU.h:116-160)
// run to the last level instead of stopping at degree-1 pairs.
//
// Lengths above 2^13 use two passes over HBM per direction: a column pass runs
// the first m1 levels on lines of stride n2 = n / 2^m1, a row pass runs the
// remaining levels on contiguous rows of n2 elements, and fuses the pointwise
// product and the inverse row levels.  Each thread holds 32 elements in
// registers (two or more independent groups) and runs four levels per LDS
// round trip.
#include <type_traits>

#include "kernels.h"

namespace annie {

struct Tw2 {
  uint32_t w, wq;
};

__device__ __forceinline__ void fwd_bfly(uint32_t& a, uint32_t& b, const Tw2 w, const uint32_t p) {
  const uint32_t t = mont_c(b, w.w, w.wq, p);
  b = submod(a, t, p);
  a = addmod(a, t, p);
}

__device__ __forceinline__ void inv_bfly(uint32_t& a, uint32_t& b, const Tw2 w, const uint32_t p) {
  const uint32_t s = addmod(a, b, p);
  const uint32_t d = submod(a, b, p);
  a = s;
  b = mont_c(d, w.w, w.wq, p);
}

__device__ __forceinline__ Tw2 ldtw(const TwDev* t, uint32_t i, uint32_t q) {
  const uint32_t w = t[i];
  return Tw2{w, w * q};
}

// Levels [d0, d0 + RR) on one group of 2^RR register elements; H0 is the heap
// index of the group's node at level d0 (children of heap node h are 2h, 2h+1).
template <int RR>
__device__ __forceinline__ void group_fwd(uint32_t* x, uint32_t H0, const TwDev* tw, uint32_t p, uint32_t q) {
#pragma unroll
  for (int e = 0; e < RR; ++e) {
    constexpr int N = 1 << RR;
    const int half = N >> (e + 1);
#pragma unroll
    for (int blk = 0; blk < (1 << e); ++blk) {
      const Tw2 w = ldtw(tw, (H0 << e) + blk, q);
#pragma unroll
      for (int jj = 0; jj < half; ++jj) {
        const int j = blk * 2 * half + jj;
        fwd_bfly(x[j], x[j + half], w, p);
      }
    }
  }
}

template <int RR>
__device__ __forceinline__ void group_inv(uint32_t* x, uint32_t H0, const TwDev* tw, uint32_t p, uint32_t q) {
#pragma unroll
  for (int e = RR - 1; e >= 0; --e) {
    constexpr int N = 1 << RR;
    const int half = N >> (e + 1);
#pragma unroll
    for (int blk = 0; blk < (1 << e); ++blk) {
      const Tw2 w = ldtw(tw, (H0 << e) + blk, q);
#pragma unroll
      for (int jj = 0; jj < half; ++jj) {
        const int j = blk * 2 * half + jj;
        inv_bfly(x[j], x[j + half], w, p);
      }
    }
  }
}

// Position inside the line of element jj of sub-group s held by thread slot tg
// in a round covering levels [d0, d0 + RR) of a line of 2^m elements.  A thread
// holds NTT_R >> RR groups of 2^RR elements, group g = tg + s * (2^m / NTT_R)
// (consecutive threads take consecutive groups, so round 0 reads and the
// last inverse round writes coalesced); group g spans
// node*2^{m-d0} + off + j*2^{m-d0-RR}.
template <int RR>
__device__ __forceinline__ uint32_t group_base(int m, int d0, uint32_t tg, int s, uint32_t& node) {
  const uint32_t g = tg + (uint32_t(s) << (m - NTT_LOGR));
  const int lst = m - d0 - RR;
  node = g >> lst;
  const uint32_t off = g & ((1u << lst) - 1u);
  return (node << (m - d0)) + off;
}

// Move the thread's 16 elements of round (d0, RR) between registers and a store.
template <int RR, class Ld>
__device__ __forceinline__ void round_load(uint32_t (&x)[NTT_R], int m, int d0, uint32_t tg, Ld&& ld) {
#pragma unroll
  for (int s = 0; s < (NTT_R >> RR); ++s) {
    uint32_t node;
    const uint32_t b = group_base<RR>(m, d0, tg, s, node);
    const int lst = m - d0 - RR;
#pragma unroll
    for (int j = 0; j < (1 << RR); ++j) x[s * (1 << RR) + j] = ld(b + (uint32_t(j) << lst));
  }
}

template <int RR, class St>
__device__ __forceinline__ void round_store(const uint32_t (&x)[NTT_R], int m, int d0, uint32_t tg, St&& st) {
#pragma unroll
  for (int s = 0; s < (NTT_R >> RR); ++s) {
    uint32_t node;
    const uint32_t b = group_base<RR>(m, d0, tg, s, node);
    const int lst = m - d0 - RR;
#pragma unroll
    for (int j = 0; j < (1 << RR); ++j) st(b + (uint32_t(j) << lst), x[s * (1 << RR) + j]);
  }
}

template <int RR>
__device__ __forceinline__ void round_fwd(uint32_t (&x)[NTT_R], int m, int d0, uint32_t tg, uint32_t lineRoot,
                                          const TwDev* tw, uint32_t p, uint32_t q) {
#pragma unroll
  for (int s = 0; s < (NTT_R >> RR); ++s) {
    uint32_t node;
    group_base<RR>(m, d0, tg, s, node);
    group_fwd<RR>(&x[s * (1 << RR)], (lineRoot << d0) + node, tw, p, q);
  }
}

template <int RR>
__device__ __forceinline__ void round_inv(uint32_t (&x)[NTT_R], int m, int d0, uint32_t tg, uint32_t lineRoot,
                                          const TwDev* tw, uint32_t p, uint32_t q) {
#pragma unroll
  for (int s = 0; s < (NTT_R >> RR); ++s) {
    uint32_t node;
    group_base<RR>(m, d0, tg, s, node);
    group_inv<RR>(&x[s * (1 << RR)], (lineRoot << d0) + node, tw, p, q);
  }
}

// Round schedule of a line of 2^M elements: rounds of four levels, the last
// one taking the remainder.  Everything is fixed at compile time, so each
// round reads and writes a statically known memory (global or LDS).
template <int M>
struct Sched {
  static constexpr int NR = (M + 3) / 4;
  template <int I>
  static constexpr int rr() { return (M - 4 * I) < 4 ? (M - 4 * I) : 4; }
};

// Forward rounds I.. of a line: round 0 loads through gld, later rounds
// through LDS; every round but the last stores to LDS, and the last leaves its
// elements in registers.
template <int M, int I, class GLd, class LLd, class LSt>
__device__ __forceinline__ void fwd_rounds(uint32_t (&x)[NTT_R], uint32_t tg, uint32_t lineRoot, const TwDev* tw,
                                           uint32_t p, uint32_t q, GLd& gld, LLd& lld, LSt& lst) {
  constexpr int RR = Sched<M>::template rr<I>();
  if constexpr (I == 0) round_load<RR>(x, M, 4 * I, tg, gld);
  else round_load<RR>(x, M, 4 * I, tg, lld);
  round_fwd<RR>(x, M, 4 * I, tg, lineRoot, tw, p, q);
  if constexpr (I + 1 < Sched<M>::NR) {
    round_store<RR>(x, M, 4 * I, tg, lst);
    __syncthreads();
    fwd_rounds<M, I + 1>(x, tg, lineRoot, tw, p, q, gld, lld, lst);
  }
}

// Inverse rounds I down to 0, starting (I = NR - 1) from the registers of the
// last forward round; round 0 stores through gst.
template <int M, int I, class GSt, class LLd, class LSt>
__device__ __forceinline__ void inv_rounds(uint32_t (&x)[NTT_R], uint32_t tg, uint32_t lineRoot, const TwDev* tw,
                                           uint32_t p, uint32_t q, GSt& gst, LLd& lld, LSt& lst) {
  constexpr int RR = Sched<M>::template rr<I>();
  if constexpr (I + 1 < Sched<M>::NR) round_load<RR>(x, M, 4 * I, tg, lld);
  round_inv<RR>(x, M, 4 * I, tg, lineRoot, tw, p, q);
  if constexpr (I > 0) {
    round_store<RR>(x, M, 4 * I, tg, lst);
    __syncthreads();
    inv_rounds<M, I - 1>(x, tg, lineRoot, tw, p, q, gst, lld, lst);
  } else {
    round_store<RR>(x, M, 0, tg, gst);
  }
}

__device__ __forceinline__ uint32_t lpad(uint32_t i) { return i + (i >> 5); }

// Row pass: the levels [m1, LN) of one contiguous row of 2^M elements per
// line, then (by mode) the pointwise product and the inverse row levels.
// Grid: (rows / lines-per-block, K, jobs).  Block: lines * 2^M / NTT_R threads.
template <int M, int MODE>
__global__ void __launch_bounds__(512) k_row(const NttArgs A) {
  extern __shared__ uint32_t lds[];
  constexpr uint32_t S = 1u << M, Tl = S / NTT_R;
  constexpr int NR = Sched<M>::NR, DL = 4 * (NR - 1), RL = Sched<M>::template rr<NR - 1>();
  const NttJob& J = A.job[blockIdx.z];
  const int k = blockIdx.y;
  const uint32_t p = A.p[k];
  const size_t n = size_t(1) << A.LN;
  const uint32_t line = threadIdx.x / Tl, tg = threadIdx.x % Tl;
  const uint32_t row = blockIdx.x * uint32_t(A.cols) + line;
  const size_t base = size_t(k) * n + size_t(row) * S;
  const uint32_t lineRoot = (1u << A.m1) + row;
  const TwDev* twf = J.fwd + size_t(k) * n;
  const TwDev* twi = J.inv + size_t(k) * n;
  uint32_t* X = lds + line * (S + (S >> 5));
  const uint32_t* src = J.src + base;
  uint32_t* dst = J.dst + base;

  auto gld = [&](uint32_t i) { return src[i]; };
  auto gst = [&](uint32_t i, uint32_t v) { dst[i] = v; };
  auto lld = [&](uint32_t i) { return X[lpad(i)]; };
  auto lst = [&](uint32_t i, uint32_t v) { X[lpad(i)] = v; };

  uint32_t x[NTT_R];
  fwd_rounds<M, 0>(x, tg, lineRoot, twf, p, A.q[k], gld, lld, lst);
  if constexpr (MODE == ROW_FWD) {
    round_store<RL>(x, M, DL, tg, gst);
  } else {
    // Pointwise product on the leaves held in registers.
    const uint32_t q = A.q[k], sc = A.sc[k], scq = A.scq[k];
    const uint32_t* op = (MODE == ROW_SQR) ? nullptr : J.opnd + base;
#pragma unroll
    for (int s = 0; s < (NTT_R >> RL); ++s) {
      uint32_t node;
      const uint32_t b = group_base<RL>(M, DL, tg, s, node);
#pragma unroll
      for (int j = 0; j < (1 << RL); ++j) {
        uint32_t& v = x[s * (1 << RL) + j];
        if constexpr (MODE == ROW_SQR) {
          v = mont_c(mont(v, v, p, q), sc, scq, p);
        } else if constexpr (MODE == ROW_MULC) {
          v = mont(v, op[b + uint32_t(j)], p, q);
        } else {
          v = mont_c(mont(v, op[b + uint32_t(j)], p, q), sc, scq, p);
        }
      }
    }
    inv_rounds<M, NR - 1>(x, tg, lineRoot, twi, p, A.q[k], gst, lld, lst);
  }
}

// Column passes: the first M levels on columns of stride n2 = n / 2^M; `cols`
// consecutive columns per block, interleaved in LDS so that global accesses
// of a wavefront touch consecutive words.  Grid: (n2 / cols, K, jobs).
template <int M>
__global__ void __launch_bounds__(1024) k_col_fwd(const NttArgs A) {
  extern __shared__ uint32_t lds[];
  constexpr int NR = Sched<M>::NR, DL = 4 * (NR - 1), RL = Sched<M>::template rr<NR - 1>();
  const NttJob& J = A.job[blockIdx.z];
  const int k = blockIdx.y;
  const uint32_t p = A.p[k];
  const size_t n = size_t(1) << A.LN;
  const uint32_t C = uint32_t(A.cols);
  const uint32_t line = threadIdx.x % C, tg = threadIdx.x / C;
  const size_t n2 = n >> M;
  const size_t base = size_t(k) * n + size_t(blockIdx.x) * C + line;
  const TwDev* twf = J.fwd + size_t(k) * n;
  const uint32_t* src = J.src + base;
  uint32_t* dst = J.dst + base;

  auto gld = [&](uint32_t i) { return src[size_t(i) * n2]; };
  auto gst = [&](uint32_t i, uint32_t v) { dst[size_t(i) * n2] = v; };
  auto lld = [&](uint32_t i) { return lds[i * C + line]; };
  auto lst = [&](uint32_t i, uint32_t v) { lds[i * C + line] = v; };

  uint32_t x[NTT_R];
  fwd_rounds<M, 0>(x, tg, 1u, twf, p, A.q[k], gld, lld, lst);
  round_store<RL>(x, M, DL, tg, gst);
}

template <int M>
__global__ void __launch_bounds__(1024) k_col_inv(const NttArgs A) {
  extern __shared__ uint32_t lds[];
  constexpr int NR = Sched<M>::NR, DL = 4 * (NR - 1), RL = Sched<M>::template rr<NR - 1>();
  const NttJob& J = A.job[blockIdx.z];
  const int k = blockIdx.y;
  const uint32_t p = A.p[k];
  const size_t n = size_t(1) << A.LN;
  const uint32_t C = uint32_t(A.cols);
  const uint32_t line = threadIdx.x % C, tg = threadIdx.x / C;
  const size_t n2 = n >> M;
  const size_t base = size_t(k) * n + size_t(blockIdx.x) * C + line;
  const TwDev* twi = J.inv + size_t(k) * n;
  const uint32_t* src = J.src + base;
  uint32_t* dst = J.dst + base;

  auto gld = [&](uint32_t i) { return src[size_t(i) * n2]; };
  auto gst = [&](uint32_t i, uint32_t v) { dst[size_t(i) * n2] = v; };
  auto lld = [&](uint32_t i) { return lds[i * C + line]; };
  auto lst = [&](uint32_t i, uint32_t v) { lds[i * C + line] = v; };

  uint32_t x[NTT_R];
  round_load<RL>(x, M, DL, tg, gld);
  inv_rounds<M, NR - 1>(x, tg, 1u, twi, p, A.q[k], gst, lld, lst);
}

// Elementwise Montgomery scaling x -> x * s * 2^-32 of K planes (used once at
// set-up to fold n^-1 into the pre-transformed constants).
__global__ void k_scale(uint32_t* x, size_t n, int K, NttArgs A, uint32_t s0, uint32_t s1, uint32_t s2, uint32_t s3,
                        uint32_t s4) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (i >= n || k >= K) return;
  const uint32_t s[KMAX] = {s0, s1, s2, s3, s4};
  const uint32_t p = A.p[k];
  x[size_t(k) * n + i] = mont(x[size_t(k) * n + i], s[k], p, A.q[k]);
}

// ---------------------------------------------------------------------------
// CRT + carry.
//
// Each thread owns CARRY_L consecutive coefficients.  Phase 1 rebuilds each
// coefficient's signed value from its K residues (Garner's mixed-radix form,
// as in ocl/kernel.cl:1418-1472), scales it (multiply by a, or subtract a raw
// digit), and splits the running sum into balanced digits of cbits bits plus
// a carry-out computed with zero carry-in.  Phase 2 adds the previous block's
// zero-in carry to the first digits; whatever remains after the first
// CARRY_L - 1 digits stays in the last digit, so every digit ends in
// [-B/2 - 1, B/2).  The value is preserved exactly, a carry never travels
// more than one block (the lazy scheme of ocl/kernel.cl:1561-1613), and the
// whole number lies in (-B^n/2, B^n/2).  The first thread of a workgroup
// recomputes its predecessor's zero-in carry from that block's residues, so
// the pass is one launch with no inter-workgroup communication; block 0 takes
// the wrapped carry of the last block (negated for the negacyclic product).
//
// Canonical range.  mod Q the result r' < 4N < B^n/4 is the only value of its
// class in (-B^n/2, B^n/2), so it comes out exact with no fix-up (its top
// digits are zero: a sign scan there would walk them all); mod R, m' (uniform
// in [0, R)) is recovered from the sign of the number, that of its most
// significant nonzero digit (the scan of src/transformCPU.h:1261-1281, which
// stops at the top digit except with probability ~2^-cbits): a negative value
// gets R added.
// ---------------------------------------------------------------------------

template <int NL>
struct Big {
  uint32_t w[NL];
};

template <int NL>
__device__ __forceinline__ void big_set_i32(Big<NL>& a, int32_t v) {
  a.w[0] = uint32_t(v);
#pragma unroll
  for (int i = 1; i < NL; ++i) a.w[i] = uint32_t(v >> 31);
}

// a = a * m + c (mod 2^{32 NL}), m and c unsigned 32-bit.
template <int NL>
__device__ __forceinline__ void big_mad(Big<NL>& a, uint32_t m, uint32_t c) {
  uint64_t carry = c;
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    const uint64_t t = uint64_t(a.w[i]) * m + carry;
    a.w[i] = uint32_t(t);
    carry = t >> 32;
  }
}

template <int NL>
__device__ __forceinline__ void big_add(Big<NL>& a, const Big<NL>& b) {
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    const uint64_t t = uint64_t(a.w[i]) + b.w[i] + c;
    a.w[i] = uint32_t(t);
    c = uint32_t(t >> 32);
  }
}

template <int NL>
__device__ __forceinline__ void big_sub(Big<NL>& a, const Big<NL>& b) {
  uint32_t br = 0;
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    const uint64_t t = uint64_t(a.w[i]) - b.w[i] - br;
    a.w[i] = uint32_t(t);
    br = uint32_t(t >> 32) & 1u;
  }
}

// Sign-extend a DIGIT_LIMBS two's-complement digit to NL limbs.
template <int NL>
__device__ __forceinline__ Big<NL> big_from_digit(const uint32_t d[DIGIT_LIMBS]) {
  Big<NL> r;
#pragma unroll
  for (int i = 0; i < NL; ++i) r.w[i] = (i < DIGIT_LIMBS) ? d[i] : uint32_t(int32_t(d[DIGIT_LIMBS - 1]) >> 31);
  return r;
}

// Split f into a balanced digit d = f mod B in [-B/2, B/2) (B = 2^cbits,
// DIGIT_LIMBS two's-complement limbs) and f <- (f - d) / B.  cbits < 32 *
// DIGIT_LIMBS; limb shifts are selected without dynamic register indexing.
template <int NL>
__device__ __forceinline__ void big_split(Big<NL>& f, int cbits, uint32_t d[DIGIT_LIMBS]) {
  const int qs = cbits >> 5, s = cbits & 31;
  const uint32_t sign = uint32_t(int32_t(f.w[NL - 1]) >> 31);
  const uint32_t mask = (s == 0) ? 0u : ((1u << s) - 1u);
  const int sl = (cbits - 1) >> 5, sb = (cbits - 1) & 31;
  const uint32_t wsl = (sl == 0) ? f.w[0] : ((sl == 1) ? f.w[1] : f.w[2]);
  const uint32_t neg = (wsl >> sb) & 1u;
  const uint32_t fill = 0u - neg;
#pragma unroll
  for (int i = 0; i < DIGIT_LIMBS; ++i) {
    const uint32_t wi = (i < NL) ? f.w[i] : sign;
    d[i] = (i < qs) ? wi : ((i == qs) ? ((wi & mask) | (fill & ~mask)) : fill);
  }
  uint32_t g[NL + 1];
#pragma unroll
  for (int i = 0; i <= NL; ++i) {
    const uint32_t a0 = (i < NL) ? f.w[i] : sign;
    const uint32_t a1 = (i + 1 < NL) ? f.w[i + 1] : sign;
    const uint32_t a2 = (i + 2 < NL) ? f.w[i + 2] : sign;
    g[i] = (qs == 0) ? a0 : ((qs == 1) ? a1 : a2);
  }
  uint32_t c = neg;
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    const uint32_t w = (s == 0) ? g[i] : __builtin_amdgcn_alignbit(g[i + 1], g[i], uint32_t(s));
    const uint64_t t = uint64_t(w) + c;
    f.w[i] = uint32_t(t);
    c = uint32_t(t >> 32);
  }
}

// Signed value of one coefficient from its K residues (each in [0, p_k)):
// mixed-radix digits u_{K-1} = r_{K-1}, u_j = (...((r_j - u_{K-1}) / p_{K-1}) - ...) / p_{j+1}
// mod p_j, with u_0 taken in (-p_0/2, p_0/2], then Horner over the primes.
template <int K, int NL>
__device__ __forceinline__ Big<NL> garner(const uint32_t r[K], const CarryArgs& A) {
  uint32_t u[K];
#pragma unroll
  for (int j = K - 1; j >= 0; --j) {
    uint32_t v = r[j];
#pragma unroll
    for (int i = K - 1; i > j; --i) v = submul_c(v, u[i], A.ginv[i][j], A.ginvq[i][j], A.p[j]);
    u[j] = v;
  }
  const int32_t i0 = (u[0] > (A.p[0] >> 1)) ? int32_t(u[0] - A.p[0]) : int32_t(u[0]);
  Big<NL> acc;
  big_set_i32(acc, i0);
#pragma unroll
  for (int j = 1; j < K; ++j) big_mad(acc, A.p[j], u[j]);
  return acc;
}

// Residue of a signed digit (|d| < 2^95) modulo p_k.
__device__ __forceinline__ uint32_t digit_residue(const uint32_t d[DIGIT_LIMBS], int k, const CarryArgs& A) {
  const uint32_t p = A.p[k];
  const bool neg = int32_t(d[2]) < 0;
  uint32_t a0 = d[0], a1 = d[1], a2 = d[2];
  if (neg) {  // |d| = ~d + 1
    a0 = ~a0;
    a1 = ~a1;
    a2 = ~a2;
    const uint64_t t0 = uint64_t(a0) + 1;
    a0 = uint32_t(t0);
    const uint64_t t1 = uint64_t(a1) + (t0 >> 32);
    a1 = uint32_t(t1);
    a2 += uint32_t(t1 >> 32);
  }
  uint32_t r = (a0 >= p) ? a0 - p : a0;
  r = addmod(r, mont_c(a1, A.c32[k], A.c32q[k], p), p);
  r = addmod(r, mont_c(a2, A.c64[k], A.c64q[k], p), p);
  return neg ? submod(0u, r, p) : r;
}

// Phase 1 for block b: digits with zero carry-in and the carry-out.
template <int K, int NL>
__device__ __forceinline__ void carry_block_local(const CarryJob& J, const CarryArgs& A, size_t n, size_t b,
                                                  uint32_t dig[CARRY_L][DIGIT_LIMBS], Big<NL>& f) {
  const size_t i0 = b * CARRY_L;
  uint32_t res[K][CARRY_L];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const uint4 v = *reinterpret_cast<const uint4*>(J.src + size_t(k) * n + i0);
    res[k][0] = v.x;
    res[k][1] = v.y;
    res[k][2] = v.z;
    res[k][3] = v.w;
  }
  uint32_t sd[DIGIT_LIMBS][CARRY_L];
  if (J.sub != nullptr) {
#pragma unroll
    for (int l = 0; l < DIGIT_LIMBS; ++l) {
      const uint4 v = *reinterpret_cast<const uint4*>(J.sub + size_t(l) * n + i0);
      sd[l][0] = v.x;
      sd[l][1] = v.y;
      sd[l][2] = v.z;
      sd[l][3] = v.w;
    }
  }
  big_set_i32(f, 0);
#pragma unroll
  for (int j = 0; j < CARRY_L; ++j) {
    uint32_t r[K];
#pragma unroll
    for (int k = 0; k < K; ++k) r[k] = res[k][j];
    Big<NL> v = garner<K, NL>(r, A);
    if (J.a != 1u) big_mad(v, J.a, 0u);
    if (J.sub != nullptr) {
      const uint32_t d[DIGIT_LIMBS] = {sd[0][j], sd[1][j], sd[2][j]};
      big_sub(v, big_from_digit<NL>(d));
    }
    big_add(f, v);
    big_split(f, A.cbits, dig[j]);
  }
}

// Phase 2: add the carry-in to the block's digits.
template <int NL>
__device__ __forceinline__ void carry_block_finish(uint32_t dig[CARRY_L][DIGIT_LIMBS], Big<NL> f, int cbits) {
#pragma unroll
  for (int j = 0; j < CARRY_L - 1; ++j) {
    big_add(f, big_from_digit<NL>(dig[j]));
    big_split(f, cbits, dig[j]);
  }
  big_add(f, big_from_digit<NL>(dig[CARRY_L - 1]));
#pragma unroll
  for (int l = 0; l < DIGIT_LIMBS; ++l) dig[CARRY_L - 1][l] = f.w[l];
}

// d += v for a signed DIGIT_LIMBS digit and small signed v.
__device__ __forceinline__ void digit_add(uint32_t d[DIGIT_LIMBS], int32_t v) {
  const uint64_t t0 = uint64_t(d[0]) + uint32_t(v);
  d[0] = uint32_t(t0);
  const uint32_t ext = uint32_t(v >> 31);
  const uint64_t t1 = uint64_t(d[1]) + ext + (t0 >> 32);
  d[1] = uint32_t(t1);
  d[2] = d[2] + ext + uint32_t(t1 >> 32);
}

// d += sgn * 2^cbits.
__device__ __forceinline__ void digit_add_base(uint32_t d[DIGIT_LIMBS], int cbits, int sgn) {
  const int qs = cbits >> 5, s = cbits & 31;
  uint32_t b[DIGIT_LIMBS];
#pragma unroll
  for (int i = 0; i < DIGIT_LIMBS; ++i) b[i] = (i == qs) ? (1u << s) : 0u;
  if (sgn < 0) {
    uint32_t br = 0;
#pragma unroll
    for (int i = 0; i < DIGIT_LIMBS; ++i) {
      const uint64_t t = uint64_t(d[i]) - b[i] - br;
      d[i] = uint32_t(t);
      br = uint32_t(t >> 32) & 1u;
    }
  } else {
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < DIGIT_LIMBS; ++i) {
      const uint64_t t = uint64_t(d[i]) + b[i] + c;
      d[i] = uint32_t(t);
      c = uint32_t(t >> 32);
    }
  }
}

__device__ __forceinline__ int digit_sign(const uint32_t d[DIGIT_LIMBS]) {
  if ((d[0] | d[1] | d[2]) == 0u) return 0;
  return (int32_t(d[2]) < 0) ? -1 : 1;
}

// Sign of the whole number: scan blocks from the top, finishing each with the
// zero-in carry of its predecessor, until a nonzero digit appears (the top
// digit is nonzero except with probability ~2^-cbits per block).
template <int K, int NL>
__device__ int number_sign(const CarryJob& J, const CarryArgs& A, size_t n) {
  const size_t nb = n / CARRY_L;
  for (size_t b = nb; b-- > 0;) {
    uint32_t d[CARRY_L][DIGIT_LIMBS], pd[CARRY_L][DIGIT_LIMBS];
    Big<NL> f, fi;
    carry_block_local<K, NL>(J, A, n, b, d, f);
    carry_block_local<K, NL>(J, A, n, (b == 0) ? nb - 1 : b - 1, pd, fi);
    if (b == 0 && J.nega) {
      Big<NL> z;
      big_set_i32(z, 0);
      big_sub(z, fi);
      fi = z;
    }
    carry_block_finish<NL>(d, fi, A.cbits);
#pragma unroll
    for (int j = CARRY_L - 1; j >= 0; --j) {
      const int sg = digit_sign(d[j]);
      if (sg != 0) return sg;
    }
  }
  return 0;
}

template <int K>
__global__ void __launch_bounds__(CARRY_WG) k_carry(const CarryArgs A) {
  constexpr int NL = K + 1;
  __shared__ Big<NL> cl[CARRY_WG];
  const CarryJob& J = A.job[blockIdx.y];
  const size_t n = size_t(1) << A.LN;
  const size_t nb = n / CARRY_L;
  const uint32_t t = threadIdx.x;
  const size_t b = size_t(blockIdx.x) * blockDim.x + t;
  const bool active = b < nb;

  uint32_t dig[CARRY_L][DIGIT_LIMBS];
  Big<NL> fo;
  if (active) carry_block_local<K, NL>(J, A, n, b, dig, fo);
  cl[t] = fo;
  __syncthreads();
  if (!active) return;

  Big<NL> fin;
  if (t > 0) {
    fin = cl[t - 1];
  } else {
    uint32_t pd[CARRY_L][DIGIT_LIMBS];
    const size_t pb = (b == 0) ? nb - 1 : b - 1;
    carry_block_local<K, NL>(J, A, n, pb, pd, fin);
    if (b == 0 && J.nega) {
      Big<NL> z;
      big_set_i32(z, 0);
      big_sub(z, fin);
      fin = z;
    }
  }
  carry_block_finish<NL>(dig, fin, A.cbits);

  if (J.canon == 1 && (b == 0 || b == nb - 1)) {
    // m' into [0, R): both ends of the number take the same decision, adding
    // R = B^n - 1 to a negative value; then m' + 2.  (mod Q needs nothing: the
    // result is already the unique member of its class in (-B^n/2, B^n/2).)
    const int sg = number_sign<K, NL>(J, A, n);
    if (sg < 0 && b == nb - 1) digit_add_base(dig[CARRY_L - 1], A.cbits, 1);
    if (b == 0) digit_add(dig[0], ((sg < 0) ? -1 : 0) + 2);
  }

  const size_t i0 = b * CARRY_L;
  if (J.raw_out) {
#pragma unroll
    for (int l = 0; l < DIGIT_LIMBS; ++l) {
      uint4 v;
      v.x = dig[0][l];
      v.y = dig[1][l];
      v.z = dig[2][l];
      v.w = dig[3][l];
      *reinterpret_cast<uint4*>(J.dst + size_t(l) * n + i0) = v;
    }
  } else {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      uint4 v;
      v.x = digit_residue(dig[0], k, A);
      v.y = digit_residue(dig[1], k, A);
      v.z = digit_residue(dig[2], k, A);
      v.w = digit_residue(dig[3], k, A);
      *reinterpret_cast<uint4*>(J.dst + size_t(k) * n + i0) = v;
    }
  }
}

// Raw signed digits -> K residue planes.
template <int K>
__global__ void k_digits_to_res(const uint32_t* digits, uint32_t* dst, const CarryArgs A) {
  const size_t n = size_t(1) << A.LN;
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t d[DIGIT_LIMBS] = {digits[i], digits[n + i], digits[2 * n + i]};
#pragma unroll
  for (int k = 0; k < K; ++k) dst[size_t(k) * n + i] = digit_residue(d, k, A);
}

// Explicit instantiations, looked up by the host engine: row lines of 2^5 ..
// 2^13 elements (single-pass lengths 2^5..2^13, second pass of 2^14..2^23) and
// column lines of 2^7 .. 2^11 elements (first pass of 2^14..2^23).
#define ANNIE_ROW(M)                                          \
  template __global__ void k_row<M, ROW_FWD>(const NttArgs);  \
  template __global__ void k_row<M, ROW_SQR>(const NttArgs);  \
  template __global__ void k_row<M, ROW_MULC>(const NttArgs); \
  template __global__ void k_row<M, ROW_MULV>(const NttArgs);
ANNIE_ROW(5)
ANNIE_ROW(6)
ANNIE_ROW(7)
ANNIE_ROW(8)
ANNIE_ROW(9)
ANNIE_ROW(10)
ANNIE_ROW(11)
ANNIE_ROW(12)
ANNIE_ROW(13)
#define ANNIE_COL(M)                                    \
  template __global__ void k_col_fwd<M>(const NttArgs); \
  template __global__ void k_col_inv<M>(const NttArgs);
ANNIE_COL(7)
ANNIE_COL(8)
ANNIE_COL(9)
ANNIE_COL(10)
ANNIE_COL(11)
template __global__ void k_carry<2>(const CarryArgs);
template __global__ void k_carry<3>(const CarryArgs);
template __global__ void k_carry<4>(const CarryArgs);
template __global__ void k_carry<5>(const CarryArgs);
template __global__ void k_digits_to_res<2>(const uint32_t*, uint32_t*, const CarryArgs);
template __global__ void k_digits_to_res<3>(const uint32_t*, uint32_t*, const CarryArgs);
template __global__ void k_digits_to_res<4>(const uint32_t*, uint32_t*, const CarryArgs);
template __global__ void k_digits_to_res<5>(const uint32_t*, uint32_t*, const CarryArgs);

}  // namespace annie
