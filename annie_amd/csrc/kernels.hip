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
// product and the inverse row levels.  A thread holds 2^LR elements in
// registers and runs up to LR levels per LDS round trip: 16 elements and four
// levels in the row passes and in columns of up to 2^9, 32 elements and five
// levels in columns of 2^10 and 2^11.
#include <type_traits>

#include "kernels.h"

namespace annie {

struct Tw2 {
  uint32_t w, wq;
};

// Word offset of the slot at list position pos (0 without a slot list).
__device__ __forceinline__ size_t slot_offset(const uint32_t* ids, size_t stride, uint32_t pos) {
  return ids ? size_t(ids[pos]) * stride : size_t(0);
}

// Forward (Cooley-Tukey) butterfly on lazily reduced residues: a and b are
// any words congruent to the values mod p (below 2^32; the primes are above
// 2^31, so a word holds at most two representatives), and so are the
// outputs.  t = r b 2^-32 mod p is in [0, p) for any b < 2^32.  a - t: one
// correction (+ p on borrow) leaves [0, 2^32).  a + t: on a carry out of 32
// bits the wrapped sum s < p gets 2^32 mod p = 2^32 - p added, which stays
// below 2^32 -- three instructions instead of the four of a full addmod.
// The inverse butterflies, whose outputs feed the CRT, take and give fully
// reduced residues; every product (mont, mont_c) accepts lazy operands.
__device__ __forceinline__ void fwd_bfly(uint32_t& a, uint32_t& b, const Tw2 w, const uint32_t p) {
  const uint32_t t = mont_c(b, w.w, w.wq, p);
  uint32_t s;
  const bool carry = __builtin_add_overflow(a, t, &s);
  b = submod(a, t, p);
  a = s + (carry ? (0u - p) : 0u);
}

__device__ __forceinline__ void inv_bfly(uint32_t& a, uint32_t& b, const Tw2 w, const uint32_t p) {
  const uint32_t s = addmod(a, b, p);
  const uint32_t d = submod(a, b, p);
  a = s;
  b = mont_c(d, w.w, w.wq, p);
}

// Roots of one register round (levels [d0, d0 + RR)) for each of the
// thread's 2^LR >> RR groups: the 2^RR - 1 roots of a group in heap order,
// level e of the round at entries 2^e - 1 .. 2^{e+1} - 2.  A round's roots are
// loaded one round ahead of their use, so their latency overlaps the previous
// round's butterflies and LDS exchange.
template <int LR, int RR>
struct RoundTw {
  static constexpr int G = (1 << LR) >> RR, W = (1 << RR) - 1;
  uint32_t w[G * W];
};

// Levels [0, RR) of one group of 2^RR register elements with the group's
// roots w (heap order, as in RoundTw).
template <int RR>
__device__ __forceinline__ void group_fwd(uint32_t* x, const uint32_t* w, uint32_t p, uint32_t q) {
#pragma unroll
  for (int e = 0; e < RR; ++e) {
    constexpr int N = 1 << RR;
    const int half = N >> (e + 1);
#pragma unroll
    for (int blk = 0; blk < (1 << e); ++blk) {
      const uint32_t r = w[(1 << e) - 1 + blk];
      const Tw2 t{r, r * q};
#pragma unroll
      for (int jj = 0; jj < half; ++jj) {
        const int j = blk * 2 * half + jj;
        fwd_bfly(x[j], x[j + half], t, p);
      }
    }
  }
}

template <int RR>
__device__ __forceinline__ void group_inv(uint32_t* x, const uint32_t* w, uint32_t p, uint32_t q) {
#pragma unroll
  for (int e = RR - 1; e >= 0; --e) {
    constexpr int N = 1 << RR;
    const int half = N >> (e + 1);
#pragma unroll
    for (int blk = 0; blk < (1 << e); ++blk) {
      const uint32_t r = w[(1 << e) - 1 + blk];
      const Tw2 t{r, r * q};
#pragma unroll
      for (int jj = 0; jj < half; ++jj) {
        const int j = blk * 2 * half + jj;
        inv_bfly(x[j], x[j + half], t, p);
      }
    }
  }
}

// Position inside the line of element jj of sub-group s held by thread slot tg
// in a round covering levels [d0, d0 + RR) of a line of 2^m elements.  A thread
// holds 2^LR elements (LR levels per full round): 2^LR >> RR groups of 2^RR
// elements, group g = tg + s * (2^m / 2^LR)
// (consecutive threads take consecutive groups, so round 0 reads and the
// last inverse round writes coalesced); group g spans
// node*2^{m-d0} + off + j*2^{m-d0-RR}.
template <int LR, int RR>
__device__ __forceinline__ uint32_t group_base(int m, int d0, uint32_t tg, int s, uint32_t& node) {
  const uint32_t g = tg + (uint32_t(s) << (m - LR));
  const int lst = m - d0 - RR;
  node = g >> lst;
  const uint32_t off = g & ((1u << lst) - 1u);
  return (node << (m - d0)) + off;
}

// Move the thread's 2^LR elements of round (d0, RR) between registers and a store.
template <int LR, int RR, class Ld>
__device__ __forceinline__ void round_load(uint32_t (&x)[1 << LR], int m, int d0, uint32_t tg, Ld&& ld) {
#pragma unroll
  for (int s = 0; s < ((1 << LR) >> RR); ++s) {
    uint32_t node;
    const uint32_t b = group_base<LR, RR>(m, d0, tg, s, node);
    const int lst = m - d0 - RR;
#pragma unroll
    for (int j = 0; j < (1 << RR); ++j) x[s * (1 << RR) + j] = ld(b + (uint32_t(j) << lst));
  }
}

template <int LR, int RR, class St>
__device__ __forceinline__ void round_store(const uint32_t (&x)[1 << LR], int m, int d0, uint32_t tg, St&& st) {
#pragma unroll
  for (int s = 0; s < ((1 << LR) >> RR); ++s) {
    uint32_t node;
    const uint32_t b = group_base<LR, RR>(m, d0, tg, s, node);
    const int lst = m - d0 - RR;
#pragma unroll
    for (int j = 0; j < (1 << RR); ++j) st(b + (uint32_t(j) << lst), x[s * (1 << RR) + j]);
  }
}

// Word i of a root table, line or operand (i < 2^30: tables and planes are at
// most 2^23 words), addressed by a 32-bit byte offset from the base, so the
// compiler can use the scalar-base + 32-bit vector-offset form instead of a
// 64-bit address per thread (fewer VGPRs: a row pass of 2^9 drops from 85 to
// 64 and runs eight waves per SIMD instead of five).
__device__ __forceinline__ uint32_t ld_w(const uint32_t* base, uint32_t i) {
  return *reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(base) + (i << 2));
}

__device__ __forceinline__ void st_w(uint32_t* base, uint32_t i, uint32_t v) {
  *reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(base) + (i << 2)) = v;
}

// Load the roots of round (d0, RR) of the line whose root node is lineRoot
// (heap index at level 0 of this pass).
template <int LR, int RR>
__device__ __forceinline__ void round_tw(RoundTw<LR, RR>& T, int m, int d0, uint32_t tg, uint32_t lineRoot,
                                         const TwDev* tw) {
#pragma unroll
  for (int s = 0; s < RoundTw<LR, RR>::G; ++s) {
    uint32_t node;
    group_base<LR, RR>(m, d0, tg, s, node);
    const uint32_t H0 = (lineRoot << d0) + node;
#pragma unroll
    for (int e = 0; e < RR; ++e)
#pragma unroll
      for (int blk = 0; blk < (1 << e); ++blk) T.w[s * RoundTw<LR, RR>::W + (1 << e) - 1 + blk] = ld_w(tw, (H0 << e) + blk);
  }
}

template <int LR, int RR>
__device__ __forceinline__ void round_fwd(uint32_t (&x)[1 << LR], const RoundTw<LR, RR>& T, uint32_t p, uint32_t q) {
#pragma unroll
  for (int s = 0; s < RoundTw<LR, RR>::G; ++s)
    group_fwd<RR>(&x[s * (1 << RR)], &T.w[s * RoundTw<LR, RR>::W], p, q);
}

template <int LR, int RR>
__device__ __forceinline__ void round_inv(uint32_t (&x)[1 << LR], const RoundTw<LR, RR>& T, uint32_t p, uint32_t q) {
#pragma unroll
  for (int s = 0; s < RoundTw<LR, RR>::G; ++s)
    group_inv<RR>(&x[s * (1 << RR)], &T.w[s * RoundTw<LR, RR>::W], p, q);
}

// Round schedule of a line of 2^M elements: rounds of LR levels (2^LR
// elements per thread), the last one taking the remainder.  Everything is
// fixed at compile time, so each round reads and writes a statically known
// memory (global or LDS).
template <int M, int LR>
struct Sched {
  static constexpr int NR = (M + LR - 1) / LR;
  template <int I>
  static constexpr int rr() { return (M - LR * I) < LR ? (M - LR * I) : LR; }
};

template <int M, int LR, int I>
using SchedTw = RoundTw<LR, Sched<M, LR>::template rr<I>()>;

// Forward rounds I.. of a line: round 0 loads through gld, later rounds
// through LDS; every round but the last stores to LDS, and the last leaves its
// elements in registers.  With AHEAD, round I's roots arrive in `pre` and each
// round first issues the loads of the next round's roots (so their latency
// overlaps a round of butterflies and the LDS exchange); otherwise each round
// loads its own roots (from LDS or global memory) and `pre` is unused.  `at_last`
// runs at the start of the last round (the caller's prefetches for what
// follows).  With PRE, round 0's elements are already in x.
template <int M, int LR, int I, bool AHEAD, bool PRE, class GLd, class LLd, class LSt, class Hook>
__device__ __forceinline__ void fwd_rounds(uint32_t (&x)[1 << LR], uint32_t tg, uint32_t lineRoot, const TwDev* tw,
                                           uint32_t p, uint32_t q, GLd& gld, LLd& lld, LSt& lst,
                                           const SchedTw<M, LR, I>& pre, Hook&& at_last) {
  constexpr int RR = Sched<M, LR>::template rr<I>();
  SchedTw<M, LR, I> own;
  if constexpr (!AHEAD) round_tw(own, M, LR * I, tg, lineRoot, tw);
  const SchedTw<M, LR, I>& cur = AHEAD ? pre : own;
  if constexpr (I + 1 < Sched<M, LR>::NR) {
    SchedTw<M, LR, I + 1> nxt;
    if constexpr (AHEAD) round_tw(nxt, M, LR * (I + 1), tg, lineRoot, tw);
    if constexpr (I == 0) {
      if constexpr (!PRE) round_load<LR, RR>(x, M, LR * I, tg, gld);
    } else {
      round_load<LR, RR>(x, M, LR * I, tg, lld);
    }
    round_fwd<LR, RR>(x, cur, p, q);
    round_store<LR, RR>(x, M, LR * I, tg, lst);
    __syncthreads();
    fwd_rounds<M, LR, I + 1, AHEAD, PRE>(x, tg, lineRoot, tw, p, q, gld, lld, lst, nxt, at_last);
  } else {
    at_last();
    if constexpr (I == 0) {
      if constexpr (!PRE) round_load<LR, RR>(x, M, LR * I, tg, gld);
    } else {
      round_load<LR, RR>(x, M, LR * I, tg, lld);
    }
    round_fwd<LR, RR>(x, cur, p, q);
  }
}

// Inverse rounds I down to 0, starting (I = NR - 1) from the registers of the
// last forward round; round 0 stores through gst.  Roots as in fwd_rounds:
// with AHEAD each round issues the loads of the next (lower) round's roots.
template <int M, int LR, int I, bool AHEAD, class GSt, class LLd, class LSt>
__device__ __forceinline__ void inv_rounds(uint32_t (&x)[1 << LR], uint32_t tg, uint32_t lineRoot, const TwDev* tw,
                                           uint32_t p, uint32_t q, GSt& gst, LLd& lld, LSt& lst,
                                           const SchedTw<M, LR, I>& pre) {
  constexpr int RR = Sched<M, LR>::template rr<I>();
  SchedTw<M, LR, I> own;
  if constexpr (!AHEAD) round_tw(own, M, LR * I, tg, lineRoot, tw);
  const SchedTw<M, LR, I>& cur = AHEAD ? pre : own;
  if constexpr (I > 0) {
    SchedTw<M, LR, I - 1> nxt;
    if constexpr (AHEAD) round_tw(nxt, M, LR * (I - 1), tg, lineRoot, tw);
    if constexpr (I + 1 < Sched<M, LR>::NR) round_load<LR, RR>(x, M, LR * I, tg, lld);
    round_inv<LR, RR>(x, cur, p, q);
    round_store<LR, RR>(x, M, LR * I, tg, lst);
    __syncthreads();
    inv_rounds<M, LR, I - 1, AHEAD>(x, tg, lineRoot, tw, p, q, gst, lld, lst, nxt);
  } else {
    if constexpr (I + 1 < Sched<M, LR>::NR) round_load<LR, RR>(x, M, LR * I, tg, lld);
    round_inv<LR, RR>(x, cur, p, q);
    round_store<LR, RR>(x, M, 0, tg, gst);
  }
}

__device__ __forceinline__ uint32_t lpad(uint32_t i) { return i + (i >> 5); }

// Row pass: the levels [m1, LN) of one contiguous row of 2^M elements per
// line, then (by mode) the pointwise product and the inverse row levels.
// 16 elements per thread (LR = 4 levels per register round; 32 elements per
// thread at rows of 2^11 measured no faster).
// Grid: (rows / lines-per-block, K, jobs).  Block: lines * 2^M / 16 threads.
// Rows of 2^11 (the 40M-digit length): five waves per SIMD, so that a one-job
// pass over five planes (10,240 waves) is exactly two resident rounds and a
// two-job pass four, instead of 2.5 and 5 at four waves (the half round left
// half the chip idle at the end of every one-job pass).
template <int M>
struct RowGeo {
  static constexpr int LR = NTT_LOGR;
  static constexpr int THREADS = 512;
  static constexpr int WAVES_PER_EU = (M == 11) ? 5 : 4;
};

// One line of a row pass (the squaring runs below; the launched pass k_row
// has the same body): levels [m1, LN) of the 2^M elements of job J's
// plane k at word `base` (slot offset + k n + row 2^M), whose root node is
// lineRoot, then (by mode) the pointwise product and the inverse levels.  tg
// is the thread's index among the 2^M / 16 threads of the line, X the line's
// LDS buffer (2^M + 2^M / 32 words).  Every thread of the workgroup must call
// it together (it synchronises the workgroup between rounds).
template <int M, int MODE, bool AH>
__device__ __forceinline__ void row_line(const NttJob& J, const NttArgs& A, int k, size_t base, uint32_t lineRoot,
                                         uint32_t tg, uint32_t* X) {
  constexpr int LR = RowGeo<M>::LR;
  constexpr int NR = Sched<M, LR>::NR, DL = LR * (NR - 1), RL = Sched<M, LR>::template rr<NR - 1>();
  const uint32_t p = A.p[k];
  const size_t n = size_t(1) << A.LN;
  const TwDev* twf = J.fwd + size_t(k) * n;
  const TwDev* twi = J.inv + size_t(k) * n;
  const uint32_t* src = J.src + base;
  uint32_t* dst = J.dst + base;

  auto gld = [&](uint32_t i) { return ld_w(src, i); };
  auto gst = [&](uint32_t i, uint32_t v) { st_w(dst, i, v); };
  auto lld = [&](uint32_t i) { return X[lpad(i)]; };
  auto lst = [&](uint32_t i, uint32_t v) { X[lpad(i)] = v; };

  // With AH, each round's roots are loaded one round ahead.
  uint32_t x[1 << LR];
  SchedTw<M, LR, 0> t0;
  if constexpr (AH) round_tw(t0, M, 0, tg, lineRoot, twf);
  if constexpr (MODE == ROW_FWD) {
    fwd_rounds<M, LR, 0, AH, false>(x, tg, lineRoot, twf, p, A.q[k], gld, lld, lst, t0, [] {});
    round_store<LR, RL>(x, M, DL, tg, gst);
  } else {
    // With AH the last forward round also issues the loads of the first
    // inverse round's roots and of the pointwise operand (MULC / MULV);
    // otherwise the operand is loaded just before the product.
    SchedTw<M, LR, NR - 1> ti;
    uint32_t opv[1 << LR];
    auto load_opv = [&] {
      if constexpr (MODE != ROW_SQR) {
        const uint32_t* op = J.opnd + base;
#pragma unroll
        for (int s = 0; s < ((1 << LR) >> RL); ++s) {
          uint32_t node;
          const uint32_t b = group_base<LR, RL>(M, DL, tg, s, node);
#pragma unroll
          for (int j = 0; j < (1 << RL); ++j) opv[s * (1 << RL) + j] = ld_w(op, b + uint32_t(j));
        }
      }
    };
    fwd_rounds<M, LR, 0, AH, false>(x, tg, lineRoot, twf, p, A.q[k], gld, lld, lst, t0, [&] {
      if constexpr (AH) {
        round_tw(ti, M, DL, tg, lineRoot, twi);
        load_opv();
      }
    });
    if constexpr (!AH) load_opv();
    // Pointwise product on the leaves held in registers.
    const uint32_t q = A.q[k], sc = A.sc[k], scq = A.scq[k];
#pragma unroll
    for (int i = 0; i < (1 << LR); ++i) {
      uint32_t& v = x[i];
      if constexpr (MODE == ROW_SQR) {
        v = mont_c(mont(v, v, p, q), sc, scq, p);
      } else if constexpr (MODE == ROW_MULC) {
        v = mont(v, opv[i], p, q);
      } else {
        v = mont_c(mont(v, opv[i], p, q), sc, scq, p);
      }
    }
    inv_rounds<M, LR, NR - 1, AH>(x, tg, lineRoot, twi, p, A.q[k], gst, lld, lst, ti);
  }
}

// The launched row pass.  Its body is row_line's, written out: with the
// inlined call the compiler allocates four more VGPRs and spills in the
// constant-multiply mode, which this pass (the largest per squaring at the
// longest lengths) cannot afford.
template <int M, int MODE>
__global__ void __launch_bounds__(RowGeo<M>::THREADS, RowGeo<M>::WAVES_PER_EU) k_row(const NttArgs A) {
  extern __shared__ uint32_t lds[];
  constexpr int LR = RowGeo<M>::LR;
  constexpr uint32_t S = 1u << M, Tl = S >> LR;
  constexpr int NR = Sched<M, LR>::NR, DL = LR * (NR - 1), RL = Sched<M, LR>::template rr<NR - 1>();
  const uint32_t nj = uint32_t(A.njobs);
  const NttJob& J = A.job[blockIdx.z % nj];
  const size_t so = slot_offset(A.slot_ids, A.slot_stride, blockIdx.z / nj);
  const int k = blockIdx.y;
  const uint32_t p = A.p[k];
  const size_t n = size_t(1) << A.LN;
  // Lines of 64 threads or more run one per block (the engine's launch), so
  // the row is block-uniform.
  constexpr bool ONE = Tl >= 64;
  const uint32_t line = ONE ? 0u : threadIdx.x / Tl, tg = ONE ? threadIdx.x : threadIdx.x % Tl;
  const uint32_t row = ONE ? blockIdx.x : blockIdx.x * uint32_t(A.cols) + line;
  const size_t base = so + size_t(k) * n + size_t(row) * S;
  const uint32_t lineRoot = (1u << A.m1) + row;
  const TwDev* twf = J.fwd + size_t(k) * n;
  const TwDev* twi = J.inv + size_t(k) * n;
  uint32_t* X = lds + line * (S + (S >> 5));
  const uint32_t* src = J.src + base;
  uint32_t* dst = J.dst + base;

  auto gld = [&](uint32_t i) { return ld_w(src, i); };
  auto gst = [&](uint32_t i, uint32_t v) { st_w(dst, i, v); };
  auto lld = [&](uint32_t i) { return X[lpad(i)]; };
  auto lst = [&](uint32_t i, uint32_t v) { X[lpad(i)] = v; };

  // Roots one round ahead, except for rows of 2^12 (three full radix-16
  // rounds), where the extra registers would leave three waves per SIMD
  // instead of four, and rows of 2^11, whose five waves per SIMD leave 96
  // registers.
  constexpr bool AH = (M != 12 && M != 11);
  uint32_t x[1 << LR];
  SchedTw<M, LR, 0> t0;
  if constexpr (AH) round_tw(t0, M, 0, tg, lineRoot, twf);
  if constexpr (MODE == ROW_FWD) {
    fwd_rounds<M, LR, 0, AH, false>(x, tg, lineRoot, twf, p, A.q[k], gld, lld, lst, t0, [] {});
    round_store<LR, RL>(x, M, DL, tg, gst);
  } else {
    // With AH the last forward round also issues the loads of the first
    // inverse round's roots and of the pointwise operand (MULC / MULV);
    // otherwise the operand is loaded just before the product.
    SchedTw<M, LR, NR - 1> ti;
    uint32_t opv[1 << LR];
    auto load_opv = [&] {
      if constexpr (MODE != ROW_SQR) {
        const uint32_t* op = J.opnd + base;
#pragma unroll
        for (int s = 0; s < ((1 << LR) >> RL); ++s) {
          uint32_t node;
          const uint32_t b = group_base<LR, RL>(M, DL, tg, s, node);
#pragma unroll
          for (int j = 0; j < (1 << RL); ++j) opv[s * (1 << RL) + j] = ld_w(op, b + uint32_t(j));
        }
      }
    };
    fwd_rounds<M, LR, 0, AH, false>(x, tg, lineRoot, twf, p, A.q[k], gld, lld, lst, t0, [&] {
      if constexpr (AH) {
        round_tw(ti, M, DL, tg, lineRoot, twi);
        load_opv();
      }
    });
    // Rows of 2^11 (96 registers): the operand is loaded one register group
    // at a time, each group's loads kept after the previous group's products
    // (the scheduling barrier), so that the 16 operand words are never live
    // together.
    constexpr bool GROUP_OP = (M == 11);
    if constexpr (!AH && !GROUP_OP) load_opv();
    // Pointwise product on the leaves held in registers.
    const uint32_t q = A.q[k], sc = A.sc[k], scq = A.scq[k];
    if constexpr (GROUP_OP && MODE != ROW_SQR) {
      const uint32_t* op = J.opnd + base;
#pragma unroll
      for (int s = 0; s < ((1 << LR) >> RL); ++s) {
        uint32_t node;
        const uint32_t b = group_base<LR, RL>(M, DL, tg, s, node);
#pragma unroll
        for (int j = 0; j < (1 << RL); ++j) {
          uint32_t& v = x[s * (1 << RL) + j];
          const uint32_t o = ld_w(op, b + uint32_t(j));
          if constexpr (MODE == ROW_MULC) v = mont(v, o, p, q);
          else v = mont_c(mont(v, o, p, q), sc, scq, p);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    } else
#pragma unroll
    for (int i = 0; i < (1 << LR); ++i) {
      uint32_t& v = x[i];
      if constexpr (MODE == ROW_SQR) {
        v = mont_c(mont(v, v, p, q), sc, scq, p);
      } else if constexpr (MODE == ROW_MULC) {
        v = mont(v, opv[i], p, q);
      } else {
        v = mont_c(mont(v, opv[i], p, q), sc, scq, p);
      }
    }
    inv_rounds<M, LR, NR - 1, AH>(x, tg, lineRoot, twi, p, A.q[k], gst, lld, lst, ti);
  }
}

// Column group of a column-pass workgroup.  Workgroups are dispatched to the
// eight XCDs round-robin (XCD = workgroup id mod 8), and each XCD has its own
// L2; a group of C columns covers 4C bytes of a 128-byte line, so the G = 32 / C
// workgroups x, x + 8, ..., x + 8 (G - 1) (same XCD, dispatched together) take
// the G parts of one line: within each run of 8G ids, x -> G (x mod 8) +
// (x div 8 mod G), a permutation of the run.
__device__ __forceinline__ uint32_t col_group(uint32_t C) {
  const uint32_t x = blockIdx.x;
  const uint32_t G = 32u / C;
  if (G < 2u || gridDim.x % (8u * G) != 0u) return x;
  return (x & ~(8u * G - 1u)) | ((x & 7u) * G) | ((x >> 3) & (G - 1u));
}

// Column passes: the first M levels on columns of stride n2 = n / 2^M; `cols`
// consecutive columns per block, interleaved in LDS so that global accesses
// of a wavefront touch consecutive words.
//
// Columns of up to 2^9 elements: 16 elements per thread, 16 columns per block;
// all columns use the same 2^M - 1 roots (heap entries 1 .. 2^M - 1), which
// the block stages in LDS after the column data while the data loads are in
// flight.  Columns of 2^10 and 2^11: 32 elements per thread (rounds of five
// levels, so 2^10 takes one LDS exchange) and 8 columns per block; a block
// of 2^10 is 256 threads and 32 KiB of LDS, so five fit a CU and a one-job
// pass of five planes of 2^21 (1,280 blocks) is resident in a single round
// (16 elements per thread and 16 columns leave 640 blocks for 512 slots).
// Their roots are read from the table in global memory (L2-resident: every
// block of a plane reads the same 2^M - 1 words).
// Grid: (n2 / cols, K, jobs); dynamic LDS cols * 2^M words (+ 2^M for staged
// roots).
template <int M>
struct ColGeo {
  static constexpr int LR = (M >= 10) ? 5 : 4;
  static constexpr bool LDS_ROOTS = (LR == 4);
  static constexpr int THREADS = (M >= 10) ? (8 << (M - LR)) : 1024;
  static constexpr int WAVES_PER_EU = (M == 10) ? 5 : ((M == 11) ? 4 : 8);
};

template <int M>
__global__ void __launch_bounds__(ColGeo<M>::THREADS, ColGeo<M>::WAVES_PER_EU) k_col_fwd(const NttArgs A) {
  extern __shared__ uint32_t lds[];
  constexpr int LR = ColGeo<M>::LR;
  constexpr uint32_t S = 1u << M;
  constexpr int NR = Sched<M, LR>::NR, DL = LR * (NR - 1), RL = Sched<M, LR>::template rr<NR - 1>();
  const uint32_t nj = uint32_t(A.njobs);
  const NttJob& J = A.job[blockIdx.z % nj];
  const size_t so = slot_offset(A.slot_ids, A.slot_stride, blockIdx.z / nj);
  const int k = blockIdx.y;
  const uint32_t p = A.p[k];
  const size_t n = size_t(1) << A.LN;
  // 8 columns per block at 32 elements per thread (a compile-time constant, so
  // that the LDS swizzle below folds); the engine's choice otherwise.
  const uint32_t C = (LR == 5) ? 8u : uint32_t(A.cols);
  const uint32_t line = threadIdx.x % C, tg = threadIdx.x / C;
  // Block-uniform base pointers and 32-bit per-thread offsets (line + i n2 <
  // 2^32), so global accesses take a scalar base and a 32-bit vector offset.
  const uint32_t n2 = uint32_t(n >> M);
  const size_t base = so + size_t(k) * n + size_t(col_group(C)) * C;
  const TwDev* twf = J.fwd + size_t(k) * n;
  const uint32_t* src = J.src + base;
  uint32_t* dst = J.dst + base;

  auto gld = [&](uint32_t i) { return src[line + i * n2]; };
  auto gst = [&](uint32_t i, uint32_t v) { dst[line + i * n2] = v; };
  // Element i of column `line` sits in LDS row i at column line ^ ((i >> LR) &
  // (C - 1)): a bijection within each row (C is a power of two), so no extra
  // space; without it the lanes of a wave that share `line` in a round whose
  // groups are 2^LR consecutive elements (i = 2^LR tg + j) all hit one bank.
  auto lld = [&](uint32_t i) { return lds[i * C + (line ^ ((i >> LR) & (C - 1u)))]; };
  auto lst = [&](uint32_t i, uint32_t v) { lds[i * C + (line ^ ((i >> LR) & (C - 1u)))] = v; };

  uint32_t x[1 << LR];
  SchedTw<M, LR, 0> t0;  // round 0's roots when they are loaded ahead
  if constexpr (ColGeo<M>::LDS_ROOTS) {
    uint32_t* roots = lds + C * S;
    round_load<LR, Sched<M, LR>::template rr<0>()>(x, M, 0, tg, gld);
    for (uint32_t i = threadIdx.x; i < S; i += blockDim.x) roots[i] = twf[i];
    __syncthreads();
    fwd_rounds<M, LR, 0, false, true>(x, tg, 1u, roots, p, A.q[k], gld, lld, lst, t0, [] {});
  } else {
    // Roots one round ahead for columns of 2^10 (96 VGPRs, five waves per
    // SIMD); columns of 2^11 would spill, so they load each round's roots.
    constexpr bool AH = (M == 10);
    if constexpr (AH) round_tw(t0, M, 0, tg, 1u, twf);
    fwd_rounds<M, LR, 0, AH, false>(x, tg, 1u, twf, p, A.q[k], gld, lld, lst, t0, [] {});
  }
  round_store<LR, RL>(x, M, DL, tg, gst);
}

template <int M>
__global__ void __launch_bounds__(ColGeo<M>::THREADS, ColGeo<M>::WAVES_PER_EU) k_col_inv(const NttArgs A) {
  extern __shared__ uint32_t lds[];
  constexpr int LR = ColGeo<M>::LR;
  constexpr uint32_t S = 1u << M;
  constexpr int NR = Sched<M, LR>::NR, DL = LR * (NR - 1), RL = Sched<M, LR>::template rr<NR - 1>();
  const uint32_t nj = uint32_t(A.njobs);
  const NttJob& J = A.job[blockIdx.z % nj];
  const size_t so = slot_offset(A.slot_ids, A.slot_stride, blockIdx.z / nj);
  const int k = blockIdx.y;
  const uint32_t p = A.p[k];
  const size_t n = size_t(1) << A.LN;
  // 8 columns per block at 32 elements per thread (a compile-time constant, so
  // that the LDS swizzle below folds); the engine's choice otherwise.
  const uint32_t C = (LR == 5) ? 8u : uint32_t(A.cols);
  const uint32_t line = threadIdx.x % C, tg = threadIdx.x / C;
  const uint32_t n2 = uint32_t(n >> M);
  const size_t base = so + size_t(k) * n + size_t(col_group(C)) * C;
  const TwDev* twi = J.inv + size_t(k) * n;
  const uint32_t* src = J.src + base;
  uint32_t* dst = J.dst + base;

  auto gld = [&](uint32_t i) { return src[line + i * n2]; };
  auto gst = [&](uint32_t i, uint32_t v) { dst[line + i * n2] = v; };
  // Element i of column `line` sits in LDS row i at column line ^ ((i >> LR) &
  // (C - 1)): a bijection within each row (C is a power of two), so no extra
  // space; without it the lanes of a wave that share `line` in a round whose
  // groups are 2^LR consecutive elements (i = 2^LR tg + j) all hit one bank.
  auto lld = [&](uint32_t i) { return lds[i * C + (line ^ ((i >> LR) & (C - 1u)))]; };
  auto lst = [&](uint32_t i, uint32_t v) { lds[i * C + (line ^ ((i >> LR) & (C - 1u)))] = v; };

  uint32_t x[1 << LR];
  SchedTw<M, LR, NR - 1> ti;  // the first inverse round's roots when loaded ahead
  round_load<LR, RL>(x, M, DL, tg, gld);
  if constexpr (ColGeo<M>::LDS_ROOTS) {
    uint32_t* roots = lds + C * S;
    for (uint32_t i = threadIdx.x; i < S; i += blockDim.x) roots[i] = twi[i];
    __syncthreads();
    inv_rounds<M, LR, NR - 1, false>(x, tg, 1u, roots, p, A.q[k], gst, lld, lst, ti);
  } else {
    constexpr bool AH = (M == 10);
    if constexpr (AH) round_tw(ti, M, DL, tg, 1u, twi);
    inv_rounds<M, LR, NR - 1, AH>(x, tg, 1u, twi, p, A.q[k], gst, lld, lst, ti);
  }
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
// whole number lies in (-B^n/2, B^n/2).  Workgroup w owns the CARRY_WG - 1
// blocks [w (CARRY_WG - 1), (w + 1)(CARRY_WG - 1)); its thread 0 runs phase 1
// on the block just before them (the last block, for w = 0) only to hand its
// zero-in carry to thread 1, so the pass is one launch with no
// inter-workgroup communication; block 0 takes the wrapped carry of the last
// block (negated for the negacyclic product).
//
// Canonical range.  mod Q the result r' < 4N < B^n/4 is the only value of its
// class in (-B^n/2, B^n/2), so it comes out exact with no fix-up (its top
// digits are zero: a sign scan there would walk them all); mod R, m' (uniform
// in [0, R)) is recovered from the sign of the number, that of its most
// significant nonzero digit (the scan of src/transformCPU.h:1261-1281, which
// stops at the top digit except with probability ~2^-cbits): a negative value
// gets R added.
//
// Multi-limb integers are two's complement, little-endian 32-bit limbs.  The
// digit width cbits = 32 QS + s has its limb shift QS as a template parameter.
// ---------------------------------------------------------------------------

// a = a * m + c (mod 2^{32 L}), m and c unsigned 32-bit.
template <int L>
__device__ __forceinline__ void big_mad(uint32_t* a, uint32_t m, uint32_t c) {
  uint64_t carry = c;
#pragma unroll
  for (int i = 0; i < L; ++i) {
    const uint64_t t = uint64_t(a[i]) * m + carry;
    a[i] = uint32_t(t);
    carry = t >> 32;
  }
}

template <int L>
__device__ __forceinline__ void big_add(uint32_t* a, const uint32_t* b) {
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < L; ++i) {
    const uint64_t t = uint64_t(a[i]) + b[i] + c;
    a[i] = uint32_t(t);
    c = uint32_t(t >> 32);
  }
}

template <int L>
__device__ __forceinline__ void big_sub(uint32_t* a, const uint32_t* b) {
  uint32_t br = 0;
#pragma unroll
  for (int i = 0; i < L; ++i) {
    const uint64_t t = uint64_t(a[i]) - b[i] - br;
    a[i] = uint32_t(t);
    br = uint32_t(t >> 32) & 1u;
  }
}

// Sign-extend a DIGIT_LIMBS digit to L limbs.
template <int L>
__device__ __forceinline__ void big_from_digit(const uint32_t (&d)[DIGIT_LIMBS], uint32_t* r) {
  const uint32_t sg = uint32_t(int32_t(d[DIGIT_LIMBS - 1]) >> 31);
#pragma unroll
  for (int i = 0; i < L; ++i) r[i] = (i < DIGIT_LIMBS) ? d[i] : sg;
}

// Split f into a balanced digit d = f mod B in [-B/2, B/2) (B = 2^cbits,
// cbits = 32 QS + s, DIGIT_LIMBS limbs) and f <- (f - d) / B.
template <int L, int QS>
__device__ __forceinline__ void big_split(uint32_t (&f)[L], uint32_t s, uint32_t (&d)[DIGIT_LIMBS]) {
  static_assert(QS < L && QS < DIGIT_LIMBS, "digit wider than the accumulator");
  const uint32_t sign = uint32_t(int32_t(f[L - 1]) >> 31);
  uint32_t neg;  // bit cbits - 1 of f
  if constexpr (QS == 0) {
    neg = (f[0] >> ((s - 1u) & 31u)) & 1u;
  } else {
    neg = (s == 0u) ? (f[QS - 1] >> 31) : ((f[QS] >> ((s - 1u) & 31u)) & 1u);
  }
  const uint32_t mask = (1u << s) - 1u;
  const uint32_t fill = 0u - neg;
#pragma unroll
  for (int i = 0; i < DIGIT_LIMBS; ++i) d[i] = (i < QS) ? f[i] : ((i == QS) ? ((f[QS] & mask) | (fill & ~mask)) : fill);
  uint32_t c = neg;
#pragma unroll
  for (int i = 0; i < L; ++i) {
    const uint32_t lo = (i + QS < L) ? f[i + QS] : sign;
    const uint32_t hi = (i + QS + 1 < L) ? f[i + QS + 1] : sign;
    const uint64_t t = uint64_t(__builtin_amdgcn_alignbit(hi, lo, s)) + c;
    f[i] = uint32_t(t);
    c = uint32_t(t >> 32);
  }
}

// Signed value of one coefficient from its K residues (each in [0, p_k)):
// mixed-radix digits u_{K-1} = r_{K-1}, u_j = (...((r_j - u_{K-1}) / p_{K-1}) - ...) / p_{j+1}
// mod p_j, with u_0 taken in (-p_0/2, p_0/2], then Horner over the primes,
// the accumulator growing by one limb per step (after step j it holds a
// signed value below 2^{32 (j+1) - 1} in magnitude).  Result sign-extended
// to L >= K limbs.
template <int K, int L>
__device__ __forceinline__ void garner(const uint32_t (&r)[K], const CarryArgs& A, uint32_t (&acc)[L]) {
  uint32_t u[K];
#pragma unroll
  for (int j = K - 1; j >= 0; --j) {
    uint32_t v = r[j];
#pragma unroll
    for (int i = K - 1; i > j; --i) v = submul_c(v, u[i], PLIT.ginv[i][j], PLIT.ginvq[i][j], PRIMES[j]);
    u[j] = v;
  }
  const int32_t i0 = (u[0] > (PRIMES[0] >> 1)) ? int32_t(u[0] - PRIMES[0]) : int32_t(u[0]);
  acc[0] = uint32_t(i0);
  uint32_t sg = uint32_t(i0 >> 31);
#pragma unroll
  for (int j = 1; j < K; ++j) {
    const uint32_t pj = PRIMES[j];
    uint64_t c = u[j];
#pragma unroll
    for (int i = 0; i < j; ++i) {
      const uint64_t t = uint64_t(acc[i]) * pj + c;
      acc[i] = uint32_t(t);
      c = t >> 32;
    }
    acc[j] = sg * pj + uint32_t(c);
    sg = uint32_t(int32_t(acc[j]) >> 31);
  }
#pragma unroll
  for (int i = K; i < L; ++i) acc[i] = sg;
}

// Residue modulo p_k of a digit of magnitude m1 2^32 + m0 < 2^64 (m1 <= 2^31)
// and sign neg: u = m 2^-32 mod p by one Montgomery reduction (m - (m0 p^-1
// mod 2^32) p is a multiple of 2^32), then u * (+-2^64) * 2^-32.
__device__ __forceinline__ uint32_t digit_residue64(uint32_t m0, uint32_t m1, bool neg, int k, const CarryArgs& A) {
  const uint32_t p = PRIMES[k];
  const uint32_t u = submod(m1, mulhi32(m0 * PLIT.q[k], p), p);
  const uint32_t w = neg ? PLIT.c32n[k] : PLIT.c32[k];
  const uint32_t wq = neg ? PLIT.c32nq[k] : PLIT.c32q[k];
  return mont_c(u, w, wq, p);
}

// Residue of a signed digit (|d| < 2^95) modulo p_k.
__device__ __forceinline__ uint32_t digit_residue_wide(const uint32_t (&d)[DIGIT_LIMBS], int k, const CarryArgs& A) {
  const uint32_t p = PRIMES[k];
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
  r = addmod(r, mont_c(a1, PLIT.c32[k], PLIT.c32q[k], p), p);
  r = addmod(r, mont_c(a2, PLIT.c64[k], PLIT.c64q[k], p), p);
  return neg ? submod(0u, r, p) : r;
}

// K residues of CARRY_L digits, written as uint4 per plane at dst + i0.
template <int K>
__device__ __forceinline__ void store_residues(const uint32_t (&dig)[CARRY_L][DIGIT_LIMBS], const CarryArgs& A,
                                               uint32_t* dst, size_t n, size_t i0, int cbits) {
  if (cbits <= 64) {
    // |d| <= 2^(cbits-1) + 1 < 2^64: the low 64 bits give the magnitude.
    uint32_t m0[CARRY_L], m1[CARRY_L];
    bool neg[CARRY_L];
#pragma unroll
    for (int j = 0; j < CARRY_L; ++j) {
      neg[j] = int32_t(dig[j][2]) < 0;
      const uint64_t lo = uint64_t(dig[j][0]) | (uint64_t(dig[j][1]) << 32);
      const uint64_t m = neg[j] ? (0ull - lo) : lo;
      m0[j] = uint32_t(m);
      m1[j] = uint32_t(m >> 32);
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      uint4 v;
      v.x = digit_residue64(m0[0], m1[0], neg[0], k, A);
      v.y = digit_residue64(m0[1], m1[1], neg[1], k, A);
      v.z = digit_residue64(m0[2], m1[2], neg[2], k, A);
      v.w = digit_residue64(m0[3], m1[3], neg[3], k, A);
      *reinterpret_cast<uint4*>(dst + size_t(k) * n + i0) = v;
    }
  } else {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      uint4 v;
      v.x = digit_residue_wide(dig[0], k, A);
      v.y = digit_residue_wide(dig[1], k, A);
      v.z = digit_residue_wide(dig[2], k, A);
      v.w = digit_residue_wide(dig[3], k, A);
      *reinterpret_cast<uint4*>(dst + size_t(k) * n + i0) = v;
    }
  }
}

// Inputs of one block: the K residues of its CARRY_L coefficients and, for a
// subtracting carry, the raw digits of the subtrahend.
template <int K>
struct CarryIn {
  uint32_t res[K][CARRY_L];
  uint32_t sd[DIGIT_LIMBS][CARRY_L];
};

template <int K>
__device__ __forceinline__ void carry_load(const CarryJob& J, size_t n, size_t b, CarryIn<K>& in) {
  const size_t i0 = b * CARRY_L;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const uint4 v = *reinterpret_cast<const uint4*>(J.src + size_t(k) * n + i0);
    in.res[k][0] = v.x;
    in.res[k][1] = v.y;
    in.res[k][2] = v.z;
    in.res[k][3] = v.w;
  }
  if (J.sub != nullptr) {
#pragma unroll
    for (int l = 0; l < DIGIT_LIMBS; ++l) {
      const uint4 v = *reinterpret_cast<const uint4*>(J.sub + size_t(l) * n + i0);
      in.sd[l][0] = v.x;
      in.sd[l][1] = v.y;
      in.sd[l][2] = v.z;
      in.sd[l][3] = v.w;
    }
  }
}

// Phase 1 on loaded inputs: digits with zero carry-in and the carry-out f.
template <int K, int L, int QS>
__device__ __forceinline__ void carry_phase1(const CarryJob& J, const CarryArgs& A, uint32_t s, const CarryIn<K>& in,
                                             uint32_t (&dig)[CARRY_L][DIGIT_LIMBS], uint32_t (&f)[L]) {
#pragma unroll
  for (int i = 0; i < L; ++i) f[i] = 0u;
#pragma unroll
  for (int j = 0; j < CARRY_L; ++j) {
    uint32_t r[K];
#pragma unroll
    for (int k = 0; k < K; ++k) r[k] = in.res[k][j];
    uint32_t v[L];
    garner<K, L>(r, A, v);
    if (J.a != 1u) big_mad<L>(v, J.a, 0u);
    if (J.sub != nullptr) {
      const uint32_t d[DIGIT_LIMBS] = {in.sd[0][j], in.sd[1][j], in.sd[2][j]};
      uint32_t e[L];
      big_from_digit<L>(d, e);
      big_sub<L>(v, e);
    }
    big_add<L>(f, v);
    big_split<L, QS>(f, s, dig[j]);
  }
}

// Phase 1 for block b.
template <int K, int L, int QS>
__device__ __forceinline__ void carry_block_local(const CarryJob& J, const CarryArgs& A, size_t n, size_t b, uint32_t s,
                                                  uint32_t (&dig)[CARRY_L][DIGIT_LIMBS], uint32_t (&f)[L]) {
  CarryIn<K> in;
  carry_load<K>(J, n, b, in);
  carry_phase1<K, L, QS>(J, A, s, in, dig, f);
}

// Phase 2: add the carry-in f to the block's digits.
template <int L, int QS>
__device__ __forceinline__ void carry_block_finish(uint32_t (&dig)[CARRY_L][DIGIT_LIMBS], uint32_t (&f)[L], uint32_t s) {
#pragma unroll
  for (int j = 0; j < CARRY_L - 1; ++j) {
    uint32_t e[L];
    big_from_digit<L>(dig[j], e);
    big_add<L>(f, e);
    big_split<L, QS>(f, s, dig[j]);
  }
  uint32_t e[L];
  big_from_digit<L>(dig[CARRY_L - 1], e);
  big_add<L>(f, e);
#pragma unroll
  for (int l = 0; l < DIGIT_LIMBS; ++l) dig[CARRY_L - 1][l] = f[l];
}

template <int L>
__device__ __forceinline__ void big_neg(uint32_t (&f)[L]) {
  uint32_t z[L];
#pragma unroll
  for (int i = 0; i < L; ++i) z[i] = 0u;
  big_sub<L>(z, f);
#pragma unroll
  for (int i = 0; i < L; ++i) f[i] = z[i];
}

// d += v for a signed DIGIT_LIMBS digit and small signed v.
__device__ __forceinline__ void digit_add(uint32_t (&d)[DIGIT_LIMBS], int32_t v) {
  const uint64_t t0 = uint64_t(d[0]) + uint32_t(v);
  d[0] = uint32_t(t0);
  const uint32_t ext = uint32_t(v >> 31);
  const uint64_t t1 = uint64_t(d[1]) + ext + (t0 >> 32);
  d[1] = uint32_t(t1);
  d[2] = d[2] + ext + uint32_t(t1 >> 32);
}

// d += 2^cbits.
__device__ __forceinline__ void digit_add_base(uint32_t (&d)[DIGIT_LIMBS], int cbits) {
  const int qs = cbits >> 5, s = cbits & 31;
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < DIGIT_LIMBS; ++i) {
    const uint64_t t = uint64_t(d[i]) + ((i == qs) ? (1u << s) : 0u) + c;
    d[i] = uint32_t(t);
    c = uint32_t(t >> 32);
  }
}

__device__ __forceinline__ int digit_sign(const uint32_t (&d)[DIGIT_LIMBS]) {
  if ((d[0] | d[1] | d[2]) == 0u) return 0;
  return (int32_t(d[2]) < 0) ? -1 : 1;
}

// Sign of the whole number: scan blocks from the top, finishing each with the
// zero-in carry of its predecessor, until a nonzero digit appears (the top
// digit is nonzero except with probability ~2^-cbits per block).
template <int K, int L, int QS>
__device__ int number_sign(const CarryJob& J, const CarryArgs& A, size_t n, uint32_t s) {
  const size_t nb = n / CARRY_L;
  for (size_t b = nb; b-- > 0;) {
    uint32_t d[CARRY_L][DIGIT_LIMBS], pd[CARRY_L][DIGIT_LIMBS];
    uint32_t f[L], fi[L];
    carry_block_local<K, L, QS>(J, A, n, b, s, d, f);
    carry_block_local<K, L, QS>(J, A, n, (b == 0) ? nb - 1 : b - 1, s, pd, fi);
    if (b == 0 && J.nega) big_neg<L>(fi);
    carry_block_finish<L, QS>(d, fi, s);
#pragma unroll
    for (int j = CARRY_L - 1; j >= 0; --j) {
      const int sg = digit_sign(d[j]);
      if (sg != 0) return sg;
    }
  }
  return 0;
}

// One finished block to the destination: raw digit planes, or K residue
// planes (and the second destination of a flagged slot).
template <int K>
__device__ __forceinline__ void carry_store(const CarryJob& J, const CarryArgs& A, size_t n, size_t b,
                                            const uint32_t (&dig)[CARRY_L][DIGIT_LIMBS], int cbits) {
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
    store_residues<K>(dig, A, J.dst, n, i0, cbits);
    if (J.dst2 != nullptr) store_residues<K>(dig, A, J.dst2, n, i0, cbits);
  }
}

// Grid: (ceil(nb / (CARRY_WG - 1)), jobs), nb = n / CARRY_L blocks.
// Workgroup w owns the CARRY_WG - 1 blocks [w (CARRY_WG - 1), (w + 1)
// (CARRY_WG - 1)); thread t > 0 takes block w (CARRY_WG - 1) + t - 1 and
// thread 0 the block before them (the last block for w = 0), only to hand
// its zero-in carry to thread 1 through LDS.
template <int K, int QS>
__global__ void __launch_bounds__(CARRY_WG) k_carry(const CarryArgs A) {
  constexpr int L = K + 1;
  __shared__ uint32_t cl[L][CARRY_WG];
  const uint32_t nj = uint32_t(A.njobs);
  CarryJob J = A.job[blockIdx.y % nj];
  int cbits = A.cbits;
  {
    const size_t so = slot_offset(A.slot_ids, A.slot_stride, blockIdx.y / nj);
    J.src += so;
    J.dst += so;
    if (J.sub != nullptr) J.sub += so;
    const uint32_t pos = blockIdx.y / nj;
    const bool bit = A.bits_mode != 0 && ((A.bits[pos >> 5] >> (pos & 31u)) & 1u) != 0u;
    if (A.a_slot != nullptr) J.a = A.a_slot[pos] & ~SLOT_SNAP;
    if (A.bits_mode == 1) J.a = bit ? A.slot_base[A.slot_ids[pos]] : 1u;
    const bool snap = (A.snap != nullptr && (A.snap[pos] & SLOT_SNAP)) || (A.bits_mode == 2 && bit);
    if (J.dst2 != nullptr) J.dst2 = snap ? J.dst2 + so : nullptr;
    if (A.cbits_slot != nullptr) cbits = int(A.cbits_slot[A.slot_ids ? A.slot_ids[blockIdx.y / nj] : 0u]);
  }
  const size_t n = size_t(1) << A.LN;
  const size_t nb = n / CARRY_L;
  const uint32_t t = threadIdx.x;
  const uint32_t s = uint32_t(cbits) & 31u;
  const size_t g = size_t(blockIdx.x) * (CARRY_WG - 1) + t;  // owned block + 1
  const size_t b = (g == 0) ? nb - 1 : g - 1;
  const bool active = g <= nb;

  uint32_t dig[CARRY_L][DIGIT_LIMBS];
  uint32_t fo[L];
  if (active) {
    carry_block_local<K, L, QS>(J, A, n, b, s, dig, fo);
#pragma unroll
    for (int i = 0; i < L; ++i) cl[i][t] = fo[i];
  }
  __syncthreads();
  if (!active || t == 0) return;

  uint32_t fin[L];
#pragma unroll
  for (int i = 0; i < L; ++i) fin[i] = cl[i][t - 1];
  if (b == 0 && J.nega) big_neg<L>(fin);
  carry_block_finish<L, QS>(dig, fin, s);

  if (J.canon == 1 && (b == 0 || b == nb - 1)) {
    // m' into [0, R): both ends of the number take the same decision, adding
    // R = B^n - 1 to a negative value; then m' + 2.  (mod Q needs nothing: the
    // result is already the unique member of its class in (-B^n/2, B^n/2).)
    const int sg = number_sign<K, L, QS>(J, A, n, s);
    if (sg < 0 && b == nb - 1) digit_add_base(dig[CARRY_L - 1], cbits);
    if (b == 0) digit_add(dig[0], ((sg < 0) ? -1 : 0) + 2);
  }

  carry_store<K>(J, A, n, b, dig, cbits);
}

// ---------------------------------------------------------------------------
// Runs of batched squarings: one workgroup owns one slot and runs whole
// squarings, step after step, with workgroup barriers where the launched
// path has kernel boundaries.  Each squaring is the launched path's sequence
// over the slot's registers (L2-resident at these lengths): the cyclic and
// negacyclic convolutions of x^2 (row_line, one line per prime plane and
// product), their carries, the m' convolution and carry (canonical, + 2), the
// (m' + 2) N convolution and the carry that subtracts T_Q.  Lengths 2^9 ..
// 2^13 (single-pass transforms); the carry keeps CARRY_L coefficients per
// block and the lazy one-block hand-off of k_carry, so the registers hold
// the same words as after the launched path.

// Threads per workgroup: one line of 2^M / 16 threads (two lines of 32 at
// 2^9), and carry blocks per thread.
template <int M>
struct RunGeo {
  static constexpr int TL = (1 << M) / NTT_R;
  static constexpr int T = TL >= 64 ? TL : 64;
  static constexpr int LINES = T / TL;
  static constexpr int NBT = (1 << M) / CARRY_L / T;
  // Waves per SIMD the register budget is set for: the row stages and the
  // carry together need up to ~160 VGPRs at 2^9 .. 2^12 (three waves), more
  // at 2^13 (two waves); a budget of 128 spills heavily.
  static constexpr int WAVES = (M == 13) ? 2 : 3;
};

// The carry of one register by one workgroup of T threads: block b = j T + t
// (j < NBT) to thread t.  Phase 1 of every block: its digits go to the
// destination as DIGIT_LIMBS raw planes (the destination holds at least
// three planes and is not the source) and its zero-in carry-out to LDS (cl:
// L words per block, limb-major).  Phase 2 reads them back and adds the
// predecessor's carry-out, one block in registers at a time.  For a
// canonical carry (m' of reduceMMGW) the finished digits go back to the
// destination, the sign of the number is that of its most significant
// nonzero digit (a workgroup maximum of (digit index + 1) * 2 + sign bit in
// `top`), and a third pass adjusts the end blocks as k_carry does and writes
// the residues.
template <int K, int QS, int NBT>
__device__ __forceinline__ void wg_carry(const CarryJob& J, const CarryArgs& A, int cbits, size_t n, uint32_t T,
                                         uint32_t* cl, int* top) {
  constexpr int L = K + 1;
  const uint32_t s = uint32_t(cbits) & 31u;
  const size_t nb = n / CARRY_L;
  const uint32_t t = threadIdx.x;
  CarryJob raw = J;
  raw.raw_out = 1;
  raw.dst2 = nullptr;
  auto load_digits = [&](size_t b, uint32_t (&dig)[CARRY_L][DIGIT_LIMBS]) {
#pragma unroll
    for (int l = 0; l < DIGIT_LIMBS; ++l) {
      const uint4 v = *reinterpret_cast<const uint4*>(J.dst + size_t(l) * n + b * CARRY_L);
      dig[0][l] = v.x;
      dig[1][l] = v.y;
      dig[2][l] = v.z;
      dig[3][l] = v.w;
    }
  };
  if (t == 0) *top = 0;
  for (int j = 0; j < NBT; ++j) {
    const size_t b = size_t(j) * T + t;
    uint32_t dig[CARRY_L][DIGIT_LIMBS], fo[L];
    carry_block_local<K, L, QS>(J, A, n, b, s, dig, fo);
    carry_store<K>(raw, A, n, b, dig, cbits);
#pragma unroll
    for (int i = 0; i < L; ++i) cl[size_t(i) * nb + b] = fo[i];
  }
  __syncthreads();
  int key = 0;
  for (int j = 0; j < NBT; ++j) {
    const size_t b = size_t(j) * T + t;
    const size_t pb = (b == 0) ? nb - 1 : b - 1;
    uint32_t dig[CARRY_L][DIGIT_LIMBS], fin[L];
    load_digits(b, dig);
#pragma unroll
    for (int i = 0; i < L; ++i) fin[i] = cl[size_t(i) * nb + pb];
    if (b == 0 && J.nega) big_neg<L>(fin);
    carry_block_finish<L, QS>(dig, fin, s);
    if (J.canon == 1) {
#pragma unroll
      for (int jj = CARRY_L - 1; jj >= 0; --jj) {
        const int sg = digit_sign(dig[jj]);
        if (sg != 0) {
          key = max(key, int(b * CARRY_L + jj + 1) * 2 + (sg < 0 ? 1 : 0));
          break;
        }
      }
      carry_store<K>(raw, A, n, b, dig, cbits);
    } else {
      carry_store<K>(J, A, n, b, dig, cbits);
    }
  }
  if (J.canon != 1) return;
  atomicMax(top, key);
  __syncthreads();
  const int kk = *top;
  const int sg = (kk == 0) ? 0 : ((kk & 1) ? -1 : 1);
  for (int j = 0; j < NBT; ++j) {
    const size_t b = size_t(j) * T + t;
    uint32_t dig[CARRY_L][DIGIT_LIMBS];
    load_digits(b, dig);
    if (sg < 0 && b == nb - 1) digit_add_base(dig[CARRY_L - 1], cbits);
    if (b == 0) digit_add(dig[0], ((sg < 0) ? -1 : 0) + 2);
    carry_store<K>(J, A, n, b, dig, cbits);
  }
}

// Grid: (slot-list length); block RunGeo<M>::T; dynamic LDS max(LINES lines
// of 2^M + 2^M / 32 words, (K + 1) n / 4 words).
template <int M, int K, int QS>
__global__ void __launch_bounds__(RunGeo<M>::T, RunGeo<M>::WAVES) k_sqr_run(const RunArgs R) {
  extern __shared__ uint32_t lds[];
  __shared__ int top;
  constexpr uint32_t S = 1u << M, Tl = RunGeo<M>::TL;
  constexpr int LINES = RunGeo<M>::LINES;
  const size_t n = S;
  const uint32_t slot = R.slot_ids[blockIdx.x];
  const size_t so = size_t(slot) * R.slot_stride;
  const int cbits = R.ca.cbits_slot ? int(R.ca.cbits_slot[slot]) : R.ca.cbits;
  const uint32_t line = threadIdx.x / Tl, tg = threadIdx.x % Tl;
  uint32_t* X = lds + line * (S + (S >> 5));
  const uint32_t* mult = R.mult + size_t(blockIdx.x) * size_t(R.steps);
  // Lines of a stage: `count` planes (job i / K, prime i % K), LINES at a
  // time; a surplus line repeats the last plane (same words to the same
  // addresses), so that every thread takes the row pass's barriers.
  auto stage = [&](auto mode, auto job_of, int count) {
    constexpr int MODE = decltype(mode)::value;
    for (int i0 = 0; i0 < count; i0 += LINES) {
      const int i = min(i0 + int(line), count - 1);
      const int k = i % K;
      row_line<M, MODE, false>(job_of(i / K), R.nt, k, size_t(k) * n, 1u, tg, X);
      __syncthreads();
    }
  };
  auto sq_job = [&](int j) {
    return j == 0 ? NttJob{R.r0 + so, R.wr + so, R.fwdR, R.invR, nullptr}
                  : NttJob{R.r0 + so, R.wq + so, R.fwdQ, R.invQ, nullptr};
  };
  auto m_job = [&](int) { return NttJob{R.tr + so, R.wr + so, R.fwdR, R.invR, R.nir + so}; };
  auto mn_job = [&](int) { return NttJob{R.tr + so, R.sq + so, R.fwdQ, R.invQ, R.nq + so}; };
  for (int step = 0; step < R.steps; ++step) {
    const uint32_t m = mult[step];
    if (m == 0u) break;
    const uint32_t a = m & ~SLOT_SNAP;
    stage(std::integral_constant<int, ROW_SQR>(), sq_job, 2 * K);
    CarryJob c{};
    c.src = R.wr + so;
    c.dst = R.tr + so;
    c.a = a;
    wg_carry<K, QS, RunGeo<M>::NBT>(c, R.ca, cbits, n, RunGeo<M>::T, lds, &top);
    __syncthreads();
    c.src = R.wq + so;
    c.dst = R.tq + so;
    c.nega = 1;
    c.raw_out = 1;
    wg_carry<K, QS, RunGeo<M>::NBT>(c, R.ca, cbits, n, RunGeo<M>::T, lds, &top);
    __syncthreads();
    stage(std::integral_constant<int, ROW_MULC>(), m_job, K);
    c = CarryJob{};
    c.src = R.wr + so;
    c.dst = R.tr + so;
    c.a = 1u;
    c.canon = 1;
    wg_carry<K, QS, RunGeo<M>::NBT>(c, R.ca, cbits, n, RunGeo<M>::T, lds, &top);
    __syncthreads();
    stage(std::integral_constant<int, ROW_MULC>(), mn_job, K);
    c = CarryJob{};
    c.src = R.sq + so;
    c.dst = const_cast<uint32_t*>(R.r0) + so;
    c.sub = R.tq + so;
    c.a = 1u;
    c.nega = 1;
    c.dst2 = (m & SLOT_SNAP) ? R.snap + so : nullptr;
    wg_carry<K, QS, RunGeo<M>::NBT>(c, R.ca, cbits, n, RunGeo<M>::T, lds, &top);
    __syncthreads();
  }
}

// Digits (DIGIT_LIMBS planes, two's complement, |d| < 2^95) -> K residue
// planes, CARRY_L digits per thread: the RNS encode of set_mpz
// (src/transformGPU.h:323-360) on the device.
template <int K>
__global__ void k_digits_to_res(const uint32_t* digits, uint32_t* dst, const CarryArgs A) {
  const size_t n = size_t(1) << A.LN;
  const size_t b = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= n / CARRY_L) return;
  const size_t i0 = b * CARRY_L;
  uint32_t dig[CARRY_L][DIGIT_LIMBS];
#pragma unroll
  for (int l = 0; l < DIGIT_LIMBS; ++l) {
    const uint4 v = *reinterpret_cast<const uint4*>(digits + size_t(l) * n + i0);
    dig[0][l] = v.x;
    dig[1][l] = v.y;
    dig[2][l] = v.z;
    dig[3][l] = v.w;
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    uint4 v;
    v.x = digit_residue_wide(dig[0], k, A);
    v.y = digit_residue_wide(dig[1], k, A);
    v.z = digit_residue_wide(dig[2], k, A);
    v.w = digit_residue_wide(dig[3], k, A);
    *reinterpret_cast<uint4*>(dst + size_t(k) * n + i0) = v;
  }
}

// Register copy for each slot of a list: words [0, len) at dst_off <- src_off
// (offsets in words from the slot base).  Grid: (blocks, slots).
__global__ void k_copy_slots(uint32_t* base, size_t dst_off, size_t src_off, size_t len, const uint32_t* slot_ids,
                             size_t slot_stride) {
  const size_t so = slot_offset(slot_ids, slot_stride, blockIdx.y);
  const uint4* src = reinterpret_cast<const uint4*>(base + so + src_off);
  uint4* dst = reinterpret_cast<uint4*>(base + so + dst_off);
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < len / 4; i += size_t(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}

// Explicit instantiations, looked up by the host engine: row lines of 2^4 ..
// 2^13 elements (single-pass lengths 2^4..2^13, second pass of 2^14..2^23) and
// column lines of 2^7 .. 2^11 elements (first pass of 2^14..2^23).
#define ANNIE_ROW(M)                                          \
  template __global__ void k_row<M, ROW_FWD>(const NttArgs);  \
  template __global__ void k_row<M, ROW_SQR>(const NttArgs);  \
  template __global__ void k_row<M, ROW_MULC>(const NttArgs); \
  template __global__ void k_row<M, ROW_MULV>(const NttArgs);
ANNIE_ROW(4)
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
#define ANNIE_CARRY(K)                                 \
  template __global__ void k_carry<K, 0>(const CarryArgs); \
  template __global__ void k_carry<K, 1>(const CarryArgs); \
  template __global__ void k_carry<K, 2>(const CarryArgs);
ANNIE_CARRY(2)
ANNIE_CARRY(3)
ANNIE_CARRY(4)
ANNIE_CARRY(5)
// Runs: the classes of lengths 2^9 .. 2^13 (K = 3, 4 with digits of 33..63
// bits; K = 5 with digits of 64..95 bits).
#define ANNIE_RUN(M)                                          \
  template __global__ void k_sqr_run<M, 3, 1>(const RunArgs); \
  template __global__ void k_sqr_run<M, 4, 1>(const RunArgs); \
  template __global__ void k_sqr_run<M, 5, 2>(const RunArgs);
ANNIE_RUN(9)
ANNIE_RUN(10)
ANNIE_RUN(11)
ANNIE_RUN(12)
ANNIE_RUN(13)
template __global__ void k_digits_to_res<2>(const uint32_t*, uint32_t*, const CarryArgs);
template __global__ void k_digits_to_res<3>(const uint32_t*, uint32_t*, const CarryArgs);
template __global__ void k_digits_to_res<4>(const uint32_t*, uint32_t*, const CarryArgs);
template __global__ void k_digits_to_res<5>(const uint32_t*, uint32_t*, const CarryArgs);

}  // namespace annie
