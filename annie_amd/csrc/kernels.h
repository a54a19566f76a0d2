// Kernel argument structures shared by the gfx950 kernels and the host engine.
#pragma once
#include <cstddef>
#include <cstdint>

#include "arith.h"

namespace annie {

// A twiddle in HBM: the root in Montgomery form (w * p^-1 mod 2^32 is formed
// on the fly with one multiply, which halves the table traffic).
typedef uint32_t TwDev;

// Elements per thread in the row passes and in columns of up to 2^9 (four
// levels per register round; columns of 2^10 and 2^11 take 32 and five).
constexpr int NTT_R = 16;
constexpr int NTT_LOGR = 4;
// Consecutive coefficients per thread in the carry kernel.
constexpr int CARRY_L = 4;
constexpr int CARRY_WG = 256;
// 32-bit limbs of a packed digit (coefficients carry at most 72 bits plus sign).
constexpr int DIGIT_LIMBS = 3;

// One transform job: K residue planes (plane stride n words).
struct NttJob {
  const uint32_t* src;  // input planes
  uint32_t* dst;        // output planes (may equal src)
  const TwDev* fwd;     // forward twiddles of prime 0 (prime k at + k * n)
  const TwDev* inv;     // inverse twiddles of prime 0
  const uint32_t* opnd; // pointwise operand planes (MULC / MULV), else null
};

// A launch runs its jobs for every slot of a slot list: grid index z (or y
// for the carry) = position * njobs + job.  A slot is one test's register
// file inside a batch allocation; the job pointers address slot 0 and slot s
// is at + s * slot_stride words.  slot_ids == null means a single slot at
// offset 0 (the one-test engine).
struct NttArgs {
  NttJob job[2];
  const uint32_t* slot_ids;
  size_t slot_stride;
  int njobs;
  uint32_t p[KMAX], q[KMAX];
  uint32_t sc[KMAX], scq[KMAX];  // n^-1 * 2^64 (Montgomery scale of a data product)
  int LN;        // log2 of the transform length
  int m;         // log2 of the line length handled by this pass
  int m1;        // levels done by the column pass (0 for a single-pass transform)
  int cols;      // columns per workgroup (column passes) / lines per workgroup (row pass)
};

enum RowMode : int { ROW_FWD = 0, ROW_SQR = 1, ROW_MULC = 2, ROW_MULV = 3 };

struct CarryJob {
  const uint32_t* src;  // K residue planes of the convolution (exact values mod p_k)
  uint32_t* dst;        // K residue planes, or DIGIT_LIMBS raw digit planes if raw_out
  const uint32_t* sub;  // DIGIT_LIMBS raw digit planes subtracted from each coefficient, or null
  uint32_t a;           // multiplier of every coefficient
  int nega;             // 1: value mod B^n + 1 (wrap negated), 0: mod B^n - 1
  int canon;            // 1: reduce m' into [0, R) then add 2; 0 otherwise (mod Q needs no fix-up)
  int raw_out;          // 1: write raw digits instead of residues
  uint32_t* dst2;       // second destination (residues) for flagged slots, or null
};

// Bit of a per-slot multiplier that also asks for the second destination.
constexpr uint32_t SLOT_SNAP = 0x80000000u;

// Batched squarings over at most this many list positions pass their
// per-position multipliers and snapshot flags as bits in the carry's kernel
// arguments instead of a device array uploaded for every squaring.
constexpr int INLINE_SLOTS = 512;

struct CarryArgs {
  CarryJob job[2];
  const uint32_t* slot_ids;
  size_t slot_stride;
  int njobs;
  const uint32_t* a_slot;  // per-position multiplier (replaces job.a) or null
  const uint32_t* snap;    // per-position flags (SLOT_SNAP): also write job.dst2, or null
  int LN;
  int cbits;  // bits per coefficient (pck * lbase)
  // Batches whose slots differ in lbase: per-slot bits per coefficient
  // (indexed by slot id), replacing cbits; all with the same cbits >> 5.
  const uint32_t* cbits_slot;
  // Per-position flags as bits (bit i for list position i), replacing
  // a_slot / snap: bits_mode 1: multiply by slot_base[slot id] where the bit
  // is set, by 1 elsewhere; 2: the bit is the position's SLOT_SNAP; 0: unused.
  int bits_mode;
  const uint32_t* slot_base;
  uint32_t bits[INLINE_SLOTS / 32];
};

// A run of batched squarings (one workgroup per listed slot): for slot list
// position i and step t < steps, r0 = m * r0^2 with m = mult[i * steps + t];
// m = 0 ends the slot's run, bit SLOT_SNAP also writes the new r0 to the
// snapshot register.  Register pointers address slot 0 (slot s at
// + s * slot_stride words).
struct RunArgs {
  NttArgs nt;            // transform constants (p, q, sc, scq, LN)
  CarryArgs ca;          // carry parameters (LN, cbits); the prime constants are literals (PLIT)
  const uint32_t* slot_ids;
  const uint32_t* mult;
  size_t slot_stride;
  int steps;
  const uint32_t* r0;    // x, and the result of each squaring
  uint32_t *wr, *wq;     // cyclic / negacyclic convolutions
  uint32_t *tr, *sq;     // T_R then m' + 2; the (m' + 2) N convolution
  uint32_t* tq;          // T_Q digits (DIGIT_LIMBS planes)
  uint32_t* snap;        // snapshot register
  const uint32_t *nir, *nq;   // transformed N^-1 mod R and N (scaled)
  const TwDev *fwdR, *invR, *fwdQ, *invQ;
};

}  // namespace annie
