/*
 * annie_hip.h -- C ABI of the MI355X (gfx950) modular-squaring engine.
 *
 * One context holds one modulus N and the register machine of the reference's
 * `transform` interface (src/transform.h:18-45): three user registers
 * Reg0..Reg2 holding values in Montgomery-McLaughlin-Gallot-Woltman form
 * x* = 2^-1 R x mod N (doc/annie.tex:138-164).  Each entry point below names
 * the reference interface it replaces.  Plain pointers and sizes only; no
 * exception crosses the ABI: every call returns a status and
 * annie_hip_last_error() describes the last failure of the calling thread.
 * A context is used by one host thread at a time.  Big integers are passed as
 * little-endian arrays of 64-bit limbs.
 */
#ifndef ANNIE_HIP_H
#define ANNIE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct annie_hip annie_hip_t;

enum {
  ANNIE_OK = 0,
  ANNIE_INIT_FAILED = 1,   /* gcd(N, R) > 1: the factor is available (reference: init_failed()) */
  ANNIE_ERR = -1,          /* HIP runtime failure */
  ANNIE_ERR_SIZE = -2,     /* transform size outside the supported range */
  ANNIE_ERR_ARG = -3,      /* invalid argument (register index, buffer too small, odd/even N ...) */
  ANNIE_ERR_NOMEM = -4     /* device allocation failed */
};

/* Register indices (src/transform.h:21). */
enum { ANNIE_REG0 = 0, ANNIE_REG1 = 1, ANNIE_REG2 = 2 };

/* Transform size chosen for N and base a, exactly as create_transformGPU
 * (src/transformGPU.h:822-850).  No GPU needed. */
int annie_hip_select_plan(const uint64_t* N, size_t nlimbs, int32_t a, int* K, int* pck, int* LN, int* lbase);

/* Host-side tables (no GPU needed), for tests.  Constants of a plan: out[0..4]
 * primes, [5..9] p^-1 mod 2^32, [10..14] n^-1 2^64 mod p, [15..19] 2^64 mod p,
 * [20..24] 2^96 mod p, [25 + 5i + j] p_i^-1 mod p_j in Montgomery form (the
 * reference's INVPi_Pj, src/transformGPU.h:699-705).  Twiddles: table 0..3 =
 * cyclic forward, cyclic inverse, negacyclic forward, negacyclic inverse
 * (Montgomery form w and w * p^-1 mod 2^32), heap-indexed, n entries. */
int annie_hip_plan_constants(int LN, int K, int pck, int lbase, uint32_t* out, size_t cap);
int annie_hip_twiddles(int LN, int k, int table, uint32_t* w, uint32_t* wq, size_t cap);

/* Replaces create_transformGPU(N, a, device) (src/transformGPU.h:822-885) and the
 * transformGPU constructor (:644-753).  min_ln is the smallest accepted log2
 * transform length (the reference accepts 10..23, or 4..23 in its QUICK_TEST
 * build, src/transformGPU.h:852-857).  Returns ANNIE_INIT_FAILED with a
 * context whose factor can be read when gcd(N, R) > 1. */
int annie_hip_create(const uint64_t* N, size_t nlimbs, int32_t a, int device, int min_ln, annie_hip_t** out);

/* transform::get_factor (src/transform.h:31; src/transformGPU.h:604). */
int annie_hip_get_factor(const annie_hip_t* ctx, uint64_t* limbs, size_t cap, size_t* nout);

/* transform::set1 -- r0 = 1 (src/transform.h:25; src/transformGPU.h:571-574). */
int annie_hip_set1(annie_hip_t* ctx);

/* transform::copy -- r_dst = r_src (src/transform.h:27; src/transformGPU.h:588-591). */
int annie_hip_copy(annie_hip_t* ctx, size_t dst, size_t src);

/* transform::square_mul -- r0 = a * r0^2 (src/transform.h:28; src/transformGPU.h:593-596). */
int annie_hip_square_mul(annie_hip_t* ctx, int32_t a);

/* No counterpart in the reference: a squaring that allocates nothing, for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops).  y = a * x^2 in MMGW form, the squaring of
 * annie_hip_square_mul on caller-owned device buffers: x and y hold K planes of
 * n u32 residues (plane-major, the layout of annie_hip_read_register), 16-byte
 * aligned; y may equal x.  Its work registers live in a static device array
 * of the library (one per GPU, 4 K n + 3 n words at the largest size, held
 * until the process ends) instead of the context's own, so the context's
 * registers are untouched.  Limit: K * n <= ANNIE_HIP_NOALLOC_MAX_KN words per
 * register (ANNIE_ERR_SIZE above it).  Enqueued on the context's stream and
 * returns without waiting.  Every call on one GPU shares that array: calls
 * whose work overlaps on one GPU (other threads, other contexts' streams)
 * race; calls on one stream are ordered. */
#define ANNIE_HIP_NOALLOC_MAX_KN ((size_t)1 << 19)
int annie_hip_square_mul_noalloc(annie_hip_t* ctx, const uint32_t* x, uint32_t* y, int32_t a);

/* The reference's hot loop body for `count` exponent bits (src/annie.h:319-341 without
 * the Gerbicz-Li products): for i < count, r0 = (bits[i] ? a : 1) * r0^2. */
int annie_hip_square_mul_bits(annie_hip_t* ctx, const uint8_t* bits, size_t count, int32_t a);

/* As annie_hip_square_mul_bits, timed with HIP events on the context's stream:
 * *ms = GPU time of the whole sequence (no counterpart in the reference; its
 * bench mode times square_mul with a wall clock, src/annie.h:515-547). */
int annie_hip_time_square_mul_bits(annie_hip_t* ctx, const uint8_t* bits, size_t count, int32_t a, float* ms);

/* Per-kernel timing with HIP events (the reference's dormant OpenCL profiling
 * queue, src/ocl.h:330-374, 579-597): begin enables it, end synchronises and
 * writes one text line per kernel kind: "name count total_ms". */
int annie_hip_profile_begin(annie_hip_t* ctx);
int annie_hip_profile_end(annie_hip_t* ctx, char* buf, size_t cap);

/* transform::mul -- r_dst *= r_src (src/transform.h:29; src/transformGPU.h:598-601). */
int annie_hip_mul(annie_hip_t* ctx, size_t dst, size_t src);

/* transform::get -- canonical value of r_src in [0, N) (src/transform.h:26;
 * src/transformGPU.h:576-586).  Writes up to cap limbs, *nout = limbs used. */
int annie_hip_get(annie_hip_t* ctx, size_t src, uint64_t* limbs, size_t cap, size_t* nout);

/* transform::save_checkpoint / read_checkpoint payload for r0 and r1
 * (src/transformGPU.h:606-641): u32 K, u32 n, then for each prime k the plane
 * of r0 followed by the plane of r1 (n u32 words each). */
size_t annie_hip_checkpoint_size(const annie_hip_t* ctx);
int annie_hip_save_checkpoint(annie_hip_t* ctx, void* buf, size_t cap);
int annie_hip_load_checkpoint(annie_hip_t* ctx, const void* buf, size_t size);

/* get_info_short fields (src/transformGPU.h:784-789). */
int annie_hip_info(const annie_hip_t* ctx, int* K, int* pck, int* LN, int* lbase);

/* Launch schedule of the NTT passes (no counterpart in the reference, whose
 * OpenCL kernels have fixed work-group shapes, src/engine.h:262-339): the
 * number of lines per row-pass workgroup for rows of fewer than 64 threads
 * (rows of up to 2^9 elements at 16 per thread) and the number of columns per
 * column-pass workgroup for columns of up to 2^9 elements; 0 keeps the
 * default.  Results do not depend on the schedule.  Values must be powers of
 * two that fit the plan (at most the rows / columns there are, 512 row-pass
 * or 1024 column-pass threads, 64 KiB of LDS per workgroup):
 * annie_hip_schedule_fits checks a pair for a transform length without a
 * context; set_schedule returns ANNIE_ERR_ARG for a pair that does not fit
 * and drops the context's captured graphs.  get_schedule reports the values
 * in effect (the defaults where 0 was set; 0 where a pass has no knob). */
int annie_hip_schedule_fits(int LN, int row_lines, int col_cols);
int annie_hip_set_schedule(annie_hip_t* ctx, int row_lines, int col_cols);
int annie_hip_get_schedule(const annie_hip_t* ctx, int* row_lines, int* col_cols);

/* Block until all queued work of the context is done. */
int annie_hip_sync(annie_hip_t* ctx);

/* Diagnostics for kernel-level parity tests (no counterpart in the reference):
 * copy register `reg` (K planes of n residues, plane-major) to host memory. */
int annie_hip_read_register(annie_hip_t* ctx, size_t reg, uint32_t* out, size_t cap_words);
int annie_hip_write_register(annie_hip_t* ctx, size_t reg, const uint32_t* in, size_t words);

/* Diagnostics (no counterpart in the reference): run one squaring of r0 up to
 * an intermediate stage (1: convolutions in registers 5/6, 2: T_R in register
 * 4 and T_Q digits, 3: m' convolution in register 5, 4: m' + 2 in register 4),
 * without updating r0; and read the T_Q digit planes. */
int annie_hip_debug_square_stage(annie_hip_t* ctx, int32_t a, int stage);
int annie_hip_read_tq(annie_hip_t* ctx, uint32_t* out, size_t cap_words);

/* ---- Batches (no counterpart in the reference, which runs one test per
 * process).  A batch holds `slots` moduli of one length class -- one plan
 * (K, pck, LN, lbase) of the size selection, src/transformGPU.h:798-857 --
 * with shared twiddles and per-slot registers, and runs each operation for a
 * list of slots with one set of launches: the register operations of
 * src/transform.h:18-45 (set1, copy, square_mul, mul, get), per slot.  Slot
 * lists hold distinct slot ids that have a modulus loaded.  A batch is
 * released with annie_hip_destroy. */
int annie_hip_batch_create(int K, int pck, int LN, int lbase, size_t slots, int device, annie_hip_t** out);
/* Device memory of a batch of plan (K, LN): bytes per slot (its registers,
 * digit buffers and slot-list staging) and bytes held once per batch (twiddle
 * tables), for admission under an HBM budget. */
int annie_hip_batch_bytes(int K, int LN, size_t* per_slot, size_t* fixed);
/* Loads modulus N (base a) into a slot: its plan must have the batch's K, pck
 * and LN, and lbase with the same pck * lbase >> 5 (the batch's lbase or
 * another of that digit-width class: each slot keeps its own).  Returns
 * ANNIE_INIT_FAILED with gcd(N, R) in `factor` when N is not invertible mod R
 * (src/transformGPU.h:652-661). */
int annie_hip_batch_load(annie_hip_t* b, size_t slot, const uint64_t* N, size_t nlimbs, int32_t a, uint64_t* factor,
                         size_t cap, size_t* nout);
int annie_hip_batch_set1(annie_hip_t* b, const uint32_t* slots, size_t count);
int annie_hip_batch_copy(annie_hip_t* b, size_t dst, size_t src, const uint32_t* slots, size_t count);
/* r0 = a[i] * r0^2 in slot slots[i], a[i] & 0x7fffffff in [1, 2^31).  When bit
 * 31 of a[i] is set the new r0 is also written to the slot's snapshot
 * register (register 3 of a batch), so that the Gerbicz-Li product
 * r1 *= r0 of that step (src/annie.h:338) can be taken later, together with
 * other slots'. */
int annie_hip_batch_square_mul(annie_hip_t* b, const uint32_t* slots, const uint32_t* a, size_t count);
/* r_dst *= r_src; src may be 3, the snapshot register. */
int annie_hip_batch_mul(annie_hip_t* b, size_t dst, size_t src, const uint32_t* slots, size_t count);
/* A run of squarings in one launch (src/annie.h:319-341 over many loop
 * iterations): for each listed slot i, r0 = m * r0^2 for m = mult[i * steps + t],
 * t = 0, 1, ..., up to the slot's first m = 0 (zeros only as a suffix); bit 31
 * of m as in annie_hip_batch_square_mul, allowed on the last squaring of the
 * slot's run only.  One workgroup per slot runs the whole squaring sequence;
 * the registers end equal, word for word, to those of the same squarings
 * issued one by one.  steps <= ANNIE_HIP_RUN_MAX_STEPS.  ANNIE_ERR_SIZE when
 * the batch's plan has no run kernel (annie_hip_batch_run_supported == 0:
 * transform lengths outside 2^9 .. 2^13, other RNS sizes or digit widths). */
#define ANNIE_HIP_RUN_MAX_STEPS 256
int annie_hip_batch_square_run(annie_hip_t* b, const uint32_t* slots, size_t count, const uint32_t* mult, size_t steps);
int annie_hip_batch_run_supported(const annie_hip_t* b);
/* Canonical values of register `reg` of each listed slot: slot i's limbs at
 * limbs + i * stride (stride limbs each), nout[i] limbs used. */
int annie_hip_batch_get(annie_hip_t* b, size_t reg, const uint32_t* slots, size_t count, uint64_t* limbs, size_t stride,
                        size_t* nout);
int annie_hip_batch_sync(annie_hip_t* b);
/* on != 0: batched squarings replay captured HIP graphs (one per slot-list
 * length, captured on first use) instead of issuing their launches one by
 * one; results are identical either way. */
int annie_hip_batch_set_graphs(annie_hip_t* b, int on);

/* Releases the context (the reference deletes the transform, src/annie.h:406). */
void annie_hip_destroy(annie_hip_t* ctx);

const char* annie_hip_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* ANNIE_HIP_H */
